"""The throughput path of batched signing, against the reference's own
outputs: 320 independent XMSS-SHA2_10_256 keys in one job take the kernels a
4096-key batch takes -- k_chains over many runs, k_level over more than 64
subtrees, k_wots for more than 256 requests -- where smaller batches take the
latency kernels.  Keys come from seeds through xmss_batch_keypair; each key
signs one log-uniform 32 B - 4 KB message at a random index, with the last
index, an exhausted key and 1022 among them (tests/golden/plain_batch_large.json,
written by the reference's xmss_core.c)."""
import ctypes as C

import pytest

import workloads as W
from common import dg, golden, product
from ffi import buf, raw

pytestmark = pytest.mark.gpu


def _arr(ptrs):
    a = (C.c_void_p * len(ptrs))()
    for i, p in enumerate(ptrs):
        a[i] = C.cast(p, C.c_void_p)
    return a


def _stats(L):
    L.xmss_mi355x_kernel_stats.restype = C.c_size_t
    L.xmss_mi355x_kernel_stats.argtypes = [C.c_char_p, C.c_size_t]
    b = C.create_string_buffer(1 << 16)
    L.xmss_mi355x_kernel_stats(b, len(b))
    return {ln.split()[0]: int(ln.split()[1]) for ln in b.value.decode().splitlines() if ln.strip()}


def _inputs(g, p):
    count = g["count"]
    seeds = [W.key_seed(g["key_seed"], k, p.n) for k in range(count)]
    msgs = [W.message(g["msg_seed"], k) for k in range(count)]
    return seeds, msgs


def test_batch_keypair_320_keys_equals_reference():
    g = golden("plain_batch_large.json")
    lib = product(0)
    L = lib.lib
    p = lib.params(g["oid"], 0)
    seeds, _ = _inputs(g, p)
    count = g["count"]
    L.xmss_batch_keypair.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_char_p, C.c_uint32, C.c_size_t]
    pks = [buf(4 + p.pk_bytes) for _ in range(count)]
    sks = [buf(4 + p.sk_bytes) for _ in range(count)]
    assert L.xmss_batch_keypair(0, _arr(pks), _arr(sks), b"".join(seeds), g["oid"], count) == 0
    for k, it in enumerate(g["items"]):
        pk = raw(pks[k], 4 + p.pk_bytes)
        assert pk[4:4 + p.n].hex() == it["root"], "key %d root" % k
        assert raw(sks[k], 4 + p.sk_bytes) == W.plain_sk(g["oid"], p, 0, seeds[k], pk[4:4 + p.n])


def test_batch_sign_320_keys_equals_reference_on_the_throughput_kernels():
    g = golden("plain_batch_large.json")
    lib = product(0)
    L = lib.lib
    oid = g["oid"]
    p = lib.params(oid, 0)
    seeds, msgs = _inputs(g, p)
    items = g["items"]
    count = len(items)
    roots = [bytes.fromhex(it["root"]) for it in items]
    sk_in = [W.plain_sk(oid, p, it["idx"], seeds[k], roots[k]) for k, it in enumerate(items)]
    sks = [buf(s) for s in sk_in]
    mbufs = [buf(m) for m in msgs]
    mlens = (C.c_ulonglong * count)(*[len(m) for m in msgs])
    L.xmss_mi355x_profile.argtypes = [C.c_int]
    L.xmss_mi355x_sign_prepare.restype = C.c_void_p
    L.xmss_mi355x_sign_prepare.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    L.xmss_mi355x_job_run.argtypes = [C.c_void_p]
    L.xmss_mi355x_sign_collect.argtypes = [C.c_void_p] * 4
    L.xmss_mi355x_job_free.argtypes = [C.c_void_p]
    L.xmss_mi355x_reset_stats()
    L.xmss_mi355x_profile(1)
    try:
        job = L.xmss_mi355x_sign_prepare(0, _arr(sks), _arr(mbufs), mlens, count)
        assert L.xmss_mi355x_job_run(job) == 0
        sms = [buf(p.sig_bytes + len(m)) for m in msgs]
        smlens = (C.c_ulonglong * count)()
        rets = (C.c_int * count)()
        assert L.xmss_mi355x_sign_collect(job, _arr(sms), smlens, rets) == 0
        L.xmss_mi355x_job_free(job)
        st = _stats(L)
    finally:
        L.xmss_mi355x_profile(0)
    # the kernels of the 4096-key batch ran, not the latency kernels
    for kname in ("chains", "ltree", "level", "wots"):
        assert st.get(kname, 0) > 0, "kernel %s not launched: %s" % (kname, sorted(st))
    signed = []
    for k, it in enumerate(items):
        assert rets[k] == it["rc"], "key %d rc" % k
        assert dg(raw(sks[k], len(sk_in[k]))) == it["sk_out"], "key %d sk" % k
        if it["rc"] == 0:
            sm = raw(sms[k], smlens[k])
            assert dg(sm) == it["sm"], "key %d (idx %d) signature" % (k, it["idx"])
            signed.append((k, sm))
    # the batch verifier: every reference signature verifies as the reference
    # says (the last index's does not), and a flipped byte is rejected
    vsm, vpk, want = [], [], []
    for j, (k, sm) in enumerate(signed):
        pk = W.with_oid(oid, roots[k] + seeds[k][2 * p.n:])
        if j % 7 == 3:
            b = bytearray(sm)
            b[4 + (j * 131) % (p.sig_bytes - 4)] ^= 0x10
            vsm.append(bytes(b))
            want.append(-1)
        else:
            vsm.append(sm)
            want.append(items[k]["verify"])
        vpk.append(pk)
    kv = len(vsm)
    L.xmss_batch_sign_open.argtypes = [C.c_int] + [C.c_void_p] * 6 + [C.c_size_t]
    smb = [buf(s) for s in vsm]
    pkb = [buf(x) for x in vpk]
    outs = [buf(len(s)) for s in vsm]
    mls = (C.c_ulonglong * kv)()
    vr = (C.c_int * kv)()
    assert L.xmss_batch_sign_open(0, _arr(outs), mls, _arr(smb), (C.c_ulonglong * kv)(*map(len, vsm)),
                                  _arr(pkb), vr, kv) == 0
    assert list(vr) == want
    for j, (k, _) in enumerate(signed):
        if want[j] == 0:
            assert raw(outs[j], mls[j]) == msgs[k]


def test_batch_sign_320_keys_as_one_graph_equals_reference():
    """The same batch without per-kernel timing: the job's launches are sized
    in a pass without launches, captured into one HIP graph and launched as
    that; a second run replays the graph and gives the same bytes."""
    g = golden("plain_batch_large.json")
    lib = product(0)
    L = lib.lib
    oid = g["oid"]
    p = lib.params(oid, 0)
    seeds, msgs = _inputs(g, p)
    items = g["items"]
    count = len(items)
    roots = [bytes.fromhex(it["root"]) for it in items]
    sk_in = [W.plain_sk(oid, p, it["idx"], seeds[k], roots[k]) for k, it in enumerate(items)]
    sks = [buf(s) for s in sk_in]
    mbufs = [buf(m) for m in msgs]
    mlens = (C.c_ulonglong * count)(*[len(m) for m in msgs])
    L.xmss_mi355x_sign_prepare.restype = C.c_void_p
    L.xmss_mi355x_sign_prepare.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    L.xmss_mi355x_job_run.argtypes = [C.c_void_p]
    L.xmss_mi355x_sign_collect.argtypes = [C.c_void_p] * 4
    L.xmss_mi355x_job_free.argtypes = [C.c_void_p]
    job = L.xmss_mi355x_sign_prepare(0, _arr(sks), _arr(mbufs), mlens, count)
    try:
        outs = []
        for run in range(2):
            assert L.xmss_mi355x_job_run(job) == 0
            sms = [buf(p.sig_bytes + len(m)) for m in msgs]
            smlens = (C.c_ulonglong * count)()
            rets = (C.c_int * count)()
            assert L.xmss_mi355x_sign_collect(job, _arr(sms), smlens, rets) == 0
            outs.append([raw(sms[k], smlens[k]) for k in range(count)])
            for k, it in enumerate(items):
                assert rets[k] == it["rc"], "run %d key %d rc" % (run, k)
                if it["rc"] == 0:
                    assert dg(outs[-1][k]) == it["sm"], "run %d key %d (idx %d) signature" % (run, k, it["idx"])
        assert outs[0] == outs[1]
        for k, it in enumerate(items):
            assert dg(raw(sks[k], len(sk_in[k]))) == it["sk_out"], "key %d sk" % k
    finally:
        L.xmss_mi355x_job_free(job)
