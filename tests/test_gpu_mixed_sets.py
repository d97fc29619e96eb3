"""The mixed-parameter workload against the reference's own outputs: the
three sets XMSSMT-SHA2_20/2_256 (SHA-256), XMSS-SHA2_10_512 (SHA-512, 131
chains) and XMSS-SHAKE_10_256 (SHAKE128), 260 plain-core signatures each, in
ONE mixed job (more than 256 requests per set: the throughput kernels), then
one mixed verification job over all of them, run directly, then as one HIP
graph over the three sets, replayed and refilled (tests/golden/mixed_sets.json,
written by the reference's xmss_core.c)."""
import ctypes as C
import random

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


def _cases(lib, g):
    cases = []  # (mt, oid, p, sk_in, msg, pk, row)
    for name, s in g.items():
        oid, mt = s["oid"], s["mt"]
        p = lib.params(oid, mt)
        seeds = [W.key_seed(s["key_seed"], k, p.n) for k in range(s["keys"])]
        roots = [bytes.fromhex(r) for r in s["roots"]]
        for i, row in enumerate(s["items"]):
            k = i % s["keys"]
            sk = W.plain_sk(oid, p, row["idx"], seeds[k], roots[k])
            pk = W.with_oid(oid, roots[k] + seeds[k][2 * p.n:])
            cases.append((mt, oid, p, sk, W.message(s["msg_seed"], i), pk, row))
    random.Random(4).shuffle(cases)
    return cases


def test_mixed_job_of_the_three_sets_equals_reference():
    g = golden("mixed_sets.json")
    lib = product(0)
    L = lib.lib
    L.xmss_mi355x_mixed_sign_prepare.restype = C.c_void_p
    L.xmss_mi355x_mixed_sign_prepare.argtypes = [C.c_void_p] * 4 + [C.c_size_t]
    L.xmss_mi355x_mixed_verify_prepare.restype = C.c_void_p
    L.xmss_mi355x_mixed_verify_prepare.argtypes = [C.c_void_p] * 4 + [C.c_size_t]
    L.xmss_mi355x_job_run.argtypes = [C.c_void_p]
    L.xmss_mi355x_sign_collect.argtypes = [C.c_void_p] * 4
    L.xmss_mi355x_verify_collect.argtypes = [C.c_void_p] * 4
    L.xmss_mi355x_job_free.argtypes = [C.c_void_p]
    cases = _cases(lib, g)
    k = len(cases)
    mts = (C.c_int * k)(*[c[0] for c in cases])
    sks = [buf(c[3]) for c in cases]
    mb = [buf(c[4]) for c in cases]
    mlens = (C.c_ulonglong * k)(*[len(c[4]) for c in cases])
    job = L.xmss_mi355x_mixed_sign_prepare(mts, _arr(sks), _arr(mb), mlens, k)
    assert L.xmss_mi355x_job_run(job) == 0
    sms = [buf(c[2].sig_bytes + len(c[4])) for c in cases]
    smlens = (C.c_ulonglong * k)()
    rets = (C.c_int * k)()
    assert L.xmss_mi355x_sign_collect(job, _arr(sms), smlens, rets) == 0
    L.xmss_mi355x_job_free(job)
    out = []
    for i, c in enumerate(cases):
        row = c[6]
        assert rets[i] == row["rc"] == 0, "item %d rc" % i
        sm = raw(sms[i], smlens[i])
        assert dg(sm) == row["sm"], "item %d (oid %d, idx %d) signature" % (i, c[1], row["idx"])
        assert dg(raw(sks[i], len(c[3]))) == row["sk_out"], "item %d sk" % i
        out.append(sm)
    # one mixed verification job: all valid, plus one flipped byte per set
    vsm = list(out)
    flipped = set()
    for name_oid in {(c[0], c[1]) for c in cases}:
        t = next(i for i, c in enumerate(cases) if (c[0], c[1]) == name_oid)
        b = bytearray(vsm[t])
        b[cases[t][2].index_bytes + 3] ^= 1
        vsm[t] = bytes(b)
        flipped.add(t)
    smb = [buf(s) for s in vsm]
    pkb = [buf(c[5]) for c in cases]
    outs = [buf(len(s)) for s in vsm]
    smls = (C.c_ulonglong * k)(*[len(s) for s in vsm])
    mls = (C.c_ulonglong * k)()
    vr = (C.c_int * k)()
    job = L.xmss_mi355x_mixed_verify_prepare(mts, _arr(smb), smls, _arr(pkb), k)
    L.xmss_mi355x_verify_refill.argtypes = [C.c_void_p] * 4
    # runs: direct launches, the sets captured into one HIP graph, two replays;
    # then a refill with the flips moved to other requests and two replays
    flips = [flipped]
    try:
        for run in range(6):
            if run == 4:
                vsm = list(out)
                moved = set()
                for t in flipped:
                    u = (t + 1) % k
                    b = bytearray(vsm[u])
                    b[cases[u][2].index_bytes + cases[u][2].n + 7] ^= 4
                    vsm[u] = bytes(b)
                    moved.add(u)
                smb = [buf(s) for s in vsm]
                smls = (C.c_ulonglong * k)(*[len(s) for s in vsm])
                outs = [buf(len(s)) for s in vsm]
                assert L.xmss_mi355x_verify_refill(job, _arr(smb), smls, _arr(pkb)) == 0
                flips.append(moved)
            assert L.xmss_mi355x_job_run(job) == 0
            assert L.xmss_mi355x_verify_collect(job, _arr(outs), mls, vr) == 0
            bad = flips[-1]
            for i, c in enumerate(cases):
                if i in bad:
                    assert vr[i] == -1 and mls[i] == 0, "run %d item %d accepted" % (run, i)
                else:
                    assert vr[i] == c[6]["verify"] == 0, "run %d verify item %d" % (run, i)
                    assert raw(outs[i], mls[i]) == c[4]
    finally:
        L.xmss_mi355x_job_free(job)
