"""Batched plain-core signing and verification on the GPU (independent keys,
random leaves, 32 B - 4 KB messages) against the reference's outputs."""
import ctypes as C

import pytest

from common import golden, oracle, product
from ffi import buf, raw

pytestmark = pytest.mark.gpu


def _arr(ptrs):
    a = (C.c_void_p * len(ptrs))()
    for i, p in enumerate(ptrs):
        a[i] = C.cast(p, C.c_void_p)
    return a


def test_batch_sign_and_verify_match_reference():
    g = golden("plain_batch.json")
    lib = product(0)
    p = lib.params(g["oid"], 0)
    items = g["items"]
    k = len(items)
    sks = [buf(bytes.fromhex(it["sk_in"])) for it in items]
    msgs = [bytes.fromhex(it["msg"]) for it in items]
    mbufs = [buf(m) for m in msgs]
    sms = [buf(p.sig_bytes + len(m)) for m in msgs]
    smlens = (C.c_ulonglong * k)()
    mlens = (C.c_ulonglong * k)(*[len(m) for m in msgs])
    rets = (C.c_int * k)()
    L = lib.lib
    L.xmss_batch_sign.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_size_t]
    L.xmss_batch_sign_open.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_size_t]
    assert L.xmss_batch_sign(0, _arr(sks), _arr(sms), smlens, _arr(mbufs), mlens, rets, k) == 0
    sm_out = []
    for i, it in enumerate(items):
        assert rets[i] == it["rc"]
        sm = raw(sms[i], smlens[i])
        assert sm.hex() == it["sm"], "item %d" % i
        assert raw(sks[i], 4 + p.sk_bytes).hex() == it["sk_out"]
        sm_out.append(sm)
    # batch verification of the same signatures (the idx-1023 one must fail)
    pks = [buf(bytes.fromhex(it["pk"])) for it in items]
    smb = [buf(s) for s in sm_out]
    outs = [buf(len(s)) for s in sm_out]
    smls = (C.c_ulonglong * k)(*[len(s) for s in sm_out])
    mls = (C.c_ulonglong * k)()
    vr = (C.c_int * k)()
    assert L.xmss_batch_sign_open(0, _arr(outs), mls, _arr(smb), smls, _arr(pks), vr, k) == 0
    for i, it in enumerate(items):
        want = 0 if it["idx"] < 1023 else -1
        assert vr[i] == want, "verify item %d" % i
        if want == 0:
            assert raw(outs[i], mls[i]) == msgs[i]


def test_three_step_job_api_with_temporary_pointer_arrays():
    """prepare / run / collect with pointer arrays that do not outlive the
    prepare call (the job keeps its own copies)."""
    g = golden("plain_batch.json")
    lib = product(0)
    L = lib.lib
    L.xmss_mi355x_sign_prepare.restype = C.c_void_p
    L.xmss_mi355x_sign_prepare.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    L.xmss_mi355x_job_run.argtypes = [C.c_void_p]
    L.xmss_mi355x_sign_collect.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.xmss_mi355x_job_free.argtypes = [C.c_void_p]
    p = lib.params(g["oid"], 0)
    items = g["items"][:4]
    k = len(items)
    sks = [buf(bytes.fromhex(it["sk_in"])) for it in items]
    msgs = [bytes.fromhex(it["msg"]) for it in items]
    mbufs = [buf(m) for m in msgs]
    mlens = (C.c_ulonglong * k)(*[len(m) for m in msgs])
    job = L.xmss_mi355x_sign_prepare(0, _arr(sks), _arr(mbufs), mlens, k)
    import gc
    gc.collect()
    assert L.xmss_mi355x_job_run(job) == 0
    sms = [buf(p.sig_bytes + len(m)) for m in msgs]
    smlens = (C.c_ulonglong * k)()
    rets = (C.c_int * k)()
    assert L.xmss_mi355x_sign_collect(job, _arr(sms), smlens, rets) == 0
    L.xmss_mi355x_job_free(job)
    for i, it in enumerate(items):
        assert rets[i] == it["rc"]
        assert raw(sms[i], smlens[i]).hex() == it["sm"]


MIXED_SETS = [("XMSS-SHA2_10_256", 0), ("XMSSMT-SHA2_20/4_256", 1), ("XMSS-SHAKE_10_256", 0),
              ("XMSSMT-SHA2_20/4_512", 1), ("XMSS-SHA2_10_192", 0), ("XMSSMT-SHAKE256_20/4_192", 1)]


def test_mixed_parameter_sets_in_one_job():
    """One job over six parameter sets of five hash families (each set on its
    own engine): every signature, sk and return value equals the single
    xmss[mt]_sign call on the same key, including a last-index (wiped key)
    and an exhausted key; the mixed verification job accepts exactly the
    valid ones and rejects an unknown OID."""
    import random
    lib = product(0)
    L = lib.lib
    L.xmss_mi355x_mixed_sign_prepare.restype = C.c_void_p
    L.xmss_mi355x_mixed_sign_prepare.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    L.xmss_mi355x_mixed_verify_prepare.restype = C.c_void_p
    L.xmss_mi355x_mixed_verify_prepare.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    L.xmss_mi355x_job_run.argtypes = [C.c_void_p]
    L.xmss_mi355x_sign_collect.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.xmss_mi355x_verify_collect.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.xmss_mi355x_job_free.argtypes = [C.c_void_p]
    rng = random.Random(4242)
    cases = []  # (mt, p, pk, sk_in, msg)
    for name, mt in MIXED_SETS:
        oid = lib.str_to_oid(name)
        p = lib.params(oid, mt)
        L.xmss_mi355x_rng_seed(oid * 7 + mt)
        rc, pk, sk = lib.keypair(oid, mt)
        assert rc == 0
        top = (1 << p.full_height) - 1
        idxs = [rng.randrange(0, top), rng.randrange(0, top)]
        if name == "XMSS-SHA2_10_256":
            idxs += [top, top + 1]  # last index (wiped key) and exhausted
        for idx in idxs:
            s = bytearray(sk)
            s[4:4 + p.index_bytes] = idx.to_bytes(p.index_bytes, "big")
            m = bytes(rng.getrandbits(8) for _ in range(int(32 * 64 ** rng.random())))
            cases.append((mt, p, pk, bytes(s), m))
    rng.shuffle(cases)
    k = len(cases)
    want = [lib.sign(c[3], c[4], c[0]) for c in cases]
    mts = (C.c_int * k)(*[c[0] for c in cases])
    sks = [buf(c[3]) for c in cases]
    mb = [buf(c[4]) for c in cases]
    mlens = (C.c_ulonglong * k)(*[len(c[4]) for c in cases])
    sms = [buf(c[1].sig_bytes + len(c[4])) for c in cases]
    smlens = (C.c_ulonglong * k)()
    rets = (C.c_int * k)()
    job = L.xmss_mi355x_mixed_sign_prepare(mts, _arr(sks), _arr(mb), mlens, k)
    assert L.xmss_mi355x_job_run(job) == 0
    assert L.xmss_mi355x_sign_collect(job, _arr(sms), smlens, rets) == 0
    L.xmss_mi355x_job_free(job)
    for i, (c, w) in enumerate(zip(cases, want)):
        assert rets[i] == w[0], "item %d rc" % i
        if w[0] == 0:
            assert raw(sms[i], smlens[i]) == w[1], "item %d signature" % i
        assert raw(sks[i], len(c[3])) == w[2], "item %d sk" % i
    # verification: the signed items, plus one with an unknown OID
    vi = [i for i in range(k) if rets[i] == 0]
    vsm = [raw(sms[i], smlens[i]) for i in vi]
    vpk = [cases[i][2] for i in vi]
    vmt = [cases[i][0] for i in vi]
    vsm.append(vsm[0])
    vpk.append(b"\xff\xff\xff\xff" + vpk[0][4:])
    vmt.append(vmt[0])
    kv = len(vsm)
    smb = [buf(s) for s in vsm]
    outs = [buf(len(s)) for s in vsm]
    smls = (C.c_ulonglong * kv)(*[len(s) for s in vsm])
    mls = (C.c_ulonglong * kv)()
    vr = (C.c_int * kv)()
    pkb = [buf(x) for x in vpk]  # must outlive the job
    L.xmss_mi355x_verify_refill.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    job = L.xmss_mi355x_mixed_verify_prepare((C.c_int * kv)(*vmt), _arr(smb), smls, _arr(pkb), kv)

    def check(flipped):
        for t, i in enumerate(vi):
            mt, p = cases[i][0], cases[i][1]
            idx = int.from_bytes(bytes(cases[i][3][4:4 + p.index_bytes]), "big")
            ok = idx < (1 << p.full_height) - 1 and t not in flipped
            assert vr[t] == (0 if ok else -1), "verify item %d" % i
            if ok:
                assert raw(outs[t], mls[t]) == cases[i][4]
        assert vr[kv - 1] == -1

    # three runs (the second and third from each set's captured graph) ...
    for _ in range(3):
        assert L.xmss_mi355x_job_run(job) == 0
        assert L.xmss_mi355x_verify_collect(job, _arr(outs), mls, vr) == 0
        check(set())
    # ... then a refill of the same shape with two signatures corrupted, twice
    flipped = {0, kv - 2}
    vsm2 = list(vsm)
    for t in flipped:
        b2 = bytearray(vsm2[t])
        b2[cases[vi[t]][1].index_bytes + 5] ^= 4
        vsm2[t] = bytes(b2)
    smb2 = [buf(s) for s in vsm2]
    assert L.xmss_mi355x_verify_refill(job, _arr(smb2), smls, _arr(pkb)) == 0
    for _ in range(2):
        assert L.xmss_mi355x_job_run(job) == 0
        assert L.xmss_mi355x_verify_collect(job, _arr(outs), mls, vr) == 0
        check(flipped)
    L.xmss_mi355x_job_free(job)


@pytest.mark.parametrize("name,mt", [("XMSS-SHA2_10_256", 0), ("XMSSMT-SHA2_20/4_256", 1),
                                     ("XMSS-SHA2_10_512", 0)])
def test_large_verify_batch_accepts_exactly_the_valid(name, mt):
    """A verification batch of 300 requests (the dense path: requests staged
    longest message first, in two groups): valid signatures of 32 B - 4 KB
    messages in shuffled order, and ones with a flipped byte in R, in a WOTS
    signature, in an auth path, in an upper layer or in the message; every
    verdict lands on its own request and equals the single sign_open call."""
    import random
    from test_gpu_opk import _sign_many
    lib = product(0)
    L = lib.lib
    oid = lib.str_to_oid(name)
    p = lib.params(oid, mt)
    L.xmss_mi355x_rng_seed(oid + 17 * mt)
    rc, pk, sk = lib.keypair(oid, mt)
    assert rc == 0
    rng = random.Random(oid * 3 + mt)
    k = 300
    msgs = [bytes(rng.getrandbits(8) for _ in range(int(32 * 128 ** rng.random()))) for _ in range(k)]
    out, _ = _sign_many(lib, sk, msgs, mt, p)
    sms = []
    ib, n, wb = p.index_bytes, p.n, p.wots_sig_bytes
    layer = wb + p.tree_height * n
    for i, (rc, sm) in enumerate(out):
        assert rc == 0
        sm = bytearray(sm)
        kind = i % 6
        if kind == 1:
            sm[ib + rng.randrange(n)] ^= 1                           # R
        elif kind == 2:
            sm[ib + n + rng.randrange(wb)] ^= 0x40                   # layer-0 WOTS signature
        elif kind == 3:
            sm[ib + n + wb + rng.randrange(p.tree_height * n)] ^= 8  # layer-0 auth path
        elif kind == 4:
            sm[ib + n + layer * (p.d - 1) + rng.randrange(layer)] ^= 2  # top layer
        elif kind == 5 and len(msgs[i]):
            sm[p.sig_bytes + rng.randrange(len(msgs[i]))] ^= 1      # message
        sms.append(bytes(sm))
    order = list(range(k))
    rng.shuffle(order)
    sms = [sms[i] for i in order]
    want = []
    for i, s in zip(order, sms):
        ok = i % 6 == 0 or (i % 6 == 5 and not msgs[i])
        want.append(0 if ok else -1)
    pkb = [buf(pk) for _ in range(k)]
    smb = [buf(s) for s in sms]
    outs = [buf(len(s)) for s in sms]
    smls = (C.c_ulonglong * k)(*[len(s) for s in sms])
    mls = (C.c_ulonglong * k)()
    vr = (C.c_int * k)()
    L.xmss_batch_sign_open.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_size_t]
    assert L.xmss_batch_sign_open(mt, _arr(outs), mls, _arr(smb), smls, _arr(pkb), vr, k) == 0
    for j in range(k):
        assert vr[j] == want[j], "request %d (signature %d, kind %d)" % (j, order[j], order[j] % 6)
        if want[j] == 0:
            assert raw(outs[j], mls[j]) == msgs[order[j]]
    # a few against the one-call API
    for j in range(0, k, 37):
        rc, m = lib.sign_open(sms[j], pk, mt)
        assert rc == vr[j]


def _verify_job_api(L):
    L.xmss_mi355x_verify_prepare.restype = C.c_void_p
    L.xmss_mi355x_verify_prepare.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    L.xmss_mi355x_verify_refill.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.xmss_mi355x_job_run.argtypes = [C.c_void_p]
    L.xmss_mi355x_verify_collect.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.xmss_mi355x_job_free.argtypes = [C.c_void_p]


@pytest.mark.parametrize("k", [300, 40])  # the large-batch path and the smaller one
def test_verify_job_replays_and_refills(k):
    """A verification job run five times (direct launches, then captured into
    a HIP graph and replayed), with the host collecting and checking between
    runs, and then refilled twice with new inputs of the same shape (another
    order and other corrupted signatures): every run's verdicts equal the
    single sign_open calls.  A refill whose public keys change parameter set
    is refused."""
    import random
    from test_gpu_opk import _sign_many
    lib = product(0)
    L = lib.lib
    _verify_job_api(L)
    oid, mt = 0x01, 0
    p = lib.params(oid, mt)
    L.xmss_mi355x_rng_seed(7 + k)
    rc, pk, sk = lib.keypair(oid, mt)
    assert rc == 0
    rng = random.Random(k)
    msgs = [bytes(rng.getrandbits(8) for _ in range(int(32 * 128 ** rng.random()))) for _ in range(k)]
    out, _ = _sign_many(lib, sk, msgs, mt, p)
    good = [sm for rc, sm in out]

    def variant(seed):
        r = random.Random(seed)
        order = list(range(k))
        r.shuffle(order)
        sms = []
        for i in order:
            sm = bytearray(good[i])
            if r.random() < 0.3:
                sm[r.randrange(p.index_bytes, p.sig_bytes)] ^= 1 << r.randrange(8)
            sms.append(bytes(sm))
        want = [lib.sign_open(s, pk, mt)[0] for s in sms]
        return sms, want

    pkb = [buf(pk) for _ in range(k)]

    def check(job, sms, want, label):
        outs = [buf(len(s)) for s in sms]
        mls = (C.c_ulonglong * k)()
        vr = (C.c_int * k)()
        assert L.xmss_mi355x_verify_collect(job, _arr(outs), mls, vr) == 0
        assert [vr[j] for j in range(k)] == want, label
        assert 0 in want and -1 in want

    sms, want = variant(1)
    smb = [buf(s) for s in sms]
    job = L.xmss_mi355x_verify_prepare(mt, _arr(smb), (C.c_ulonglong * k)(*[len(s) for s in sms]), _arr(pkb), k)
    for run in range(5):
        assert L.xmss_mi355x_job_run(job) == 0
        check(job, sms, want, "run %d" % run)
    for seed in (2, 3):
        sms2, want2 = variant(seed)
        smb2 = [buf(s) for s in sms2]
        assert L.xmss_mi355x_verify_refill(job, _arr(smb2), (C.c_ulonglong * k)(*[len(s) for s in sms2]),
                                           _arr(pkb)) == 0
        for run in range(2):
            assert L.xmss_mi355x_job_run(job) == 0
            check(job, sms2, want2, "refill %d run %d" % (seed, run))
    # another parameter set at position 0: refused, the job unchanged
    other = bytearray(pk)
    other[3] = 0x02
    bad = [buf(bytes(other))] + pkb[1:]
    assert L.xmss_mi355x_verify_refill(job, _arr(smb2), (C.c_ulonglong * k)(*[len(s) for s in sms2]),
                                       _arr(bad)) == -1
    assert L.xmss_mi355x_job_run(job) == 0
    check(job, sms2, want2, "after a refused refill")
    L.xmss_mi355x_job_free(job)


@pytest.mark.parametrize("k", [300, 40])
def test_verify_jobs_pipelined(k):
    """Two verification jobs alternating over a stream of six batches, each
    batch staged (refill) while the other job's launched work runs, and
    launched before or after the other job is collected:
    launch(A); loop { refill(B); launch(B); collect(A); swap } and
    launch(A); loop { refill(B); collect(A); launch(B); swap } in turn.
    Every batch's verdicts equal the single sign_open calls.  A refill of a
    job still in flight, a launch with no wait and a free of a job in flight
    wait for its work by themselves."""
    import random
    from test_gpu_opk import _sign_many
    lib = product(0)
    L = lib.lib
    _verify_job_api(L)
    L.xmss_mi355x_job_launch.argtypes = [C.c_void_p]
    L.xmss_mi355x_job_wait.argtypes = [C.c_void_p]
    oid, mt = 0x01, 0
    p = lib.params(oid, mt)
    L.xmss_mi355x_rng_seed(11 + k)
    rc, pk, sk = lib.keypair(oid, mt)
    assert rc == 0
    rng = random.Random(100 + k)
    msgs = [bytes(rng.getrandbits(8) for _ in range(int(32 * 128 ** rng.random()))) for _ in range(k)]
    out, _ = _sign_many(lib, sk, msgs, mt, p)
    good = [sm for rc, sm in out]
    pkb = [buf(pk) for _ in range(k)]

    def variant(seed):
        r = random.Random(seed)
        order = list(range(k))
        r.shuffle(order)
        sms = []
        for i in order:
            sm = bytearray(good[i])
            if r.random() < 0.3:
                sm[r.randrange(p.index_bytes, p.sig_bytes)] ^= 1 << r.randrange(8)
            sms.append(bytes(sm))
        want = [lib.sign_open(s, pk, mt)[0] for s in sms]
        return [buf(s) for s in sms], (C.c_ulonglong * k)(*[len(s) for s in sms]), want

    def collect(job, batch, label):
        smb, smls, want = batch
        outs = [buf(smls[j]) for j in range(k)]
        mls = (C.c_ulonglong * k)()
        vr = (C.c_int * k)()
        assert L.xmss_mi355x_verify_collect(job, _arr(outs), mls, vr) == 0
        assert [vr[j] for j in range(k)] == want, label
        for j in range(k):
            if want[j] == 0:
                assert raw(outs[j], mls[j]) == bytes(smb[j][p.sig_bytes:smls[j]]), label
        assert 0 in want and -1 in want

    batches = [variant(s) for s in range(1, 7)]
    jobs = [L.xmss_mi355x_verify_prepare(mt, _arr(b[0]), b[1], _arr(pkb), k) for b in batches[:2]]
    held = [batches[0], batches[1]]  # the inputs each job holds
    cur = 0
    assert L.xmss_mi355x_job_launch(jobs[cur]) == 0
    for i in range(1, len(batches)):
        nxt = 1 - cur
        b = batches[i]
        assert L.xmss_mi355x_verify_refill(jobs[nxt], _arr(b[0]), b[1], _arr(pkb)) == 0
        held[nxt] = b
        if i % 2:
            assert L.xmss_mi355x_job_launch(jobs[nxt]) == 0
            collect(jobs[cur], held[cur], "batch %d" % (i - 1))
        else:
            collect(jobs[cur], held[cur], "batch %d" % (i - 1))
            assert L.xmss_mi355x_job_launch(jobs[nxt]) == 0
        cur = nxt
    assert L.xmss_mi355x_job_wait(jobs[cur]) == 0
    collect(jobs[cur], held[cur], "last batch")
    # refill while in flight: waits, then stages the new inputs
    assert L.xmss_mi355x_job_launch(jobs[0]) == 0
    b = batches[2]
    assert L.xmss_mi355x_verify_refill(jobs[0], _arr(b[0]), b[1], _arr(pkb)) == 0
    assert L.xmss_mi355x_job_run(jobs[0]) == 0
    collect(jobs[0], b, "refilled in flight")
    assert L.xmss_mi355x_job_launch(jobs[1]) == 0
    L.xmss_mi355x_job_free(jobs[1])  # in flight
    L.xmss_mi355x_job_free(jobs[0])


def test_sign_job_launched_beside_a_verify_job():
    """A signing job and a verification job launched together and collected
    in the other order: the signatures equal those of job_run."""
    import random
    lib = product(0)
    L = lib.lib
    _verify_job_api(L)
    L.xmss_mi355x_job_launch.argtypes = [C.c_void_p]
    L.xmss_mi355x_job_wait.argtypes = [C.c_void_p]
    L.xmss_mi355x_sign_prepare.restype = C.c_void_p
    L.xmss_mi355x_sign_prepare.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    L.xmss_mi355x_sign_collect.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    oid, mt = 0x01, 0
    p = lib.params(oid, mt)
    rng = random.Random(77)
    k = 6
    keys = []
    for i in range(k):
        L.xmss_mi355x_rng_seed(500 + i)
        rc, pk, sk = lib.keypair(oid, mt)
        assert rc == 0
        sk = sk[:4] + (rng.randrange(1000)).to_bytes(4, "big") + sk[8:]
        keys.append((pk, sk))
    msgs = [bytes(rng.getrandbits(8) for _ in range(rng.randrange(1, 3000))) for _ in range(k)]

    def sign_job(launch_only):
        skb = [buf(sk) for _, sk in keys]
        msb = [buf(m) for m in msgs]
        mls = (C.c_ulonglong * k)(*[len(m) for m in msgs])
        job = L.xmss_mi355x_sign_prepare(mt, _arr(skb), _arr(msb), mls, k)
        return job, skb, msb, mls

    def sign_collect(job, skb):
        sms = [buf(p.sig_bytes + len(m)) for m in msgs]
        smls = (C.c_ulonglong * k)()
        rets = (C.c_int * k)()
        assert L.xmss_mi355x_sign_collect(job, _arr(sms), smls, rets) == 0
        assert all(rets[i] == 0 for i in range(k))
        return [raw(sms[i], smls[i]) for i in range(k)], [raw(skb[i], 4 + p.sk_bytes) for i in range(k)]

    job, skb, _m1, _l1 = sign_job(False)
    assert L.xmss_mi355x_job_run(job) == 0
    want_sm, want_sk = sign_collect(job, skb)
    L.xmss_mi355x_job_free(job)
    # a verification job of those signatures, and the same signing job again
    vsmb = [buf(s) for s in want_sm]
    vpkb = [buf(pk) for pk, _ in keys]
    vjob = L.xmss_mi355x_verify_prepare(mt, _arr(vsmb), (C.c_ulonglong * k)(*[len(s) for s in want_sm]),
                                        _arr(vpkb), k)
    sjob, skb2, _m2, _l2 = sign_job(True)
    assert L.xmss_mi355x_job_launch(sjob) == 0
    assert L.xmss_mi355x_job_launch(vjob) == 0
    outs = [buf(len(s)) for s in want_sm]
    mls = (C.c_ulonglong * k)()
    vr = (C.c_int * k)()
    assert L.xmss_mi355x_verify_collect(vjob, _arr(outs), mls, vr) == 0
    assert [vr[i] for i in range(k)] == [0] * k
    assert [raw(outs[i], mls[i]) for i in range(k)] == msgs
    assert L.xmss_mi355x_job_wait(sjob) == 0
    got_sm, got_sk = sign_collect(sjob, skb2)
    assert got_sm == want_sm and got_sk == want_sk
    L.xmss_mi355x_job_free(sjob)
    L.xmss_mi355x_job_free(vjob)


def test_empty_and_all_rejected_batches():
    """Batches that reach no device work: zero inputs (every batch entry point
    returns 0), and a batch whose keys are all exhausted or of an unknown OID
    (each ret and sk as xmss_sign leaves them: -2 with the key wiped, -1 with
    the sk untouched)."""
    lib = product(0)
    o = oracle()
    L = lib.lib
    vp = C.c_void_p
    L.xmss_batch_sign.argtypes = [C.c_int, vp, vp, vp, vp, vp, vp, C.c_size_t]
    L.xmss_batch_sign_open.argtypes = [C.c_int, vp, vp, vp, vp, vp, vp, C.c_size_t]
    L.xmss_batch_keypair.argtypes = [C.c_int, vp, vp, vp, C.c_uint32, C.c_size_t]
    L.xmss_mi355x_sign_many.argtypes = [C.c_int, vp, vp, vp, vp, vp, C.c_size_t, vp]
    for nm in ("xmss_mi355x_sign_prepare", "xmss_mi355x_verify_prepare"):
        getattr(L, nm).restype = vp
        getattr(L, nm).argtypes = [C.c_int, vp, vp, vp, C.c_size_t]
    for nm in ("xmss_mi355x_mixed_sign_prepare", "xmss_mi355x_mixed_verify_prepare"):
        getattr(L, nm).restype = vp
        getattr(L, nm).argtypes = [vp, vp, vp, vp, C.c_size_t]
    L.xmss_mi355x_job_run.argtypes = [vp]
    L.xmss_mi355x_sign_collect.argtypes = [vp, vp, vp, vp]
    L.xmss_mi355x_verify_collect.argtypes = [vp, vp, vp, vp]
    L.xmss_mi355x_job_free.argtypes = [vp]
    e = _arr([])
    z = (C.c_ulonglong * 1)()
    r = (C.c_int * 1)()
    assert L.xmss_batch_sign(0, e, e, z, e, z, r, 0) == 0
    assert L.xmss_batch_sign_open(1, e, z, e, z, e, r, 0) == 0
    assert L.xmss_batch_keypair(0, e, e, None, 0x00000001, 0) == 0
    sk1 = buf((1).to_bytes(4, "big") + bytes(132))
    assert L.xmss_mi355x_sign_many(0, sk1, e, z, e, z, 0, r) == 0
    for prep, collect in (("xmss_mi355x_sign_prepare", "xmss_mi355x_sign_collect"),
                          ("xmss_mi355x_verify_prepare", "xmss_mi355x_verify_collect")):
        job = getattr(L, prep)(0, e, z, e, 0)
        assert job
        assert L.xmss_mi355x_job_run(job) == 0
        assert getattr(L, collect)(job, e, z, r) == 0
        L.xmss_mi355x_job_free(job)
    mts = (C.c_int * 1)()
    job = L.xmss_mi355x_mixed_verify_prepare(mts, e, z, e, 0)
    assert job and L.xmss_mi355x_job_run(job) == 0
    L.xmss_mi355x_job_free(job)

    # keys past their last index (XMSS-SHA2_10_256: 1024, 2^32 - 1) and an unknown OID
    sk = (1).to_bytes(4, "big") + bytes((37 * i + 11) & 255 for i in range(132))
    cases = []
    for idx in (1024, 0xFFFFFFFF):
        s = bytearray(sk)
        s[4:8] = idx.to_bytes(4, "big")
        cases.append(bytes(s))
    bad = bytearray(sk)
    bad[0:4] = (0x7FFF).to_bytes(4, "big")
    cases.append(bytes(bad))
    k = len(cases)
    sks = [buf(s) for s in cases]
    msgs = [bytes([i]) * 32 for i in range(k)]
    mbufs = [buf(m) for m in msgs]
    sms = [buf(2500 + 32) for _ in range(k)]
    smlens = (C.c_ulonglong * k)()
    mlens = (C.c_ulonglong * k)(*[32] * k)
    rets = (C.c_int * k)()
    mts = (C.c_int * k)()
    job = L.xmss_mi355x_mixed_sign_prepare(mts, _arr(sks), _arr(mbufs), mlens, k)
    assert job and L.xmss_mi355x_job_run(job) == 0
    assert L.xmss_mi355x_sign_collect(job, _arr(sms), smlens, rets) == 0
    L.xmss_mi355x_job_free(job)
    for i in range(2):
        rc, _, sko = o.sign(cases[i], msgs[i], 0, 0)
        assert rc == -2 and rets[i] == -2
        assert raw(sks[i], len(cases[i])) == sko
    assert rets[2] == -1
    assert raw(sks[2], len(cases[2])) == cases[2]
