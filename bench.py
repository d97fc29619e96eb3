"""Benchmark of the MI355X XMSS engine on its headline workloads.

Metric: XMSS-SHA2_16_256 keygen seconds; signs/s + verifies/s at batch 4096;
p99 sign latency.

  step     one XMSS-SHA2_16_256 keygen (65,536 leaves, 298,582,010 SHA-256
           compressions): the value, in seconds.
  batch    4,096 independent XMSS-SHA2_10_256 keys, one plain-core signature
           each at a random leaf, messages log-uniform 32 B - 4 KB; then the
           4,096 verifications.  Inputs are staged in HBM before the timed
           region (xmss_mi355x_sign_prepare), results copied out after it.
  latency  1,024 sequential BDS-core signatures of one XMSS-SHA2_16_256 key
           (p50 / p99 per call, each a full xmss_sign call).

Prints one JSON line.  Kernel durations come from HIP events recorded on the
library's stream around each launch.
"""
import argparse
import ctypes as C
import gc
import json
import math
import os
import random
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tests"))
from ffi import Api, ORACLE_SO, PRODUCT_DIR, REF_DIR, Oracle, buf, load_ref, raw  # noqa: E402

METRIC = "XMSS-SHA2_16_256 keygen s; signs/s + verifies/s at batch 4096; p99 sign latency"
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # single-issue VALU lane-ops/s (256 CU x 64 lanes x 2.4 GHz): 39.32 T
# The roofline's ceiling: twice that, the VALU lane rate behind MI355X's
# published 157.3 TFLOP/s FP32 vector peak (78.6 T FMA lanes/s).
VALU_CEILING_TOPS = 2 * VALU_PEAK_TOPS
OPS_PER_SHA256 = 1384  # VALU ops of one SHA-256 compression (64 x 14 round + 48 x 10 schedule + 8)
SEED = 0x584D5353


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_product(core):
    api = Api(os.path.join(PRODUCT_DIR, "libxmss_mi355x_fast.so" if core else "libxmss_mi355x.so"), core)
    L = api.lib
    L.xmss_mi355x_device_ok.restype = C.c_int
    L.xmss_mi355x_profile.argtypes = [C.c_int]
    L.xmss_mi355x_kernel_stats.argtypes = [C.c_char_p, C.c_size_t]
    L.xmss_mi355x_kernel_stats.restype = C.c_size_t
    L.xmss_mi355x_rng_seed.argtypes = [C.c_uint64]
    for nm in ("xmss_mi355x_sign_prepare", "xmss_mi355x_verify_prepare"):
        getattr(L, nm).restype = C.c_void_p
        getattr(L, nm).argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t] if "sign" in nm else \
            [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    L.xmss_mi355x_job_run.argtypes = [C.c_void_p]
    L.xmss_mi355x_sign_collect.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.xmss_mi355x_verify_collect.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.xmss_mi355x_verify_refill.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.xmss_mi355x_job_launch.argtypes = [C.c_void_p]
    L.xmss_mi355x_job_wait.argtypes = [C.c_void_p]
    L.xmss_mi355x_job_free.argtypes = [C.c_void_p]
    L.xmss_batch_keypair.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_char_p, C.c_uint32, C.c_size_t]
    return api


def tuned_plans(L):
    """The verification schedules the library's autotuner settled on, one
    line per shape (parameter family, tree height, layers, batch size)."""
    L.xmss_mi355x_tuned_plans.argtypes = [C.c_char_p, C.c_size_t]
    L.xmss_mi355x_tuned_plans.restype = C.c_size_t
    n = L.xmss_mi355x_tuned_plans(None, 0)
    b = C.create_string_buffer(n + 1)
    L.xmss_mi355x_tuned_plans(b, n + 1)
    return b.value.decode().splitlines()


def kernel_stats(L):
    n = L.xmss_mi355x_kernel_stats(None, 0)
    b = C.create_string_buffer(n + 1)
    L.xmss_mi355x_kernel_stats(b, n + 1)
    out = {}
    for line in b.value.decode().splitlines():
        name, cnt, ms = line.split()
        out[name] = (int(cnt), float(ms))
    return out


def ptrs(bufs):
    a = (C.c_void_p * len(bufs))()
    for i, b in enumerate(bufs):
        a[i] = C.cast(b, C.c_void_p)
    return a


def keygen_h16(api):
    """One XMSS-SHA2_16_256 keygen (plain core: seed keypair, xmss_core.c:134)."""
    rc, pk, sk = api.seed_keypair(0x02, 0, bytes(range(96)))
    assert rc == 0
    return pk


def bench_keygen(api, steps, warmup):
    L = api.lib
    for _ in range(warmup):
        keygen_h16(api)
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        pk = keygen_h16(api)
        times.append(time.perf_counter() - t0)
    # one profiled keygen: per-kernel HIP-event durations
    L.xmss_mi355x_reset_stats()
    L.xmss_mi355x_profile(1)
    keygen_h16(api)
    L.xmss_mi355x_profile(0)
    st = kernel_stats(L)
    L.xmss_mi355x_reset_stats()
    return times, st, pk


def batch_inputs(count, rng):
    """XMSS-SHA2_10_256 keys from their own 96-byte seeds, a uniform leaf in
    [0, 1023] each and log-uniform 32 B - 4 KB messages."""
    seeds = bytes(rng.getrandbits(8) for _ in range(96 * count))
    idxs = [rng.randrange(0, 1024) for _ in range(count)]
    msgs = [bytes(rng.getrandbits(8) for _ in range(int(32 * 128 ** rng.random()))) for _ in range(count)]
    return seeds, idxs, msgs


def bench_batch(api, count, stream_batches=16):
    """Batch keygen, then one plain-core signature per key, then verification
    of all of them.  Returns timings, kernel stats and the verify results."""
    L = api.lib
    rng = random.Random(SEED + 7)
    seeds, idxs, msgs = batch_inputs(count, rng)
    p = api.params(0x01, 0)
    pks = [buf(4 + p.pk_bytes) for _ in range(count)]
    sks = [buf(4 + p.sk_bytes) for _ in range(count)]
    pkp, skp = ptrs(pks), ptrs(sks)
    t0 = time.perf_counter()
    assert L.xmss_batch_keypair(0, pkp, skp, seeds, 0x01, count) == 0
    keygen_t = time.perf_counter() - t0
    for i in range(count):
        sks[i][4:8] = list(idxs[i].to_bytes(4, "big"))
    ms = [buf(m) for m in msgs]
    mlens = (C.c_ulonglong * count)(*[len(m) for m in msgs])
    sms = [buf(p.sig_bytes + len(m)) for m in msgs]
    smlens = (C.c_ulonglong * count)()
    rets = (C.c_int * count)()
    msp, smp = ptrs(ms), ptrs(sms)
    t0 = time.perf_counter()
    job = L.xmss_mi355x_sign_prepare(0, skp, msp, mlens, count)
    t1 = time.perf_counter()
    L.xmss_mi355x_reset_stats()
    L.xmss_mi355x_profile(1)
    L.xmss_mi355x_job_run(job)
    t2 = time.perf_counter()
    L.xmss_mi355x_profile(0)
    st = kernel_stats(L)
    L.xmss_mi355x_sign_collect(job, smp, smlens, rets)
    L.xmss_mi355x_job_free(job)
    t3 = time.perf_counter()
    assert all(rets[i] == 0 for i in range(count))
    sign_run, sign_total = t2 - t1, t3 - t0

    sm_bytes = [raw(sms[i], smlens[i]) for i in range(count)]
    smb = [buf(x) for x in sm_bytes]
    outs = [buf(len(x)) for x in sm_bytes]
    smls = (C.c_ulonglong * count)(*[len(x) for x in sm_bytes])
    mls = (C.c_ulonglong * count)()
    vr = (C.c_int * count)()
    job = L.xmss_mi355x_verify_prepare(0, ptrs(smb), smls, ptrs(pks), count)
    # the same staged job runs twice untimed (the first run loads the kernels
    # and sizes the buffers, the second captures the launches into a HIP
    # graph), then three timed runs replay the graph: the median
    L.xmss_mi355x_job_run(job)
    L.xmss_mi355x_job_run(job)
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        L.xmss_mi355x_job_run(job)
        runs.append(time.perf_counter() - t0)
    ver_t = sorted(runs)[1]
    L.xmss_mi355x_verify_collect(job, ptrs(outs), mls, vr)
    # a stream of same-shape batches: refill (host -> HBM), run, collect
    # (HBM -> host) per batch, here the same signatures in another order
    perm = list(range(count))
    random.Random(3).shuffle(perm)
    smb2 = [smb[i] for i in perm]
    smls2 = (C.c_ulonglong * count)(*[smls[i] for i in perm])
    pks2 = [pks[i] for i in perm]
    outs2 = [outs[i] for i in perm]
    vr2 = (C.c_int * count)()
    mls2 = (C.c_ulonglong * count)()
    smp2, pkp2, outp2 = ptrs(smb2), ptrs(pks2), ptrs(outs2)
    t0 = time.perf_counter()
    assert L.xmss_mi355x_verify_refill(job, smp2, smls2, pkp2) == 0
    L.xmss_mi355x_job_run(job)
    L.xmss_mi355x_verify_collect(job, outp2, mls2, vr2)
    ver_refill_t = time.perf_counter() - t0
    assert [vr2[j] for j in range(count)] == [vr[i] for i in perm]
    # the same stream through two jobs: each batch is staged while the other
    # job's launched work runs, and collected after the next one is launched
    stream_t = None
    if stream_batches:
        stream_t = bench_verify_stream(L, job, count, (ptrs(smb), smls, ptrs(pks), ptrs(outs)),
                                       (smp2, smls2, pkp2, outp2), [vr[i] for i in range(count)],
                                       [vr[i] for i in perm], nbatch=stream_batches)
    L.xmss_mi355x_job_free(job)
    # leaf 1023 is the last index: the reference signs it with the wiped key,
    # so exactly those signatures must be rejected (xmss_core.c:222-231)
    for i in range(count):
        want = 0 if idxs[i] < 1023 else -1
        assert vr[i] == want, (i, idxs[i], vr[i])
        if want == 0:
            assert raw(outs[i], mls[i]) == msgs[i]
    # nominal work per request (cached SHA-256 compressions, n = 32): a
    # signature rebuilds its 1024-leaf tree (4,665,338) and adds R (2), H_msg
    # and its WOTS signature (2 + 4 x E[steps] = 2 + 4 x 7.5 x 67); a
    # verification is H_msg, the remaining chain steps (4 x 7.5 x 67), the
    # L-tree (66 x 6) and the auth path (10 x 6)
    hmsg = sum(math.ceil((128 + len(m) + 9) / 64) for m in msgs)
    sign_comps = count * (4665338 + 2 + 2 + 2010) + hmsg
    ver_comps = count * (2010 + 396 + 60) + hmsg
    frac = lambda comps, t: round(comps * OPS_PER_SHA256 / t / 1e12 / VALU_CEILING_TOPS, 4)
    nominal = {"sign_comps": sign_comps, "sign_frac": frac(sign_comps, sign_run),
               "verify_comps": ver_comps, "verify_frac": frac(ver_comps, ver_t)}
    return (keygen_t, sign_run, sign_total, (ver_t, ver_refill_t, stream_t), st,
            sum(1 for i in idxs if i == 1023), nominal)


def bench_verify_stream(L, job_a, count, in_a, in_b, want_a, want_b, nbatch=16, njobs=3):
    """Seconds per batch of a stream of `nbatch` same-shape verification
    batches (the two input sets alternating) through `njobs` jobs in
    rotation: batch i is refilled into job i mod njobs, then batch
    i - njobs + 1 is collected and batch i launched.  Host staging and the
    PCIe copy of one batch overlap the GPU work of the others, and with three
    jobs the GPU has the next batch queued while the host collects; every
    batch's verdicts are checked after the timed region."""
    jobs = [job_a]
    for _ in range(njobs - 1):
        j = L.xmss_mi355x_verify_prepare(0, in_b[0], in_b[1], in_b[2], count)
        L.xmss_mi355x_job_run(j)  # sizes the buffers, then captures the graph
        L.xmss_mi355x_job_run(j)
        jobs.append(j)
    sets = [in_a, in_b]
    mls = [(C.c_ulonglong * count)() for _ in range(nbatch)]
    vrs = [(C.c_int * count)() for _ in range(nbatch)]
    t0 = time.perf_counter()
    for i in range(nbatch + njobs - 1):
        if i < nbatch:
            src = sets[i % 2]
            assert L.xmss_mi355x_verify_refill(jobs[i % njobs], src[0], src[1], src[2]) == 0
        k = i - (njobs - 1)
        if k >= 0:
            L.xmss_mi355x_verify_collect(jobs[k % njobs], sets[k % 2][3], mls[k], vrs[k])
        if i < nbatch:
            L.xmss_mi355x_job_launch(jobs[i % njobs])
    t = time.perf_counter() - t0
    for j in jobs[1:]:
        L.xmss_mi355x_job_free(j)
    for i in range(nbatch):
        assert [vrs[i][j] for j in range(count)] == (want_a if i % 2 == 0 else want_b), "stream batch %d" % i
    return t / nbatch


def bench_plain_single(api, count):
    """The plain one-request-per-launch form of the batch workload: `count`
    XMSS-SHA2_10_256 keys (leaves uniform in [0, 1022], messages log-uniform
    32 B - 4 KB), each signed and then verified by its own xmss_sign /
    xmss_sign_open call through the plain-core library.  Returns signs/s,
    verifies/s and p50/p99 per-call latency in ms."""
    L = api.lib
    rng = random.Random(SEED + 17)
    seeds, idxs, msgs = batch_inputs(count, rng)
    idxs = [min(i, 1022) for i in idxs]
    p = api.params(0x01, 0)
    pks = [buf(4 + p.pk_bytes) for _ in range(count)]
    sks = [buf(4 + p.sk_bytes) for _ in range(count)]
    assert L.xmss_batch_keypair(0, ptrs(pks), ptrs(sks), seeds, 0x01, count) == 0
    slat, vlat, sms = [], [], []
    for i in range(count):
        sks[i][4:8] = list(idxs[i].to_bytes(4, "big"))
        t0 = time.perf_counter()
        rc, sm, _ = api.sign(raw(sks[i], 4 + p.sk_bytes), msgs[i], 0)
        slat.append(time.perf_counter() - t0)
        assert rc == 0
        sms.append(sm)
    for i in range(count):
        t0 = time.perf_counter()
        vr, mm = api.sign_open(sms[i], raw(pks[i], 4 + p.pk_bytes), 0)
        vlat.append(time.perf_counter() - t0)
        assert vr == 0 and mm == msgs[i]
    q = lambda v, f: sorted(v)[min(len(v) - 1, int(math.ceil(f * len(v))) - 1)]
    return {"keys": count, "signs_per_s": round(count / sum(slat), 1),
            "verifies_per_s": round(count / sum(vlat), 1),
            "sign_p50_ms": round(1000 * q(slat, 0.5), 3), "sign_p99_ms": round(1000 * q(slat, 0.99), 3),
            "verify_p50_ms": round(1000 * q(vlat, 0.5), 3), "verify_p99_ms": round(1000 * q(vlat, 0.99), 3)}


def bench_opk(api, count):
    """One XMSS-SHA2_10_256 key signing `count` messages (log-uniform 32 B -
    4 KB) at consecutive indices from 0 with xmss_mi355x_sign_many (one tree
    pass shared by all), against the same key signing one xmss_sign call per
    message (timed on the first 32).  Returns signs/s of both."""
    L = api.lib
    L.xmss_mi355x_sign_many.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_size_t, C.c_void_p]
    rng = random.Random(SEED + 19)
    p = api.params(0x01, 0)
    rc, pk, sk = api.seed_keypair(0x01, 0, bytes(rng.getrandbits(8) for _ in range(96)))
    assert rc == 0
    oid = (1).to_bytes(4, "big")
    pk, sk = oid + pk, oid + sk
    msgs = [bytes(rng.getrandbits(8) for _ in range(int(32 * 128 ** rng.random()))) for _ in range(count)]
    sms = [buf(p.sig_bytes + len(m)) for m in msgs]
    mbs = [buf(m) for m in msgs]
    smp = (C.c_void_p * count)(*[C.cast(x, C.c_void_p) for x in sms])
    mp = (C.c_void_p * count)(*[C.cast(x, C.c_void_p) for x in mbs])
    mlens = (C.c_ulonglong * count)(*[len(m) for m in msgs])
    smlens = (C.c_ulonglong * count)()
    rets = (C.c_int * count)()
    # warm-up on a copy of the key (device buffers grow on first use)
    assert L.xmss_mi355x_sign_many(0, buf(sk), smp, smlens, mp, mlens, count, rets) == 0
    skb = buf(sk)
    t0 = time.perf_counter()
    assert L.xmss_mi355x_sign_many(0, skb, smp, smlens, mp, mlens, count, rets) == 0
    t_many = time.perf_counter() - t0
    assert all(rets[i] == 0 for i in range(count))
    vr, mm = api.sign_open(raw(sms[count - 1], smlens[count - 1]), pk, 0)
    assert vr == 0 and mm == msgs[count - 1]
    k = min(32, count)
    t0 = time.perf_counter()
    s = sk
    for m in msgs[:k]:
        rc, _, s = api.sign(s, m, 0)
        assert rc == 0
    t_one = time.perf_counter() - t0
    return {"signs": count, "signs_per_s": round(count / t_many, 1), "seconds": round(t_many, 6),
            "one_call_each_signs_per_s": round(k / t_one, 1)}


MIXED = [("XMSSMT-SHA2_20/2_256", 1, 0x01), ("XMSS-SHA2_10_512", 0, 0x04), ("XMSS-SHAKE_10_256", 0, 0x07)]


def warm_run(L, job, runs=5):
    """Median seconds of a staged verification job's runs after two untimed
    ones (the first sizes its buffers, the second captures its HIP graph)."""
    L.xmss_mi355x_job_run(job)
    L.xmss_mi355x_job_run(job)
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        L.xmss_mi355x_job_run(job)
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[runs // 2]


def bench_mixed(api, per_set):
    """The mixed-parameter batch: per_set keys of each of XMSSMT-SHA2_20/2_256,
    XMSS-SHA2_10_512 and XMSS-SHAKE_10_256, one plain-core signature each at
    a uniform leaf with a log-uniform 32 B - 4 KB message, then verification.
    Timed twice: one batch per parameter set back to back (per-set rates), and
    all 3 x per_set requests as one mixed job (each set on its own engine,
    sharing the GPU).  Signing is timed over one run of each job,
    verification as the median of warm runs (graph replays)."""
    L = api.lib
    L.xmss_mi355x_mixed_sign_prepare.restype = C.c_void_p
    L.xmss_mi355x_mixed_sign_prepare.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    L.xmss_mi355x_mixed_verify_prepare.restype = C.c_void_p
    L.xmss_mi355x_mixed_verify_prepare.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    rng = random.Random(SEED + 11)
    res = {}
    total_sign = total_ver = 0.0
    allk = []  # (mt, p, pk bytes, sk bytes, msg)
    for name, mt, oid in MIXED:
        p = api.params(oid, mt)
        seeds = bytes(rng.getrandbits(8) for _ in range(3 * p.n * per_set))
        pks = [buf(4 + p.pk_bytes) for _ in range(per_set)]
        sks = [buf(4 + p.sk_bytes) for _ in range(per_set)]
        assert L.xmss_batch_keypair(mt, ptrs(pks), ptrs(sks), seeds, oid, per_set) == 0
        idxs = [rng.randrange(0, (1 << p.full_height) - 1) for _ in range(per_set)]
        for i in range(per_set):
            sks[i][4:4 + p.index_bytes] = list(idxs[i].to_bytes(p.index_bytes, "big"))
        msgs = [bytes(rng.getrandbits(8) for _ in range(int(32 * 128 ** rng.random()))) for _ in range(per_set)]
        for i in range(per_set):
            allk.append((mt, p, raw(pks[i], 4 + p.pk_bytes), raw(sks[i], 4 + p.sk_bytes), msgs[i]))
        ms = [buf(m) for m in msgs]
        mlens = (C.c_ulonglong * per_set)(*[len(m) for m in msgs])
        sms = [buf(p.sig_bytes + len(m)) for m in msgs]
        smlens = (C.c_ulonglong * per_set)()
        rets = (C.c_int * per_set)()
        job = L.xmss_mi355x_sign_prepare(mt, ptrs(sks), ptrs(ms), mlens, per_set)
        t0 = time.perf_counter()
        L.xmss_mi355x_job_run(job)
        ts = time.perf_counter() - t0
        L.xmss_mi355x_sign_collect(job, ptrs(sms), smlens, rets)
        L.xmss_mi355x_job_free(job)
        assert all(rets[i] == 0 for i in range(per_set))
        smb = [buf(raw(sms[i], smlens[i])) for i in range(per_set)]
        outs = [buf(smlens[i]) for i in range(per_set)]
        mls = (C.c_ulonglong * per_set)()
        vr = (C.c_int * per_set)()
        job = L.xmss_mi355x_verify_prepare(mt, ptrs(smb), smlens, ptrs(pks), per_set)
        tv = warm_run(L, job)
        L.xmss_mi355x_verify_collect(job, ptrs(outs), mls, vr)
        L.xmss_mi355x_job_free(job)
        assert all(vr[i] == 0 for i in range(per_set))
        res[name] = {"signs_per_s": round(per_set / ts, 1), "verifies_per_s": round(per_set / tv, 1)}
        total_sign += ts
        total_ver += tv
        log("  %s: sign %.3f s, verify %.4f s" % (name, ts, tv))
    res["total_sign_s"] = round(total_sign, 4)
    res["total_verify_s"] = round(total_ver, 5)
    # the same requests, shuffled, as one mixed job
    rng.shuffle(allk)
    k = len(allk)
    mts = (C.c_int * k)(*[x[0] for x in allk])
    sks = [buf(x[3]) for x in allk]
    ms = [buf(x[4]) for x in allk]
    mlens = (C.c_ulonglong * k)(*[len(x[4]) for x in allk])
    sms = [buf(x[1].sig_bytes + len(x[4])) for x in allk]
    smlens = (C.c_ulonglong * k)()
    rets = (C.c_int * k)()
    job = L.xmss_mi355x_mixed_sign_prepare(mts, ptrs(sks), ptrs(ms), mlens, k)
    t0 = time.perf_counter()
    L.xmss_mi355x_job_run(job)
    ts = time.perf_counter() - t0
    L.xmss_mi355x_sign_collect(job, ptrs(sms), smlens, rets)
    L.xmss_mi355x_job_free(job)
    assert all(rets[i] == 0 for i in range(k))
    pks = [buf(x[2]) for x in allk]
    smb = [buf(raw(sms[i], smlens[i])) for i in range(k)]
    outs = [buf(smlens[i]) for i in range(k)]
    mls = (C.c_ulonglong * k)()
    vr = (C.c_int * k)()
    job = L.xmss_mi355x_mixed_verify_prepare(mts, ptrs(smb), smlens, ptrs(pks), k)
    tv = warm_run(L, job)
    L.xmss_mi355x_verify_collect(job, ptrs(outs), mls, vr)
    L.xmss_mi355x_job_free(job)
    assert all(vr[i] == 0 for i in range(k))
    res["one_job"] = {"requests": k, "sign_s": round(ts, 4), "verify_s": round(tv, 5),
                      "signs_per_s": round(k / ts, 1), "verifies_per_s": round(k / tv, 1)}
    log("  one mixed job: sign %.3f s, verify %.4f s" % (ts, tv))
    return res


SERVE_SETS = [("XMSS-SHA2_16_256", 0, 0x02), ("XMSSMT-SHA2_20/4_256", 1, 0x02), ("XMSSMT-SHA2_40/4_256", 1, 0x04)]


def bench_serve(rate, seconds, nkeys, budget):
    """Serving: Poisson arrivals at `rate` requests/s for `seconds`, 70% sign /
    30% verify, over nkeys BDS keys split evenly among XMSS-SHA2_16_256,
    XMSSMT-SHA2_20/4_256 and XMSSMT-SHA2_40/4_256; messages log-uniform
    32 B - 4 KB.  Latency is submission to completion, measured in the library."""
    import queue
    import threading
    api = load_product(1)
    L = api.lib
    L.xmss_mi355x_server_create.restype = C.c_void_p
    L.xmss_mi355x_server_create.argtypes = [C.c_size_t, C.c_int]
    L.xmss_mi355x_server_add_key.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    for nm in ("xmss_mi355x_server_sign", "xmss_mi355x_server_verify"):
        getattr(L, nm).restype = C.c_uint64
        getattr(L, nm).argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_ulonglong]
    L.xmss_mi355x_server_result.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.POINTER(C.c_ulonglong),
                                            C.POINTER(C.c_double)]
    L.xmss_mi355x_server_stats.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                           C.POINTER(C.c_size_t), C.POINTER(C.c_uint64)]
    L.xmss_mi355x_server_destroy.argtypes = [C.c_void_p]
    rng = random.Random(SEED + 13)
    t0 = time.perf_counter()
    keys = []
    made = {}
    for i in range(nkeys):
        name, mt, oid = SERVE_SETS[i % 3]
        p = api.params(oid, mt)
        if mt == 0 and name in made:  # the XMSS BDS keypair uses fixed seeds: every such key is the same
            pk, sk = made[name]
        else:
            L.xmss_mi355x_rng_seed(SEED + 100 + i)
            rc, pk, sk = api.keypair(oid, mt, sk_init=bytes(4 + p.sk_bytes))
            assert rc == 0
            made[name] = (pk, sk)
        keys.append((mt, p, pk, sk))
    keygen_s = time.perf_counter() - t0
    srv = L.xmss_mi355x_server_create(budget, 256)
    for mt, p, pk, sk in keys:
        L.xmss_mi355x_server_add_key(srv, mt, buf(sk), buf(pk))
    pending = queue.Queue()
    done_sigs = []  # (key, sm) available for verification
    lat = {"s": [], "v": []}
    by_set = {}  # (kind, parameter set) -> latencies
    fails = [0]

    def collect():
        while True:
            item = pending.get()
            if item is None:
                return
            kind, k, rid = item
            p = keys[k][1]
            out = buf(p.sig_bytes + 4200)
            olen, ms = C.c_ulonglong(0), C.c_double(0)
            rc = L.xmss_mi355x_server_result(srv, rid, out, C.byref(olen), C.byref(ms))
            if rc != 0:
                fails[0] += 1
            lat[kind].append(ms.value)
            by_set.setdefault((kind, SERVE_SETS[k % 3][0]), []).append(ms.value)
            if kind == "s" and rc == 0:
                done_sigs.append((k, raw(out, olen.value)))

    th = threading.Thread(target=collect)
    th.start()
    n = 0
    start = time.perf_counter()
    t_next = start
    while True:
        t_next += rng.expovariate(rate)
        if t_next - start > seconds:
            break
        now = time.perf_counter()
        if t_next > now:
            time.sleep(t_next - now)
        if rng.random() < 0.7 or not done_sigs:
            k = rng.randrange(nkeys)
            m = bytes(rng.getrandbits(8) for _ in range(int(32 * 128 ** rng.random())))
            pending.put(("s", k, L.xmss_mi355x_server_sign(srv, k, m, len(m))))
        else:
            k, sm = done_sigs[rng.randrange(len(done_sigs))]
            pending.put(("v", k, L.xmss_mi355x_server_verify(srv, k, sm, len(sm))))
        n += 1
    pending.put(None)
    th.join()
    elapsed = time.perf_counter() - start
    launches, calls, peak, waits = C.c_uint64(), C.c_uint64(), C.c_size_t(), C.c_uint64()
    L.xmss_mi355x_server_stats(srv, C.byref(launches), C.byref(calls), C.byref(peak), C.byref(waits))
    L.xmss_mi355x_server_destroy(srv)
    q = lambda v, f: sorted(v)[min(len(v) - 1, int(math.ceil(f * len(v))) - 1)] if v else None
    return {"offered_per_s": rate, "seconds": seconds, "keys": nkeys, "requests": n,
            "sustained_ops_per_s": round(n / elapsed, 1), "failures": fails[0],
            "sign_p50_ms": round(q(lat["s"], 0.5), 3), "sign_p99_ms": round(q(lat["s"], 0.99), 3),
            "verify_p50_ms": round(q(lat["v"], 0.5), 3) if lat["v"] else None,
            "verify_p99_ms": round(q(lat["v"], 0.99), 3) if lat["v"] else None,
            "by_set_p50_p99_ms": {"%s %s" % ("sign" if kd == "s" else "verify", nm): [round(q(v, 0.5), 3),
                                                                                   round(q(v, 0.99), 3)]
                                  for (kd, nm), v in sorted(by_set.items())},
            "hbm_budget_bytes": budget, "peak_reserved_bytes": peak.value, "admission_waits": waits.value,
            "device_calls": calls.value, "launch_groups": launches.value, "key_setup_s": round(keygen_s, 2)}


def bench_latency(nsigns, fixed_len=None):
    """Sequential BDS signatures of one XMSS-SHA2_16_256 key (fast core),
    messages log-uniform 32 B - 4 KB (or all fixed_len bytes)."""
    api = load_product(1)
    p = api.params(0x02, 0)
    rc, pk, sk = api.keypair(0x02, 0, sk_init=bytes(4 + p.sk_bytes))
    assert rc == 0
    rng = random.Random(SEED + 9)
    lat = []
    last = None
    # the interpreter's cyclic collector stays out of the timed calls (the
    # heap of the earlier sections would otherwise add its pauses to the tail)
    gc.collect()
    gc.disable()
    try:
        for _ in range(nsigns):
            mlen = fixed_len if fixed_len else int(32 * 128 ** rng.random())
            m = bytes(rng.getrandbits(8) for _ in range(mlen))
            t0 = time.perf_counter()
            rc, sm, sk = api.sign(sk, m, 0)
            lat.append(time.perf_counter() - t0)
            assert rc == 0
            last = (sm, m)
    finally:
        gc.enable()
    vlat = []
    for _ in range(16):
        t0 = time.perf_counter()
        vr, mm = api.sign_open(last[0], pk, 0)
        vlat.append(time.perf_counter() - t0)
        assert vr == 0 and mm == last[1]
    lat.sort()
    vlat.sort()
    q = lambda v, f: v[min(len(v) - 1, int(math.ceil(f * len(v))) - 1)]
    return q(lat, 0.5), q(lat, 0.99), q(vlat, 0.5)


def cpu_baseline():
    """The reference C code (oracle/_ref, xmss_core.c keygen) on one host core:
    XMSS-SHA2_10_256 keygens, scaled by 64.0 = 298,582,010 / 4,665,338
    compressions to an XMSS-SHA2_16_256 keygen."""
    ref = load_ref(0)
    kind = "reference"
    if ref is None:
        kind = "port"
        orc = Oracle(ORACLE_SO)
        run = lambda: orc.seed_keypair(0x01, 0, bytes(range(96)))
    else:
        run = lambda: ref.seed_keypair(0x01, 0, bytes(range(96)))
    n, t0 = 0, time.perf_counter()
    while n < 2 or time.perf_counter() - t0 < 12.0:  # about 10 s of CPU work
        run()
        n += 1
        if n >= 10:
            break
    dt = (time.perf_counter() - t0) / n
    scale = 298582010 / 4665338
    return {"value": round(dt * scale, 3), "unit": "s per XMSS-SHA2_16_256 keygen", "cores": 1, "kind": kind,
            "sample": "%d XMSS-SHA2_10_256 keygens (%.2f s each) on 1 core, x%.2f compressions" % (n, dt, scale)}


KEYGEN_SEED = bytes(range(96))  # SK_SEED || SK_PRF || PUB_SEED of the keygen step
CPU_PARTS = 16                  # the keygen's 65,536 leaves in 16 subtrees of 4,096


def _ref_hash_fns(ref):
    from ffi import Params, u8p
    P = C.POINTER(Params)
    ref.lib.gen_leaf_wots.argtypes = [P, u8p, u8p, u8p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    ref.lib.thash_h.argtypes = [P, u8p, u8p, u8p, C.POINTER(C.c_uint32)]


def cpu_keygen_worker(part):
    """Root of leaves [part * 4096, (part + 1) * 4096) of the keygen step's
    XMSS-SHA2_16_256 tree with the reference's own gen_leaf_wots and thash_h
    (oracle/_ref, OpenSSL SHA-256), merged as treehash does
    (xmss_core.c:19-86); prints the root and the seconds taken."""
    from ffi import addr
    ref = load_ref(0)
    _ref_hash_fns(ref)
    p = ref.params(0x02, 0)
    n = p.n
    sk_seed, pub_seed = buf(KEYGEN_SEED[:n]), buf(KEYGEN_SEED[2 * n:3 * n])
    height = p.tree_height - int(math.log2(CPU_PARTS))
    leaf0 = part << height
    t0 = time.perf_counter()
    stack = []  # (height, node)
    for idx in range(leaf0, leaf0 + (1 << height)):
        leaf = buf(n)
        ref.lib.gen_leaf_wots(C.byref(p), leaf, sk_seed, pub_seed, addr([0, 0, 0, 1, idx]), addr([0, 0, 0, 0, idx]))
        node, h = raw(leaf, n), 0
        while stack and stack[-1][0] == h:
            out = buf(n)
            ref.lib.thash_h(C.byref(p), out, buf(stack.pop()[1] + node), pub_seed, addr([0, 0, 0, 2, 0, h, idx >> (h + 1)]))
            node, h = raw(out, n), h + 1
        stack.append((h, node))
    print(stack[0][1].hex(), "%.6f" % (time.perf_counter() - t0), flush=True)


def cpu_keygen(pk_gpu):
    """One whole XMSS-SHA2_16_256 keygen on the host: 16 child processes each
    build a 4,096-leaf subtree with the reference's hashing code, the parent
    hashes the top four levels.  Its root must equal the GPU keygen's."""
    import subprocess
    from ffi import addr
    ref = load_ref(0)
    if ref is None:
        return None
    _ref_hash_fns(ref)
    p = ref.params(0x02, 0)
    n = p.n
    t0 = time.perf_counter()
    procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-keygen-worker", str(k)],
                              stdout=subprocess.PIPE, text=True) for k in range(CPU_PARTS)]
    outs = [pr.communicate()[0].split() for pr in procs]
    level = [bytes.fromhex(o[0]) for o in outs]
    h = p.tree_height - int(math.log2(CPU_PARTS))
    pub_seed = buf(KEYGEN_SEED[2 * n:3 * n])
    while len(level) > 1:
        nxt = []
        for i in range(len(level) // 2):
            out = buf(n)
            ref.lib.thash_h(C.byref(p), out, buf(level[2 * i] + level[2 * i + 1]), pub_seed,
                            addr([0, 0, 0, 2, 0, h, i]))
            nxt.append(raw(out, n))
        level, h = nxt, h + 1
    wall = time.perf_counter() - t0
    assert level[0] == pk_gpu[:n], "CPU keygen root differs from the GPU's"
    each = [float(o[1]) for o in outs]
    return {"value": round(wall, 3), "unit": "s per XMSS-SHA2_16_256 keygen", "cores": CPU_PARTS,
            "kind": "reference",
            "sample": "one whole XMSS-SHA2_16_256 keygen: %d processes x 4,096 leaves with the reference's "
                      "gen_leaf_wots / thash_h (%.2f s each on average), top 4 levels in the parent; root equal "
                      "to the GPU's" % (CPU_PARTS, sum(each) / len(each))}


def cpu_sign_worker():
    """One plain-core XMSS-SHA2_10_256 signature by the reference C code
    (oracle/_ref) at a random leaf, in a process of its own: prints seconds.
    The sk's root field is arbitrary bytes: signing time does not depend on
    it (the signature rebuilds the whole tree either way)."""
    ref = load_ref(0)
    p = ref.params(0x01, 0)
    rng = random.Random(os.getpid())
    sk = (1).to_bytes(4, "big") + rng.randrange(1023).to_bytes(4, "big") + \
        bytes(rng.getrandbits(8) for _ in range(4 * p.n))
    m = bytes(rng.getrandbits(8) for _ in range(int(32 * 128 ** rng.random())))
    t0 = time.perf_counter()
    rc, _, _ = ref.sign(sk, m, 0)
    assert rc == 0
    print("%.6f" % (time.perf_counter() - t0), flush=True)


def cpu_batch_baseline(cores):
    """The batch-sign workload (XMSS-SHA2_10_256 plain signatures of
    independent keys) on the reference C code across `cores` host processes,
    one signature each, started as child processes (no GPU use in them)."""
    import subprocess
    if load_ref(0) is None:
        return None
    t0 = time.perf_counter()
    procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-sign-worker"],
                              stdout=subprocess.PIPE, text=True) for _ in range(cores)]
    each = [float(pr.communicate()[0].strip().splitlines()[-1]) for pr in procs]
    wall = time.perf_counter() - t0
    return {"value": round(cores / wall, 3), "unit": "XMSS-SHA2_10_256 plain signs/s", "cores": cores,
            "kind": "reference",
            "sample": "%d signatures, one per process (%.2f s each on average), %.2f s wall" % (
                cores, sum(each) / len(each), wall)}


def copy_calibration(nbytes):
    """One plain device-to-device copy of nbytes (reads nbytes, writes nbytes):
    the reference point for HBM byte counters."""
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpyDtoD(b, a, C.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log("copy calibration: %d bytes device-to-device" % nbytes)


def profiled(field, kernel="k_chains"):
    """A per-launch figure of the keygen chain kernel from the committed
    rocprofv3 PMC summary: `traffic_bytes_per_launch` (FETCH_SIZE + WRITE_SIZE,
    calibrated on a copy kernel) or `valu_lane_ops_per_launch` (SQ_INSTS_VALU
    x 64)."""
    path = os.path.join(ROOT, "profiles", "keygen_xmss_sha2_16_256_pmc.md")
    try:
        for line in open(path):
            if line.startswith("%s %s:" % (field, kernel)):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


VALU_CLASSES = {1: "v_alignbit_b32", 2: "v_add3_u32", 3: "v_bitop3_b32", 4: "v_add_u32/v_xor_b32",
                5: "v_perm_b32/v_bfi_b32"}


def valu_peak(api, kind=0, runs=3):
    """Measured int32 VALU rate (T lane-ops/s): best of `runs` launches of the
    library's probe kernel (4096 workgroups x 256 lanes x 4000 x 96
    instructions).  Kind 0 issues the instruction mix of the SHA-256 chain
    step (per 12: 5 v_alignbit, 2 v_add3, 3 v_bitop3, 1 v_add, 1 v_lshrrev);
    kinds 1-5 one instruction class each (VALU_CLASSES)."""
    L = api.lib
    L.xmss_mi355x_valu_probe.argtypes = [C.c_int, C.c_uint, C.c_uint, C.POINTER(C.c_double),
                                         C.POINTER(C.c_double)]
    best = 0.0
    for _ in range(runs):
        rate, ms = C.c_double(0), C.c_double(0)
        if L.xmss_mi355x_valu_probe(kind, 4096, 4000, C.byref(rate), C.byref(ms)) != 0:
            return None
        best = max(best, rate.value)
    return best / 1e12


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--latency-signs", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sign-worker", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-keygen-worker", type=int, default=-1, help=argparse.SUPPRESS)
    ap.add_argument("--opk", type=int, default=1023,
                    help="consecutive signatures of one XMSS-SHA2_10_256 key through sign_many (0: skip)")
    ap.add_argument("--single", type=int, default=256,
                    help="keys signed one xmss_sign call each, beside the batch (0: skip)")
    ap.add_argument("--stream-batches", type=int, default=16,
                    help="batches of the PCIe-inclusive verification stream (0: skip)")
    ap.add_argument("--mixed", type=int, default=1024, help="keys per parameter set of the mixed batch (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=4000.0, help="offered requests/s of the serving run (0: skip)")
    ap.add_argument("--serve-seconds", type=float, default=5.0)
    ap.add_argument("--serve-budget", type=int, default=64 << 20, help="HBM budget of the serving run, bytes")
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="MiB of one device-to-device copy, run before the keygens (HBM counter calibration)")
    args = ap.parse_args()
    if args.cpu_sign_worker:
        cpu_sign_worker()
        return
    if args.cpu_keygen_worker >= 0:
        cpu_keygen_worker(args.cpu_keygen_worker)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: keys are independent and shard by key id with no "
              "communication, so no multi-GPU run is defined here (use --gpus 1).", file=sys.stderr)
        sys.exit(2)

    api = load_product(0)
    if api.lib.xmss_mi355x_device_ok() != 1:
        raise SystemExit("no usable GPU")
    if args.copy_calib:
        copy_calibration(args.copy_calib << 20)
    log("keygen XMSS-SHA2_16_256 ...")
    times, kst, pk = bench_keygen(api, args.steps, args.warmup)
    assert len(pk) == 64
    ms_per_step = 1000 * sum(times) / len(times)
    chains = kst.get("chains")
    leaves = 1 << 16
    comps_chains = leaves * 67 * 62  # PRF_keygen (2) + 15 thash_f (4 each) per chain
    roof = None
    if chains:
        avg_ms = chains[1] / chains[0]
        peak_meas = valu_peak(api)
        # achieved: the fixed nominal work of one launch (62 SHA-256
        # compressions of 1384 VALU ops per (leaf, chain) item) over its
        # duration; the ceiling is a fixed figure, so that a cut in executed
        # instructions shows as a higher fraction.  Beside it: the lane-ops the
        # kernel executes (static count = SQ_INSTS_VALU x 64 in the profile;
        # 11% fewer than nominal: PRF rounds shared along each chain, constant
        # rounds and schedule words folded) and the probe's rate for the
        # kernel's own instruction mix.
        ops = profiled("valu_lane_ops_per_launch")
        nominal = comps_chains * OPS_PER_SHA256
        rate = lambda o: o / (avg_ms / 1000) / 1e12
        achieved = rate(nominal)
        roof = {"bound": "valu", "achieved": round(achieved, 3), "peak": round(VALU_CEILING_TOPS, 3),
                "unit": "TOP/s", "frac": round(achieved / VALU_CEILING_TOPS, 4),
                "traffic": profiled("traffic_bytes_per_launch"),
                "kernel": "k_chains<Sha256Fam<32>>", "kernel_ms": round(avg_ms, 4),
                "ops_per_launch": nominal,
                "ops_counted": "nominal: %d compressions x %d VALU ops" % (comps_chains, OPS_PER_SHA256),
                "peak_source": "2 x 256 CU x 64 lanes x 2.4 GHz (the lane rate of the 157.3 TFLOP/s FP32 peak)",
                "executed_ops_per_launch": ops,
                "executed_ops_per_compression": round(ops / comps_chains, 1) if ops else None,
                "executed_rate": round(rate(ops), 3) if ops else None,
                "mix_peak_measured": round(peak_meas, 3) if peak_meas else None,
                "frac_of_mix_peak": round(rate(ops) / peak_meas, 4) if ops and peak_meas else None,
                "single_issue_peak": round(VALU_PEAK_TOPS, 2),
                "class_rates": {nm: round(valu_peak(api, k, runs=1) or 0, 2) for k, nm in VALU_CLASSES.items()}}
    total_comps = 298582010
    keygen_frac = total_comps * OPS_PER_SHA256 / (ms_per_step / 1000) / 1e12 / VALU_CEILING_TOPS

    bkeygen_t = sign_run = sign_total = ver_t = ver_refill_t = ver_stream_t = None
    bst, n_last, bnom = {}, None, None
    if args.batch:
        log("batch keygen/sign/verify x%d ..." % args.batch)
        (bkeygen_t, sign_run, sign_total, (ver_t, ver_refill_t, ver_stream_t), bst, n_last,
         bnom) = bench_batch(api, args.batch, args.stream_batches)
    rate = lambda t: None if not t else round(args.batch / t, 2)
    single = None
    if args.single:
        log("plain one-request-per-launch x%d ..." % args.single)
        single = bench_plain_single(api, args.single)
    opk = None
    if args.opk:
        log("one key, %d consecutive signatures ..." % args.opk)
        opk = bench_opk(api, args.opk)
    mixed = None
    if args.mixed:
        log("mixed batch x%d per set ..." % args.mixed)
        mixed = bench_mixed(api, args.mixed)
    serve = None
    if args.serve_rate:
        log("serving at %.0f req/s for %.0f s ..." % (args.serve_rate, args.serve_seconds))
        serve = bench_serve(args.serve_rate, args.serve_seconds, 256, args.serve_budget)
    log("latency (BDS) x%d ..." % args.latency_signs)
    p50, p99, v50 = bench_latency(args.latency_signs) if args.latency_signs else (None, None, None)
    p50s, p99s, _ = bench_latency(args.latency_signs, 32) if args.latency_signs else (None, None, None)
    cpu = None if args.no_cpu else cpu_keygen(pk)
    cpu_1core = None if args.no_cpu else cpu_baseline()
    cpu_batch = None if args.no_cpu else cpu_batch_baseline(min(16, os.cpu_count() or 1))

    out = {
        "metric": METRIC,
        "value": round(ms_per_step / 1000, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded keys and messages)",
        "config": {"workload": "XMSS-SHA2_16_256 keygen (65,536 leaves); batch 4096 x XMSS-SHA2_10_256 "
                               "plain sign + verify; 1024 sequential BDS signs (messages log-uniform 32 B - 4 KB)",
                   "parallelism": "single GPU"},
        "keygen_nominal_frac": round(keygen_frac, 4),
        "keygens_per_s_batch4096": rate(bkeygen_t),
        "signs_per_s_batch4096": rate(sign_run),
        "signs_per_s_batch4096_incl_pcie": rate(sign_total),
        "verifies_per_s_batch4096": rate(ver_t),
        "verifies_per_s_batch4096_refill_incl_pcie": rate(ver_refill_t),
        "verifies_per_s_batch4096_stream_incl_pcie": rate(ver_stream_t),
        "p50_sign_ms": None if p50 is None else round(1000 * p50, 3),
        "p99_sign_ms": None if p99 is None else round(1000 * p99, 3),
        "p50_verify_ms": None if v50 is None else round(1000 * v50, 3),
        "bds_sign_32B_messages_ms": None if p50s is None else {"p50": round(1000 * p50s, 3),
                                                                "p99": round(1000 * p99s, 3)},
        "roofline": roof,
        "batch4096_nominal_frac": bnom,
        "cpu_baseline": cpu,
        "cpu_baseline_1core": cpu_1core,
        "cpu_baseline_batch_sign": cpu_batch,
        "kernels_keygen_ms": {k: round(v[1], 4) for k, v in kst.items()},
        "kernels_batch_sign_ms": {k: round(v[1], 3) for k, v in bst.items()},
        "batch_last_index_rejected": n_last,
        "plain_one_per_call": single,
        "one_key_many_signs": opk,
        "mixed_batch": mixed,
        "serving": serve,
        "verify_tuned_plans": tuned_plans(api.lib),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
