"""The verification-batch schedules the autotuner chooses among, each against
the reference's outputs: the three mixed-workload sets (XMSSMT-SHA2_20/2_256,
XMSS-SHA2_10_512, XMSS-SHAKE_10_256; 260 plain-core signatures each, signed
here and checked against tests/golden/mixed_sets.json, written by the
reference's xmss_core.c) are verified as one batch per set with every
candidate plan forced -- direct launches, graph capture and replay -- with a
flipped byte in R, a WOTS chain value and an auth node of three requests.
Then one shape is left to the tuner: every run, tuning runs included, must
give the same verdicts, and the settled plan is one of the candidates."""
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


def _lib():
    api = product(0)
    L = api.lib
    L.xmss_mi355x_sign_prepare.restype = C.c_void_p
    L.xmss_mi355x_sign_prepare.argtypes = [C.c_int] + [C.c_void_p] * 3 + [C.c_size_t]
    L.xmss_mi355x_verify_prepare.restype = C.c_void_p
    L.xmss_mi355x_verify_prepare.argtypes = [C.c_int] + [C.c_void_p] * 3 + [C.c_size_t]
    L.xmss_mi355x_job_run.argtypes = [C.c_void_p]
    L.xmss_mi355x_sign_collect.argtypes = [C.c_void_p] * 4
    L.xmss_mi355x_verify_collect.argtypes = [C.c_void_p] * 4
    L.xmss_mi355x_job_free.argtypes = [C.c_void_p]
    L.xmss_mi355x_set_verify_plan.argtypes = [C.c_int]
    L.xmss_mi355x_verify_plans.argtypes = [C.c_void_p, C.c_size_t]
    L.xmss_mi355x_verify_plans.restype = C.c_size_t
    L.xmss_mi355x_tuned_plans.argtypes = [C.c_char_p, C.c_size_t]
    L.xmss_mi355x_tuned_plans.restype = C.c_size_t
    return api


@pytest.fixture(scope="module")
def signed():
    """{set name: (mt, p, [sm], [pk], [msg])}, every signature equal to the reference's."""
    api = _lib()
    L = api.lib
    out = {}
    for name, s in golden("mixed_sets.json").items():
        oid, mt = s["oid"], s["mt"]
        p = api.params(oid, mt)
        seeds = [W.key_seed(s["key_seed"], k, p.n) for k in range(s["keys"])]
        roots = [bytes.fromhex(r) for r in s["roots"]]
        rows = s["items"]
        k = len(rows)
        sks = [buf(W.plain_sk(oid, p, r["idx"], seeds[i % s["keys"]], roots[i % s["keys"]]))
               for i, r in enumerate(rows)]
        msgs = [W.message(s["msg_seed"], i) for i in range(k)]
        mb = [buf(m) for m in msgs]
        mlens = (C.c_ulonglong * k)(*[len(m) for m in msgs])
        job = L.xmss_mi355x_sign_prepare(mt, _arr(sks), _arr(mb), mlens, k)
        assert L.xmss_mi355x_job_run(job) == 0
        sms = [buf(p.sig_bytes + len(m)) for m in msgs]
        smlens = (C.c_ulonglong * k)()
        rets = (C.c_int * k)()
        assert L.xmss_mi355x_sign_collect(job, _arr(sms), smlens, rets) == 0
        L.xmss_mi355x_job_free(job)
        sm = []
        for i, r in enumerate(rows):
            assert rets[i] == r["rc"] == 0
            x = raw(sms[i], smlens[i])
            assert dg(x) == r["sm"], "%s item %d signature" % (name, i)
            sm.append(x)
        pks = [W.with_oid(oid, roots[i % s["keys"]] + seeds[i % s["keys"]][2 * p.n:]) for i in range(k)]
        out[name] = (mt, p, sm, pks, msgs)
    return out


def _corrupt(p, sms):
    """Flips a byte of R, of a WOTS chain value and of an auth node in three
    requests; returns the new list and the flipped indices."""
    sms = list(sms)
    rng = random.Random(11)
    picks = rng.sample(range(len(sms)), 3)
    offs = [p.index_bytes + 3, p.index_bytes + p.n + 5 * p.n + 1,
            p.index_bytes + p.n + p.wots_sig_bytes + 2]
    for i, off in zip(picks, offs):
        b = bytearray(sms[i])
        b[off] ^= 0x10
        sms[i] = bytes(b)
    return sms, set(picks)


def _verify(L, mt, sms, pks, runs):
    """Verdicts and outputs of `runs` runs of one verification job."""
    k = len(sms)
    smb = [buf(s) for s in sms]
    pkb = [buf(x) for x in pks]
    smls = (C.c_ulonglong * k)(*[len(s) for s in sms])
    job = L.xmss_mi355x_verify_prepare(mt, _arr(smb), smls, _arr(pkb), k)
    assert job
    res = []
    try:
        for _ in range(runs):
            assert L.xmss_mi355x_job_run(job) == 0
            outs = [buf(len(s)) for s in sms]
            mls = (C.c_ulonglong * k)()
            vr = (C.c_int * k)()
            assert L.xmss_mi355x_verify_collect(job, _arr(outs), mls, vr) == 0
            res.append(([vr[i] for i in range(k)], [raw(outs[i], mls[i]) for i in range(k)]))
    finally:
        L.xmss_mi355x_job_free(job)
    return res


def _check(res, msgs, flipped, what):
    for r, (vr, ms) in enumerate(res):
        for i, m in enumerate(msgs):
            if i in flipped:
                assert vr[i] == -1 and ms[i] == b"", "%s run %d: corrupted request %d accepted" % (what, r, i)
            else:
                assert vr[i] == 0 and ms[i] == m, "%s run %d: request %d rejected" % (what, r, i)


def test_every_candidate_plan_verifies_like_the_reference(signed):
    L = _lib().lib
    n = L.xmss_mi355x_verify_plans(None, 0)
    plans = (C.c_uint * n)()
    assert L.xmss_mi355x_verify_plans(plans, n) == n and n >= 2
    try:
        for plan in list(plans):
            L.xmss_mi355x_set_verify_plan(plan)
            for name, (mt, p, sms, pks, msgs) in signed.items():
                bad, flipped = _corrupt(p, sms)
                print("plan %d, %s" % (plan, name), flush=True)
                _check(_verify(L, mt, bad, pks, 3), msgs, flipped, "plan %d %s" % (plan, name))
    finally:
        L.xmss_mi355x_set_verify_plan(-1)


def test_tuned_plan_gives_the_same_verdicts(signed):
    L = _lib().lib
    n = L.xmss_mi355x_verify_plans(None, 0)
    plans = (C.c_uint * n)()
    L.xmss_mi355x_verify_plans(plans, n)
    L.xmss_mi355x_set_verify_plan(-1)
    mt, p, sms, pks, msgs = signed["XMSSMT-SHA2_20/2_256"]
    bad, flipped = _corrupt(p, sms)
    # runs 1 (direct), 2 (tuning: every candidate, then the kept one), 3-4 (replays)
    _check(_verify(L, mt, bad, pks, 4), msgs, flipped, "tuned")
    need = L.xmss_mi355x_tuned_plans(None, 0)
    text = C.create_string_buffer(need)
    L.xmss_mi355x_tuned_plans(text, need)
    # XMSSMT-SHA2_20/2_256: th = 10, d = 2; 260 requests -> floor(log2) = 8
    lines = [x for x in text.value.decode().splitlines() if "th=10 d=2 log2count=8 one_stream=0" in x]
    assert len(lines) == 1, text.value
    assert int(lines[0].split("plan=")[1]) in list(plans)


_CHILD = r'''
import ctypes as C, os, random, sys
sys.path.insert(0, sys.argv[1])
from common import product
from ffi import buf
L = product(0).lib
L.xmss_mi355x_tuned_plans.argtypes = [C.c_char_p, C.c_size_t]
L.xmss_mi355x_tuned_plans.restype = C.c_size_t
def plans():
    n = L.xmss_mi355x_tuned_plans(None, 0)
    b = C.create_string_buffer(n)
    L.xmss_mi355x_tuned_plans(b, n)
    return b.value.decode()
if sys.argv[2] == "run":
    # 300 well-formed but forged XMSS-SHA2_10_256 signed messages: every one
    # is rejected, and the batch still takes the lane-dense path and its tuning
    L.xmss_mi355x_verify_prepare.restype = C.c_void_p
    L.xmss_mi355x_verify_prepare.argtypes = [C.c_int] + [C.c_void_p] * 3 + [C.c_size_t]
    L.xmss_mi355x_job_run.argtypes = [C.c_void_p]
    L.xmss_mi355x_verify_collect.argtypes = [C.c_void_p] * 4
    L.xmss_mi355x_job_free.argtypes = [C.c_void_p]
    rng = random.Random(9)
    k = 300
    sms = [bytes([0, 0, 1, rng.randrange(256)]) + bytes(rng.getrandbits(8) for _ in range(2496 + 40)) for _ in range(k)]
    pks = [b"\x00\x00\x00\x01" + bytes(rng.getrandbits(8) for _ in range(64)) for _ in range(k)]
    arr = lambda xs: (C.c_void_p * k)(*[C.cast(x, C.c_void_p) for x in xs])
    smb, pkb = [buf(s) for s in sms], [buf(p) for p in pks]
    job = L.xmss_mi355x_verify_prepare(0, arr(smb), (C.c_ulonglong * k)(*map(len, sms)), arr(pkb), k)
    for _ in range(3):
        assert L.xmss_mi355x_job_run(job) == 0
    outs = [buf(len(s)) for s in sms]
    mls = (C.c_ulonglong * k)()
    vr = (C.c_int * k)()
    assert L.xmss_mi355x_verify_collect(job, arr(outs), mls, vr) == 0
    L.xmss_mi355x_job_free(job)
    assert all(v == -1 for v in vr)
print(plans(), end="")
'''


def test_tuned_plans_persist_in_the_tune_file(tmp_path):
    """A process that tunes a new shape appends its decision to
    XMSS_MI355X_TUNE_FILE; the next process reads it at start (before any job)."""
    import os
    import subprocess
    import sys
    tests = os.path.dirname(os.path.abspath(__file__))
    env = dict(os.environ, XMSS_MI355X_TUNE_FILE=str(tmp_path / "tune.txt"))
    env.pop("XMSS_MI355X_TUNE", None)
    first = subprocess.run([sys.executable, "-c", _CHILD, tests, "run"], env=env, capture_output=True, text=True,
                           timeout=240)
    assert first.returncode == 0, first.stderr[-2000:]
    line = [x for x in first.stdout.splitlines() if "fam=1 th=10 d=1 log2count=8 one_stream=0" in x]
    assert len(line) == 1, first.stdout
    assert (tmp_path / "tune.txt").read_text().startswith("verify ")
    second = subprocess.run([sys.executable, "-c", _CHILD, tests, "read"], env=env, capture_output=True, text=True,
                            timeout=240)
    assert second.returncode == 0, second.stderr[-2000:]
    assert line[0] in second.stdout.splitlines(), second.stdout
