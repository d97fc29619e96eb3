"""Batched plain-core signing and verification on the GPU (independent keys,
random leaves, 32 B - 4 KB messages) against the reference's outputs."""
import ctypes as C

import pytest

from common import golden, product
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
