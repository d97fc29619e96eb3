"""Device failures come back as return codes, never as an abort of the
caller's process: with the library's device-memory limit set below what a
call needs, that call returns -3 (a one-call signature leaves its sk as it
was), and once the limit is lifted the same calls work and equal the oracle."""
import ctypes as C

import numpy as np
import pytest

from common import oracle, product
from ffi import buf, raw

pytestmark = pytest.mark.gpu


def _limit_api(L):
    L.xmss_mi355x_set_device_limit.restype = C.c_size_t
    L.xmss_mi355x_set_device_limit.argtypes = [C.c_size_t]
    L.xmss_mi355x_device_bytes.restype = C.c_size_t


def test_device_limit_turns_allocation_failures_into_minus_3():
    api = product(0)
    L = api.lib
    _limit_api(L)
    oid, mt = api.str_to_oid("XMSSMT-SHA2_20/2_256"), 1
    p = api.params(oid, mt)
    L.xmss_mi355x_rng_seed(99)
    rc, pk, sk = api.keypair(oid, mt)
    assert rc == 0
    # 2048 consecutive signatures need far more scratch than anything before
    cnt = 2048
    stride = p.sig_bytes + 32
    msg = buf(b"m" * 32)
    area = (C.c_ubyte * (cnt * stride))()
    smp = (C.c_void_p * cnt).from_buffer((np.arange(cnt, dtype=np.uint64) * stride + C.addressof(area)))
    mp = (C.c_void_p * cnt).from_buffer(np.full(cnt, C.addressof(msg), dtype=np.uint64))
    mlens = (C.c_ulonglong * cnt).from_buffer(np.full(cnt, 32, dtype=np.uint64))
    L.xmss_mi355x_sign_many.argtypes = [C.c_int] + [C.c_void_p] * 5 + [C.c_size_t, C.c_void_p]
    skb = buf(sk)
    smlens = (C.c_ulonglong * cnt)()
    rets = (C.c_int * cnt)()
    old = L.xmss_mi355x_set_device_limit(L.xmss_mi355x_device_bytes() + (4 << 20))
    try:
        assert L.xmss_mi355x_sign_many(mt, skb, smp, smlens, mp, mlens, cnt, rets) == -3
        assert set(rets) == {-3}
        assert raw(skb, len(sk)) == sk  # no signature left the call: the sk as it was
        # a job whose device buffers do not fit: its live inputs return -3
        L.xmss_batch_sign_open.argtypes = [C.c_int] + [C.c_void_p] * 6 + [C.c_size_t]
        k = 4096
        sig = bytes(p.sig_bytes)
        sms = [buf(sig) for _ in range(k)]
        pks = [buf(pk) for _ in range(k)]
        outs = [buf(p.sig_bytes) for _ in range(k)]
        arr = lambda xs: (C.c_void_p * len(xs))(*[C.cast(x, C.c_void_p) for x in xs])
        mls = (C.c_ulonglong * k)()
        vr = (C.c_int * k)()
        C.memset(vr, 0x55, C.sizeof(vr))  # every verdict must be written
        assert L.xmss_batch_sign_open(mt, arr(outs), mls, arr(sms), (C.c_ulonglong * k)(*[p.sig_bytes] * k),
                                      arr(pks), vr, k) == -3
        assert set(vr) == {-3}
    finally:
        L.xmss_mi355x_set_device_limit(old)
    # the library goes on: the same signatures now succeed, the key advances
    assert L.xmss_mi355x_sign_many(mt, skb, smp, smlens, mp, mlens, cnt, rets) == 0
    assert set(rets) == {0}
    assert raw(skb, len(sk))[4:4 + p.index_bytes] == cnt.to_bytes(p.index_bytes, "big")
    orc = oracle()
    for i in (0, cnt - 1):
        sm = bytes(area[i * stride:i * stride + smlens[i]])
        assert orc.sign_open(sm, pk, mt) == (0, b"m" * 32)
    # and a one-call signature under a limit that admits nothing new returns
    # -3 or succeeds (its buffers may already exist), never aborts; with the
    # limit lifted it equals the oracle
    sk1 = raw(skb, len(sk))
    rc, sm, sk2 = api.sign(sk1, b"x" * 100, mt)
    assert rc == 0
    rc_o, sm_o, sk2_o = orc.sign(sk1, b"x" * 100, mt, 0)
    assert (rc_o, sm_o, sk2_o) == (rc, sm, sk2)
