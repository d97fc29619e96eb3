"""Checks the libraries make on the host before any device work (no GPU):
over-long messages, signed messages shorter than a signature, and the
device-memory limit API."""
import ctypes as C
import os

import pytest

from ffi import Api, PRODUCT_DIR, buf, raw, ull

LIBS = ["libxmss_mi355x.so", "libxmss_mi355x_fast.so"]
BIG = (1 << 32) + 5  # more than the 0xFFFFF000-byte limit; no buffer that long exists


def _api(lib):
    return Api(os.path.join(PRODUCT_DIR, lib), int(lib.endswith("_fast.so")))


@pytest.mark.parametrize("lib", LIBS)
@pytest.mark.parametrize("name", ["XMSS-SHA2_10_256", "XMSSMT-SHA2_20/4_256"])
def test_sign_refuses_messages_over_the_limit(lib, name):
    """The length check comes before anything is written: the caller's sk and
    sm stay as they were (the buffers are far shorter than the claimed
    length, so any copy would read and write out of bounds)."""
    api = _api(lib)
    mt = name.startswith("XMSSMT")
    oid = api.str_to_oid(name)
    p = api.params(oid, mt)
    sk0 = oid.to_bytes(4, "big") + bytes((i * 7) & 255 for i in range(p.sk_bytes))
    sk = buf(sk0)
    sm = buf(p.sig_bytes + 64)
    C.memset(sm, 0xA5, p.sig_bytes + 64)
    smlen = ull(12345)
    m = buf(64)
    f = api.lib.xmssmt_sign if mt else api.lib.xmss_sign
    assert f(sk, sm, C.byref(smlen), m, BIG) == -1
    assert raw(sk, len(sk0)) == sk0
    assert raw(sm, p.sig_bytes + 64) == b"\xa5" * (p.sig_bytes + 64)
    assert smlen.value == 12345


@pytest.mark.parametrize("lib", LIBS)
def test_sign_open_refuses_short_and_over_long_signed_messages(lib):
    """smlen < sig_bytes (where the reference underflows *mlen,
    xmss_commons.c:159) and a message part over the limit: -1, *mlen = 0,
    nothing read past the buffers."""
    api = _api(lib)
    oid = api.str_to_oid("XMSS-SHA2_10_256")
    p = api.params(oid, 0)
    pk = buf(oid.to_bytes(4, "big") + bytes(64))
    m = buf(p.sig_bytes + 64)
    mlen = ull(777)
    sm = buf(p.sig_bytes + 64)
    for smlen in (0, 1, p.sig_bytes - 1, p.sig_bytes + BIG):
        mlen.value = 777
        assert api.lib.xmss_sign_open(m, C.byref(mlen), sm, smlen, pk) == -1, smlen
        assert mlen.value == 0


def test_batch_sign_refuses_an_over_long_message():
    """A sign job leaves an over-long request's sk unchanged and returns -1
    for it; the rest of the job is unaffected by it (here they report the
    missing GPU)."""
    api = _api("libxmss_mi355x.so")
    L = api.lib
    oid = api.str_to_oid("XMSS-SHA2_10_256")
    p = api.params(oid, 0)
    sk0 = oid.to_bytes(4, "big") + (5).to_bytes(4, "big") + bytes(range(128))
    sks = [buf(sk0), buf(sk0)]
    ms = [buf(32), buf(32)]
    arr = lambda xs: (C.c_void_p * len(xs))(*[C.cast(x, C.c_void_p) for x in xs])
    mlens = (ull * 2)(BIG, 32)
    L.xmss_mi355x_device_ok.restype = C.c_int
    L.xmss_mi355x_sign_prepare.restype = C.c_void_p
    L.xmss_mi355x_sign_prepare.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    L.xmss_mi355x_sign_collect.argtypes = [C.c_void_p] * 4
    L.xmss_mi355x_job_free.argtypes = [C.c_void_p]
    if L.xmss_mi355x_device_ok() == 1:
        pytest.skip("a GPU is present (the GPU tests cover this job)")
    job = L.xmss_mi355x_sign_prepare(0, arr(sks), arr(ms), mlens, 2)
    sms = [buf(p.sig_bytes + 32), buf(p.sig_bytes + 32)]
    smlens = (ull * 2)()
    rets = (C.c_int * 2)()
    assert L.xmss_mi355x_sign_collect(job, arr(sms), smlens, rets) == 0
    L.xmss_mi355x_job_free(job)
    assert rets[0] == -1 and raw(sks[0], len(sk0)) == sk0
    assert rets[1] == -3


@pytest.mark.parametrize("lib", LIBS)
def test_device_limit_api(lib):
    so = C.CDLL(os.path.join(PRODUCT_DIR, lib))
    so.xmss_mi355x_set_device_limit.restype = C.c_size_t
    so.xmss_mi355x_set_device_limit.argtypes = [C.c_size_t]
    so.xmss_mi355x_device_bytes.restype = C.c_size_t
    assert so.xmss_mi355x_set_device_limit(1 << 20) == 0
    assert so.xmss_mi355x_set_device_limit(0) == 1 << 20
    assert so.xmss_mi355x_device_bytes() >= 0
