"""The allocation-free verification entry point (xmss_mi355x_sign_open_static)
through the C ABI: on two calls in a row, the second with a smaller input,
its return value, *mlen and every byte of m equal xmss[mt]_sign_open's on the
same signed message -- valid and corrupted signatures, XMSS and XMSS^MT --
and the library holds no more device memory after the calls than before.
A request larger than its static scratch returns -4."""
import ctypes as C
import random

import pytest

from common import product
from ffi import buf, raw

pytestmark = pytest.mark.gpu


def _static(L):
    L.xmss_mi355x_sign_open_static.restype = C.c_int
    L.xmss_mi355x_sign_open_static.argtypes = [C.c_int, C.c_void_p, C.POINTER(C.c_ulonglong), C.c_void_p,
                                               C.c_ulonglong, C.c_void_p]
    L.xmss_mi355x_device_bytes.restype = C.c_size_t


def _both(lib, mt, sm, pk):
    """(rc, mlen, m bytes) of the static entry point and of sign_open; m is
    pre-filled with the same pattern for both, so untouched bytes compare too."""
    L = lib.lib
    out = []
    for static in (True, False):
        cap = len(sm) + 64
        m = buf(bytes((i * 13) & 255 for i in range(cap)))
        mlen = C.c_ulonglong(999)
        if static:
            before = L.xmss_mi355x_device_bytes()
            rc = L.xmss_mi355x_sign_open_static(mt, m, C.byref(mlen), buf(sm), len(sm), buf(pk))
            assert L.xmss_mi355x_device_bytes() == before, "the static entry point allocated"
        else:
            f = L.xmssmt_sign_open if mt else L.xmss_sign_open
            rc = f(m, C.byref(mlen), buf(sm), len(sm), buf(pk))
        out.append((rc, mlen.value, raw(m, cap)))
    return out


@pytest.mark.parametrize("name", ["XMSS-SHA2_10_256", "XMSSMT-SHA2_20/4_256", "XMSSMT-SHAKE256_40/8_192",
                                  "XMSS-SHA2_10_512"])
def test_static_verify_equals_sign_open_on_two_calls(name):
    lib = product(1)
    L = lib.lib
    _static(L)
    mt = int(name.startswith("XMSSMT"))
    oid = lib.str_to_oid(name)
    p = lib.params(oid, mt)
    L.xmss_mi355x_rng_seed(oid * 11 + mt)
    rc, pk, sk = lib.keypair(oid, mt, sk_init=bytes(4 + p.sk_bytes))
    assert rc == 0
    rng = random.Random(oid)
    big = bytes(rng.getrandbits(8) for _ in range(3000))
    small = bytes(rng.getrandbits(8) for _ in range(40))
    rc, sm_big, sk = lib.sign(sk, big, mt)
    assert rc == 0
    rc, sm_small, sk = lib.sign(sk, small, mt)
    assert rc == 0
    bad = bytearray(sm_small)
    bad[p.index_bytes + p.n + 7] ^= 0x20
    # the library's device memory stays where it is across each static call
    for sm, ok, msg in ((sm_big, True, big), (sm_small, True, small), (bytes(bad), False, small)):
        got, want = _both(lib, mt, sm, pk)
        assert got == want
        assert got[0] == (0 if ok else -1)
        if ok:
            assert got[2][:got[1]] == msg


def test_static_verify_refuses_a_request_beyond_its_scratch():
    lib = product(1)
    L = lib.lib
    _static(L)
    oid = lib.str_to_oid("XMSS-SHA2_10_256")
    p = lib.params(oid, 0)
    L.xmss_mi355x_rng_seed(5)
    rc, pk, sk = lib.keypair(oid, 0, sk_init=bytes(4 + p.sk_bytes))
    msg = bytes(20 << 20)  # 20 MiB: more than the 16 MiB array holds
    rc, sm, sk = lib.sign(sk, msg, 0)
    assert rc == 0
    m = buf(len(sm))
    mlen = C.c_ulonglong(5)
    assert L.xmss_mi355x_sign_open_static(0, m, C.byref(mlen), buf(sm), len(sm), buf(pk)) == -4
    assert mlen.value == 0
    # the ordinary entry point takes it
    assert lib.sign_open(sm, pk, 0) == (0, msg)
