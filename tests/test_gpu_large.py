"""A verification batch too large for the latency kernels' scratch (60,000
XMSS-SHA2_16_256 signatures: the pre-images of all of them would exceed the
4 GiB the engine allows itself), which takes the one-workgroup-per-request
path: exactly the corrupted signatures are rejected, and sampled verdicts
equal the one-call API's.  The signatures come from one key through
sign_many (one tree pass)."""
import ctypes as C
import random

import pytest

from common import product
from ffi import buf, raw
from test_gpu_opk import _sign_many

pytestmark = pytest.mark.gpu


def test_verify_batch_beyond_latency_scratch():
    api = product(0)
    L = api.lib
    oid, mt = 0x02, 0  # XMSS-SHA2_16_256
    p = api.params(oid, mt)
    api.lib.xmss_mi355x_rng_seed(4242)
    rc, pk, sk = api.keypair(oid, mt)
    assert rc == 0
    rng = random.Random(5)
    k = 60000
    msgs = [bytes(rng.getrandbits(8) for _ in range(rng.randrange(1, 48))) for _ in range(k)]
    out, _ = _sign_many(api, sk, msgs, mt, p)
    assert all(rc == 0 for rc, _ in out)
    sms = [bytearray(sm) for _, sm in out]
    bad = set(rng.sample(range(k), 40))
    for i in bad:
        sms[i][rng.randrange(p.index_bytes, p.sig_bytes)] ^= 1 << rng.randrange(8)
    smb = [buf(bytes(s)) for s in sms]
    outs = [buf(len(s)) for s in sms]
    pkb = buf(pk)
    arr = lambda xs: (C.c_void_p * len(xs))(*[C.cast(x, C.c_void_p) for x in xs])
    smls = (C.c_ulonglong * k)(*[len(s) for s in sms])
    mls = (C.c_ulonglong * k)()
    vr = (C.c_int * k)()
    L.xmss_batch_sign_open.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_size_t]
    assert L.xmss_batch_sign_open(mt, arr(outs), mls, arr(smb), smls, arr([pkb] * k), vr, k) == 0
    got_bad = {i for i in range(k) if vr[i] != 0}
    assert got_bad == bad
    for i in list(bad)[:5] + rng.sample(range(k), 20):
        rc, m = api.sign_open(bytes(sms[i]), pk, mt)
        assert rc == vr[i], i
        if rc == 0:
            assert raw(outs[i], mls[i]) == m == msgs[i]
