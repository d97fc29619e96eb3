"""One key, many messages (xmss_mi355x_sign_many, the plain core): the
signatures, return values and final secret key must equal those of the same
number of successive xmss[mt]_sign calls -- across an XMSS^MT layer-0
subtree boundary and across the end of an XMSS key's life (the last index
signs with the wiped key, then -2)."""
import ctypes as C
import random

import pytest

from common import product
from ffi import buf, raw

pytestmark = pytest.mark.gpu


def _sign_many(api, sk, msgs, mt, p):
    L = api.lib
    L.xmss_mi355x_sign_many.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_size_t, C.c_void_p]
    n = len(msgs)
    skb = buf(sk)
    sms = [buf(p.sig_bytes + len(m)) for m in msgs]
    mbs = [buf(m) for m in msgs]
    smp = (C.c_void_p * n)(*[C.cast(x, C.c_void_p) for x in sms])
    mp = (C.c_void_p * n)(*[C.cast(x, C.c_void_p) for x in mbs])
    mlens = (C.c_ulonglong * n)(*[len(m) for m in msgs])
    smlens = (C.c_ulonglong * n)()
    rets = (C.c_int * n)()
    assert L.xmss_mi355x_sign_many(mt, skb, smp, smlens, mp, mlens, n, rets) == 0
    out = [(rets[i], raw(sms[i], smlens[i]) if rets[i] == 0 else None) for i in range(n)]
    return out, raw(skb, len(sk))


@pytest.mark.parametrize("name,mt,start,count", [("XMSS-SHA2_10_256", 0, 1000, 30),
                                                 ("XMSSMT-SHA2_20/4_256", 1, 25, 20),
                                                 ("XMSS-SHAKE_10_256", 0, 3, 5)])
def test_sign_many_equals_successive_signs(name, mt, start, count):
    api = product(0)
    oid = api.str_to_oid(name)
    p = api.params(oid, mt)
    api.lib.xmss_mi355x_rng_seed(0x0FF + oid + mt)
    rc, pk, sk = api.keypair(oid, mt)
    assert rc == 0
    sk = bytearray(sk)
    sk[4:4 + p.index_bytes] = start.to_bytes(p.index_bytes, "big")
    sk = bytes(sk)
    rng = random.Random(start)
    msgs = [bytes(rng.getrandbits(8) for _ in range(rng.randrange(0, 600))) for _ in range(count)]
    want, skw = [], sk
    for m in msgs:
        r, sm, skw = api.sign(skw, m, mt)
        want.append((r, sm if r == 0 else None))
    got, skg = _sign_many(api, sk, msgs, mt, p)
    assert [g[0] for g in got] == [w[0] for w in want]
    for i, (g, w) in enumerate(zip(got, want)):
        assert g[1] == w[1], "signature %d" % i
    assert skg == skw
    ok = [i for i, (r, sm) in enumerate(got) if r == 0 and start + i < (1 << p.full_height) - 1]
    for i in ok[:3] + ok[-2:]:
        vr, mm = api.sign_open(got[i][1], pk, mt)
        assert vr == 0 and mm == msgs[i]
