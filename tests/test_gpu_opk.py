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


def test_sign_many_whole_life_equals_reference():
    """The fixed-seed XMSS-SHA2_10_256 key of the BDS life fixture in
    plain-core form, signing its 1025 messages in one call: indices 0..1022
    equal the reference's signatures byte for byte (a BDS signature equals the
    plain one for the same seeds and index), index 1023 (wiped key) equals the
    reference plain core's, and the last call is -2 with the key wiped."""
    import workloads as W
    from common import dg, golden
    api = product(0)
    bds = golden("bds_xmss_sha2_10_256.json")
    life = golden("whole_life.json")["bds10_last_plain"]
    p = api.params(0x01, 0)
    root = bytes.fromhex(bds["pk"])[4:4 + p.n]
    sk = W.plain_sk(0x01, p, 0, b"\x02" * 64 + b"\x04" * 32, root)
    gen = W.bds10_messages()
    msgs = [next(gen) for _ in range(len(bds["signs"]))]
    got, sko = _sign_many(api, sk, msgs, 0, p)
    assert [g[0] for g in got] == [r["rc"] for r in bds["signs"]]
    for i in range(1023):
        assert dg(got[i][1]) == bds["signs"][i]["sm"], "signature %d" % i
    assert got[1023][0] == life["rc"] == 0 and dg(got[1023][1]) == life["sm"]
    assert sko == (1).to_bytes(4, "big") + b"\xff" * 4 + bytes(4 * p.n)


def test_xmssmt_whole_life_with_sign_many():
    """cuxmss/test/xmss_max_signatures.c:35-102 for XMSSMT-SHA2_20/2_256: a
    seeded key signs one random 32-byte message 2^20 times (the last index
    with the wiped key) and then fails twice with -2.  Beyond the return codes
    the reference test checks: every signature of the first 2^20 - 1 verifies
    in a batch, the signatures at sampled indices and at the last index equal
    the reference's (tests/golden/whole_life.json), and the sk after the two
    failures equals the reference's."""
    import numpy as np
    from common import dg, golden
    g = golden("whole_life.json")
    api = product(0)
    L = api.lib
    oid, mt = g["oid"], 1
    p = api.params(oid, mt)
    L.xmss_mi355x_rng_seed(g["rng_seed"])
    m = buf(32)
    L.randombytes.argtypes = [C.c_void_p, C.c_ulonglong]
    L.randombytes(m, 32)  # drawn before the keypair, as the reference test does
    msg = raw(m, 32)
    assert msg.hex() == g["msg"]
    rc, pk, sk = api.keypair(oid, mt)
    assert rc == 0 and pk.hex() == g["pk"] and sk.hex() == g["sk"]
    top = 1 << p.full_height
    samples = {s["idx"]: s for s in g["samples"]}
    L.xmss_mi355x_sign_many.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_size_t, C.c_void_p]
    L.xmss_batch_sign_open.argtypes = [C.c_int] + [C.c_void_p] * 6 + [C.c_size_t]
    stride = p.sig_bytes + 32
    chunk = 1 << 16
    skb = buf(sk)
    mb = buf(msg)
    total = top + 2
    pkb = buf(pk)
    for c0 in range(0, total, chunk):
        cnt = min(chunk, total - c0)
        area = (C.c_ubyte * (cnt * stride))()
        base = C.addressof(area)
        ptrs = (np.arange(cnt, dtype=np.uint64) * stride + base).astype(np.uint64)
        smp = (C.c_void_p * cnt).from_buffer(ptrs)
        mp = (C.c_void_p * cnt).from_buffer(np.full(cnt, C.addressof(mb), dtype=np.uint64))
        mlens = (C.c_ulonglong * cnt).from_buffer(np.full(cnt, 32, dtype=np.uint64))
        smlens = (C.c_ulonglong * cnt)()
        rets = (C.c_int * cnt)()
        assert L.xmss_mi355x_sign_many(mt, skb, smp, smlens, mp, mlens, cnt, rets) == 0
        r = np.frombuffer(rets, dtype=np.int32)
        want = np.where(np.arange(c0, c0 + cnt) < top, 0, -2)
        assert (r == want).all(), "return codes of chunk at %d" % c0
        for i in range(c0, c0 + cnt):
            if i in samples:
                sm = bytes(area[(i - c0) * stride:(i - c0) * stride + smlens[i - c0]])
                assert dg(sm) == samples[i]["sm"], "signature %d" % i
        if c0 <= top - 1 < c0 + cnt:  # the last index: the wiped key's signature
            i = top - 1 - c0
            assert dg(bytes(area[i * stride:i * stride + smlens[i]])) == g["last"][0]["sm"]
        # every valid signature of the chunk verifies in one batch
        nv = min(cnt, top - 1 - c0)
        if nv > 0:
            outs = (C.c_ubyte * (nv * stride))()
            optr = (C.c_void_p * nv).from_buffer(
                (np.arange(nv, dtype=np.uint64) * stride + C.addressof(outs)).astype(np.uint64))
            pks = (C.c_void_p * nv).from_buffer(np.full(nv, C.addressof(pkb), dtype=np.uint64))
            mls = (C.c_ulonglong * nv)()
            vr = (C.c_int * nv)()
            assert L.xmss_batch_sign_open(mt, optr, mls, smp, smlens, pks, vr, nv) == 0
            assert (np.frombuffer(vr, dtype=np.int32) == 0).all(), "verification of chunk at %d" % c0
        del area
    assert raw(skb, len(sk)).hex() == g["last"][-1]["sk_out"]
