"""The CPU oracle (oracle/) pinned against the reference's own outputs.

Pins, per function of the path:
  * SHA-256 / SHA-512 / SHAKE128 / SHAKE256: hashlib over every padding
    boundary (FIPS 180-4 / FIPS 202 library routines) + FIPS known answers.
  * prf, prf_keygen, thash_f, thash_h, hash_message, WOTS+ and gen_leaf_wots
    of all seven (func, n) families: vectors written by the reference build.
  * WOTS / L-tree / treehash / sign: the reference's KAT emitter output
    (cuxmss/test/vectors.c) and full sm/sk bytes from the reference build.
  * treehash vs brute force on tiny trees built from gen_leaf + thash_h.
  * BDS core: the reference's serialized sk state after every signature.
"""
import ctypes as C
import hashlib
import os
import random

import pytest

from common import SEED, dg, golden, oracle
from ffi import Params, addr, buf, raw


# ---------------------------------------------------------------- primitives
BOUNDARIES = [0, 1, 3, 4, 55, 56, 63, 64, 65, 111, 112, 119, 120, 127, 128, 129, 135, 136, 137,
              167, 168, 169, 200, 255, 256, 257, 1000, 4096]


def test_primitives_match_hashlib():
    o = oracle()
    rng = random.Random(1)
    for L in BOUNDARIES:
        d = bytes(rng.getrandbits(8) for _ in range(L))
        assert o.sha256(d) == hashlib.sha256(d).digest(), L
        assert o.sha512(d) == hashlib.sha512(d).digest(), L
        for k in (10, 24, 32, 64, 200):
            assert o.shake128(d, k) == hashlib.shake_128(d).digest(k), (L, k)
            assert o.shake256(d, k) == hashlib.shake_256(d).digest(k), (L, k)


def test_primitives_fips_known_answers():
    o = oracle()
    assert o.sha256(b"abc").hex() == "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad"
    assert o.sha512(b"abc").hex().startswith("ddaf35a193617abacc417349ae204131")
    assert o.shake128(b"", 16).hex() == "7f9c2ba4e88f827d616045507605853e"
    assert o.shake256(b"", 16).hex() == "46b9dd2b0ba88d13233b3feb743eeb24"


# ---------------------------------------------------------------- hash layer
@pytest.mark.parametrize("fam", ["SHA2_256", "SHA2_512", "SHA2_192", "SHAKE128_256", "SHAKE256_512",
                                 "SHAKE256_256", "SHAKE256_192"])
def test_hash_layer_vectors(fam):
    o = oracle()
    g = golden("hash_layer.json")[fam]
    p = o.params(g["oid"], 0)
    n = p.n
    L = o.lib
    for c in g["cases"]:
        out = buf(n)
        L.xo_prf(C.byref(p), out, buf(bytes.fromhex(c["in32"])), buf(bytes.fromhex(c["key"])))
        assert raw(out, n).hex() == c["prf"]
        L.xo_prf_keygen(C.byref(p), out, buf(bytes.fromhex(c["seed2"] + c["in32"])), buf(bytes.fromhex(c["key"])))
        assert raw(out, n).hex() == c["prf_keygen"]
        a = addr(c["addr"])
        L.xo_thash_f(C.byref(p), out, buf(bytes.fromhex(c["x"])), buf(bytes.fromhex(c["key"])), a)
        assert raw(out, n).hex() == c["thash_f"]
        a = addr(c["addr"])
        L.xo_thash_h(C.byref(p), out, buf(bytes.fromhex(c["x2"])), buf(bytes.fromhex(c["key"])), a)
        assert raw(out, n).hex() == c["thash_h"]
    for hm in g["hash_message"]:
        m = bytes.fromhex(hm["m"])
        pre = p.padding_len + 3 * n
        mb = buf(bytes(pre) + m)
        out = buf(n)
        L.xo_hash_message(C.byref(p), out, buf(bytes.fromhex(hm["R"])), buf(bytes.fromhex(hm["root"])),
                          hm["idx"], mb, len(m))
        assert raw(out, n).hex() == hm["out"], len(m)
    w = g["wots"]
    ss, ps, msg = (buf(bytes.fromhex(w[k])) for k in ("sk_seed", "pub_seed", "msg"))
    pk, sig, pk2, leaf = buf(p.wots_sig_bytes), buf(p.wots_sig_bytes), buf(p.wots_sig_bytes), buf(n)
    L.xo_wots_pkgen(C.byref(p), pk, ss, ps, addr(w["ots_addr"]))
    assert raw(pk, p.wots_sig_bytes).hex() == w["pk"]
    L.xo_wots_sign(C.byref(p), sig, msg, ss, ps, addr(w["ots_addr"]))
    assert raw(sig, p.wots_sig_bytes).hex() == w["sig"]
    L.xo_wots_pk_from_sig(C.byref(p), pk2, sig, msg, ps, addr(w["ots_addr"]))
    assert raw(pk2, p.wots_sig_bytes).hex() == w["pk_from_sig"] == w["pk"]
    L.xo_gen_leaf_wots(C.byref(p), leaf, ss, ps, addr(w["ltree_addr"]), addr(w["ots_addr"]))
    assert raw(leaf, n).hex() == w["leaf"]


# ---------------------------------------------------------------- KAT emitter
def _kat_lines():
    with open(os.path.join(os.path.dirname(__file__), "golden", "kat_vectors.txt")) as f:
        return [ln.split() for ln in f.read().splitlines() if ln.strip()]


def _h10(b):
    return hashlib.shake_128(bytes(b)).hexdigest(10)


def test_kat_wots_lines():
    """WOTS+ lines of cuxmss/test/vectors.c:71-112."""
    o = oracle()
    L = o.lib
    for row in _kat_lines():
        if row[0] != "WOTS+":
            continue
        oid = int(row[1])
        p = o.params(oid, 0)
        a = addr([500000000 * i for i in range(8)])
        a2 = addr([400000000 * i for i in range(8)])
        m = bytes((3 * i) & 255 for i in range(p.n))
        ps = bytes((2 * i) & 255 for i in range(p.n))
        ss = bytes(i for i in range(p.n))
        pk, sig, leaf = buf(p.wots_sig_bytes), buf(p.wots_sig_bytes), buf(p.n)
        L.xo_wots_pkgen(C.byref(p), pk, buf(ss), buf(ps), a)
        L.xo_wots_sign(C.byref(p), sig, buf(m), buf(ss), buf(ps), a)
        L.xo_gen_leaf_wots(C.byref(p), leaf, buf(ss), buf(ps), a, a2)
        assert [_h10(raw(pk, p.wots_sig_bytes)), _h10(raw(sig, p.wots_sig_bytes)), _h10(raw(leaf, p.n))] == row[2:]


def _kat_xmss(o, oid, mt):
    p = o.params(oid, mt)
    _, pk, sk = o.seed_keypair(oid, mt, bytes(i & 255 for i in range(3 * p.n)))
    sk = bytearray(sk)
    sk[:p.index_bytes] = (1 << (p.full_height - 1)).to_bytes(p.index_bytes, "big")
    skb, sm, smlen = buf(bytes(sk)), buf(p.sig_bytes + 1), C.c_ulonglong(0)
    o.lib.xo_plain_sign(C.byref(p), skb, sm, C.byref(smlen), buf(b"\x25"), 1)
    return [_h10(pk), _h10(raw(sm, p.sig_bytes))]


def test_kat_xmssmt_lines():
    """XMSSMT lines of vectors.c:29-69 (seed[i] = i, idx = 2^(h-1), msg {37})."""
    o = oracle()
    for row in _kat_lines():
        if row[0] == "XMSSMT":
            assert _kat_xmss(o, int(row[1]), 1) == row[2:], row[1]


@pytest.mark.parametrize("oid", [1, 13, 19])
def test_kat_xmss_lines(oid):
    """XMSS lines of vectors.c for the n = 32 / 24 sets (the n = 64 ones are
    pinned by the hash-layer vectors and the 20/4_512 sets)."""
    o = oracle()
    row = [r for r in _kat_lines() if r[0] == "XMSS" and int(r[1]) == oid][0]
    assert _kat_xmss(o, oid, 0) == row[2:]


# ---------------------------------------------------------------- brute force
def test_treehash_against_brute_force_tiny_trees():
    """Root and auth path of treehash == a direct level-by-level tree of
    gen_leaf_wots leaves and thash_h, for custom heights 2..4."""
    o = oracle()
    L = o.lib
    for fh in (2, 3, 4):
        p = Params()
        p.func, p.n, p.padding_len, p.full_height, p.d, p.wots_w, p.bds_k = 0, 32, 32, fh, 1, 16, 0
        L.xo_initialize_params(C.byref(p), 0)
        ss = bytes(range(32))
        ps = bytes(range(32, 64))
        leaves = []
        for i in range(1 << fh):
            leaf = buf(32)
            L.xo_gen_leaf_wots(C.byref(p), leaf, buf(ss), buf(ps), addr([0, 0, 0, 1, i]), addr([0, 0, 0, 0, i]))
            leaves.append(raw(leaf, 32))
        levels = [leaves]
        for h in range(fh):
            nxt = []
            for i in range(len(levels[-1]) // 2):
                out = buf(32)
                L.xo_thash_h(C.byref(p), out, buf(levels[-1][2 * i] + levels[-1][2 * i + 1]), buf(ps),
                             addr([0, 0, 0, 2, 0, h, i]))
                nxt.append(raw(out, 32))
            levels.append(nxt)
        for leaf_idx in range(1 << fh):
            root, auth = buf(32), buf(32 * fh)
            L.xo_treehash(C.byref(p), root, auth, buf(ss), buf(ps), leaf_idx, addr([0] * 8))
            assert raw(root, 32) == levels[-1][0]
            want = b"".join(levels[h][(leaf_idx >> h) ^ 1] for h in range(fh))
            assert raw(auth, 32 * fh) == want
        # every aligned subtree's root (the CPU baseline's split keygen) is
        # the node of the directly built tree above it
        L.xo_subtree_root.argtypes = [C.POINTER(Params), C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                      C.c_uint, C.c_void_p]
        for h in range(fh + 1):
            for i in range(1 << (fh - h)):
                out = buf(32)
                L.xo_subtree_root(C.byref(p), out, buf(ss), buf(ps), i << h, h, addr([0] * 8))
                assert raw(out, 32) == levels[h][i], (fh, h, i)


# ---------------------------------------------------------------- plain core
def test_plain_config1_keypair_first_sign_and_edges():
    o = oracle()
    g = golden("plain_xmss_sha2_10_256.json")
    o.seed(SEED)
    rc, pk, sk = o.keypair(0x01, 0, 0)
    assert (pk.hex(), sk.hex()) == (g["pk"], g["sk"])
    s = g["signs"][0]
    rc, sm, _ = o.sign(sk, bytes.fromhex(s["msg"]), 0, 0)
    assert sm.hex() == s["sm"]
    assert o.sign_open(sm, pk, 0) == (0, bytes.fromhex(s["msg"]))
    for e in g["edge"][1:]:  # idx 1023 (wiped key still signs), then -2
        rc, sm, sko = o.sign(bytes.fromhex(e["sk_in"]), bytes.fromhex(e["msg"]), 0, 0)
        assert rc == e["rc"] and sko.hex() == e["sk_out"]
        if rc == 0:
            assert sm.hex() == e["sm"]
            assert o.sign_open(sm, pk, 0)[0] == e["verify"] == -1


@pytest.mark.parametrize("name", ["XMSS-SHA2_10_192", "XMSSMT-SHA2_20/4_256", "XMSSMT-SHA2_40/8_256",
                                  "XMSSMT-SHA2_60/12_256", "XMSSMT-SHA2_20/4_512", "XMSSMT-SHAKE_20/4_256",
                                  "XMSSMT-SHAKE_20/4_512", "XMSSMT-SHA2_20/4_192", "XMSSMT-SHAKE256_20/4_256",
                                  "XMSSMT-SHAKE256_20/4_192"])
def test_plain_parameter_sets(name):
    o = oracle()
    g = golden("plain_sets.json")[name]
    mt = g["mt"]
    o.seed(g["rng_seed"])
    rc, pk, sk = o.keypair(g["oid"], mt, 0)
    assert (pk.hex(), sk.hex()) == (g["pk"], g["sk"])
    p = o.params(g["oid"], mt)
    for s in g["signs"]:
        skx = bytearray(sk)
        skx[4:4 + p.index_bytes] = s["idx"].to_bytes(p.index_bytes, "big")
        m = bytes.fromhex(s["msg"])
        rc, sm, sko = o.sign(bytes(skx), m, mt, 0)
        assert (dg(sm), dg(sko)) == (s["sm"], s["sk_out"])
        assert o.sign_open(sm, pk, mt) == (0, m)
        bad = bytearray(sm)
        bad[p.index_bytes + 5] ^= 0x40  # one bit of R
        assert o.sign_open(bytes(bad), pk, mt) == (-1, b"")


# ---------------------------------------------------------------- BDS core
def _bds_replay(g, mt):
    o = oracle()
    p = o.params(g["oid"], mt, 1)
    if mt:
        o.seed(g["rng_seed"])
    rc, pk, sk = o.keypair(g["oid"], mt, 1, sk_init=bytes(4 + p.sk_bytes))
    assert (pk.hex(), sk.hex()) == (g["pk"], g["sk"])
    rng = random.Random(g["msg_seed"])
    for i, row in enumerate(g["signs"]):
        m = bytes(rng.getrandbits(8) for _ in range(32))
        rc, sm, sk = o.sign(sk, m, mt, 1)
        assert rc == row["rc"], i
        assert (dg(sm), dg(sk)) == (row["sm"], row["sk"]), i


def test_bds_xmss_whole_life():
    _bds_replay(golden("bds_xmss_sha2_10_256.json"), 0)


def test_bds_xmssmt_1100_signs():
    _bds_replay(golden("bds_xmssmt_sha2_20_4_256.json"), 1)


def test_bds_signatures_equal_plain_signatures():
    """Invariant: for the same seeds and index, BDS and plain signatures are
    byte-identical (only sk differs)."""
    o = oracle()
    g = golden("bds_xmssmt_sha2_20_4_256.json")
    o.seed(g["rng_seed"])
    p = o.params(g["oid"], 1, 1)
    _, pk, skb = o.keypair(g["oid"], 1, 1, sk_init=bytes(4 + p.sk_bytes))
    pp = o.params(g["oid"], 1, 0)
    skp = skb[:4 + pp.sk_bytes]
    for i in range(5):
        m = bytes([i]) * 40
        _, sm_b, skb = o.sign(skb, m, 1, 1)
        _, sm_p, skp = o.sign(skp, m, 1, 0)
        assert sm_b == sm_p
