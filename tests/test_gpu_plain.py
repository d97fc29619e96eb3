"""Plain core on the GPU: byte parity with the reference's outputs.

Fixtures come from the reference's own C code (tests/golden/make_golden.py);
the oracle recomputes what the fixtures leave open.
"""
import struct

import pytest

from common import SEED, dg, golden, oracle, product

pytestmark = pytest.mark.gpu

XMSS_SHA2_10_256 = 0x01


def test_config1_keypair_and_16_signs():
    """XMSS-SHA2_10_256: seeded keypair, then leaves 0..15 (sign + verify)."""
    g = golden("plain_xmss_sha2_10_256.json")
    lib = product(0)
    lib.lib.xmss_mi355x_rng_seed(SEED)
    rc, pk, sk = lib.keypair(XMSS_SHA2_10_256, 0)
    assert rc == 0
    assert pk.hex() == g["pk"]
    assert sk.hex() == g["sk"]
    for i, s in enumerate(g["signs"]):
        m = bytes.fromhex(s["msg"])
        rc, sm, sk = lib.sign(sk, m, 0)
        assert rc == 0
        assert sm.hex() == s["sm"], "signature %d differs" % i
        vr, mm = lib.sign_open(sm, pk, 0)
        assert vr == 0 and mm == m
    assert sk.hex() == g["sk_after"]
    # bit flip in the message must fail and zero the output length
    bad = bytearray(sm)
    bad[-1] ^= 1
    vr, mm = lib.sign_open(bytes(bad), pk, 0)
    assert vr == -1 and mm == b""


def test_index_edge_cases():
    """idx 1022 signs normally; 1023 signs with the wiped key (idx field
    0xffffffff, fails to verify); the next call returns -2."""
    g = golden("plain_xmss_sha2_10_256.json")
    lib = product(0)
    pk = bytes.fromhex(g["pk"])
    for e in g["edge"]:
        rc, sm, sko = lib.sign(bytes.fromhex(e["sk_in"]), bytes.fromhex(e["msg"]), 0)
        assert rc == e["rc"]
        assert sko.hex() == e["sk_out"]
        if rc == 0:
            assert sm.hex() == e["sm"]
            assert lib.sign_open(sm, pk, 0)[0] == e["verify"]


@pytest.mark.parametrize("name", [
    "XMSS-SHA2_10_192", "XMSS-SHA2_10_512", "XMSS-SHAKE_10_256", "XMSS-SHAKE_10_512",
    "XMSS-SHAKE256_10_256", "XMSS-SHAKE256_10_192",
    "XMSSMT-SHA2_20/2_256", "XMSSMT-SHA2_20/4_256", "XMSSMT-SHA2_40/8_256", "XMSSMT-SHA2_60/12_256",
    "XMSSMT-SHA2_20/4_512", "XMSSMT-SHAKE_20/4_256", "XMSSMT-SHAKE_20/4_512", "XMSSMT-SHA2_20/4_192",
    "XMSSMT-SHAKE256_20/4_256", "XMSSMT-SHAKE256_20/4_192"])
def test_parameter_sets(name):
    """Every (func, n) family, XMSS and XMSS^MT: keypair, two signs at random
    indices with 32 B - 4 KB messages, verification."""
    g = golden("plain_sets.json")[name]
    mt = g["mt"]
    lib = product(0)
    lib.lib.xmss_mi355x_rng_seed(g["rng_seed"])
    rc, pk, sk = lib.keypair(g["oid"], mt)
    assert rc == 0
    assert pk.hex() == g["pk"] and sk.hex() == g["sk"]
    p = lib.params(g["oid"], mt)
    for s in g["signs"]:
        skx = bytearray(sk)
        skx[4:4 + p.index_bytes] = s["idx"].to_bytes(p.index_bytes, "big")
        m = bytes.fromhex(s["msg"])
        rc, sm, sko = lib.sign(bytes(skx), m, mt)
        assert rc == 0
        assert dg(sm) == s["sm"]
        assert dg(sko) == s["sk_out"]
        assert lib.sign_open(sm, pk, mt) == (0, m)


def test_seed_keypair_matches_oracle_small_trees():
    """xmssmt_core_seed_keypair against the oracle for every family at tree
    height 5 (XMSS^MT 20/4 sets)."""
    o = oracle()
    lib = product(0)
    for oid in (0x02, 0x0a, 0x12, 0x1a, 0x22, 0x2a, 0x32):
        p = o.params(oid, 1)
        seed = bytes((7 * i + oid) & 255 for i in range(3 * p.n))
        _, pk_o, sk_o = o.seed_keypair(oid, 1, seed)
        _, pk_g, sk_g = lib.seed_keypair(oid, 1, seed)
        assert pk_g == pk_o and sk_g == sk_o, hex(oid)


@pytest.mark.parametrize("name", ["XMSSMT-SHA2_20/4_256", "XMSSMT-SHA2_60/12_256", "XMSSMT-SHAKE_40/8_256"])
def test_mt_verify_each_layer_corrupted(name):
    """One-call XMSS^MT verification (upper layers' chains hashed for every
    start digit beside layer 0, then looked up): a valid signature verifies;
    one flipped byte in any layer's WOTS signature or auth path, or in R,
    is rejected -- each verdict and output equal to the oracle's."""
    import random
    o = oracle()
    lib = product(0)
    oid = lib.str_to_oid(name)
    p = lib.params(oid, 1)
    lib.lib.xmss_mi355x_rng_seed(oid * 7 + 3)
    rc, pk, sk = lib.keypair(oid, 1)
    assert rc == 0
    rng = random.Random(oid)
    skx = bytearray(sk)
    idx = rng.randrange(1 << p.full_height)
    skx[4:4 + p.index_bytes] = idx.to_bytes(p.index_bytes, "big")
    m = bytes(rng.getrandbits(8) for _ in range(77))
    rc, sm, _ = lib.sign(bytes(skx), m, 1)
    assert rc == 0
    assert lib.sign_open(sm, pk, 1) == (0, m) == o.sign_open(sm, pk, 1)
    ib, n, wb = p.index_bytes, p.n, p.wots_sig_bytes
    layer = wb + p.tree_height * n
    cases = [ib + rng.randrange(n)]  # R
    for L in range(p.d):
        cases.append(ib + n + L * layer + rng.randrange(wb))                               # WOTS signature
        cases.append(ib + n + L * layer + wb + rng.randrange(p.tree_height * n))          # auth path
    for pos in cases:
        bad = bytearray(sm)
        bad[pos] ^= 1 << rng.randrange(8)
        got = lib.sign_open(bytes(bad), pk, 1)
        assert got[0] == -1, (name, pos)
        assert got == o.sign_open(bytes(bad), pk, 1), (name, pos)
