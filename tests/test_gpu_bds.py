"""BDS core (libxmss_mi355x_fast.so) on the GPU: the serialized traversal
state in sk and every signature must match the reference's
xmss_core_fast.c byte for byte, over a key's whole life."""
import pytest

from common import dg, golden, product
from ffi import Params

pytestmark = pytest.mark.gpu


def _replay(lib, g, mt):
    import random
    p = lib.params(g["oid"], mt)
    if mt:
        lib.lib.xmss_mi355x_rng_seed(g["rng_seed"])
    rc, pk, sk = lib.keypair(g["oid"], mt, sk_init=bytes(4 + p.sk_bytes))
    assert rc == 0
    assert pk.hex() == g["pk"]
    assert sk.hex() == g["sk"]
    rng = random.Random(g["msg_seed"])
    for i, row in enumerate(g["signs"]):
        m = bytes(rng.getrandbits(8) for _ in range(32))
        rc, sm, sk = lib.sign(sk, m, mt)
        assert rc == row["rc"], "sign %d rc" % i
        if "sm_hex" in row:
            assert sm.hex() == row["sm_hex"], "sign %d sm" % i
            assert sk.hex() == row["sk_hex"], "sign %d sk" % i
        assert dg(sm) == row["sm"], "sign %d sm" % i
        assert dg(sk) == row["sk"], "sign %d sk" % i
        if rc == 0 and i % 97 == 0:
            vr, mm = lib.sign_open(sm, pk, mt)
            assert vr == (0 if i < (1 << p.full_height) - 1 else -1) and (vr != 0 or mm == m)


def test_bds_xmss_sha2_10_256_whole_life():
    """Fixed-seed keypair, all 1024 signatures (the last with the wiped key),
    then -2."""
    _replay(product(1), golden("bds_xmss_sha2_10_256.json"), 0)


def test_bds_xmssmt_sha2_20_4_256_1100_signs():
    """Seeded keypair, 1100 signatures: crosses 34 layer-0 and one layer-1
    subtree switch (NEXT-tree swap, WOTS re-signing of the new root)."""
    _replay(product(1), golden("bds_xmssmt_sha2_20_4_256.json"), 1)


def test_bds_xmss_sha2_16_256_keygen_and_signs():
    """XMSS-SHA2_16_256 (65,536 leaves): the BDS keypair and 16 signatures."""
    _replay(product(1), golden("bds_xmss_sha2_16_256.json"), 0)
