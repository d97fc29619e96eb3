"""Full-size trees on the GPU: the reference ships zero-seed keypairs of
height-20 (sub)trees as data (cuxmss/gpu/*_key.h) for SHA-256 and SHAKE256,
n = 24 and 32, and (h, d) = (20, 1), (40, 2), (60, 3).  The XMSS-SHA2_20_256
one was re-derived by the reference CPU code (1994 s on one core); the GPU
must reproduce every one of the twelve roots from seed 0^(3n) with
xmssmt_core_seed_keypair (2^20 leaves each)."""
import pytest

from common import golden, product

pytestmark = pytest.mark.gpu


NAMES = sorted(golden("ref_gpu_keys.json")["pk"])


@pytest.mark.parametrize("name", NAMES)
def test_zero_seed_height20_root(name):
    pk_want = golden("ref_gpu_keys.json")["pk"][name]
    lib = product(0)
    oid = lib.str_to_oid(name)
    mt = name.startswith("XMSSMT")
    p = lib.params(oid, mt)
    rc, pk, sk = lib.seed_keypair(oid, mt, bytes(3 * p.n))
    assert rc == 0
    assert pk.hex() == pk_want
