"""BDS core (libxmss_mi355x_fast.so) on the GPU: the serialized traversal
state in sk and every signature must match the reference's
xmss_core_fast.c byte for byte, over a key's whole life."""
import pytest

import random

from common import dg, golden, oracle, product
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


def test_bds_xmssmt_sha2_40_4_256_1100_signs():
    """XMSSMT-SHA2_40/4_256 (four layers of 2^10-leaf subtrees, the served
    set of the largest height): seeded keypair, 1100 signatures across the
    first layer-0 subtree switch."""
    _replay(product(1), golden("bds_xmssmt_sha2_40_4_256.json"), 1)


@pytest.mark.parametrize("name", ["XMSS-SHA2_10_256", "XMSSMT-SHA2_20/4_256"])
def test_bds_keypair_into_a_buffer_that_is_not_zero(name):
    """The BDS keypair leaves some state fields unwritten
    (xmss_core_fast.c:656-701): they keep the caller's bytes, as the
    reference's do; four signatures after it equal the reference's too."""
    import workloads as W
    g = golden("nonzero_sk_keypairs.json")[name]
    lib = product(1)
    mt = g["mt"]
    p = lib.params(g["oid"], mt)
    init = random.Random(g["init_seed"]).randbytes(4 + p.sk_bytes)
    assert any(init[4:])
    lib.lib.xmss_mi355x_rng_seed(g["rng_seed"])
    rc, pk, sk = lib.keypair(g["oid"], mt, sk_init=init)
    assert rc == 0
    assert pk.hex() == g["pk"]
    assert sk.hex() == g["sk"]
    for i, row in enumerate(g["signs"]):
        rc, sm, sk = lib.sign(sk, W.fixed_message(W.SEED + 952, i), mt)
        assert rc == row["rc"]
        assert dg(sm) == row["sm"], "sign %d" % i
        assert dg(sk) == row["sk"], "sign %d sk" % i


def test_bds_xmss_sha2_16_256_keygen_and_signs():
    """XMSS-SHA2_16_256 (65,536 leaves): the BDS keypair and 16 signatures."""
    _replay(product(1), golden("bds_xmss_sha2_16_256.json"), 0)


MT_FAMILIES = ["XMSSMT-SHA2_20/4_192", "XMSSMT-SHA2_20/4_512", "XMSSMT-SHAKE_20/4_256",
               "XMSSMT-SHAKE_20/4_512", "XMSSMT-SHAKE256_20/4_256", "XMSSMT-SHAKE256_20/4_192"]


@pytest.mark.parametrize("name", MT_FAMILIES)
def test_bds_every_hash_family_against_oracle(name):
    """BDS keypair and 40 signatures of an XMSSMT-*_20/4 set of every other
    hash family (SHA-256 n = 24, SHA-512, SHAKE128, SHAKE256 n = 24/32/64):
    the signatures cross the layer-0 subtree switch at index 32, messages are
    32 B - 2 KB; sk and signed message equal the oracle's BDS core after
    every signature."""
    lib, orc = product(1), oracle()
    oid = orc.str_to_oid(name)
    p = lib.params(oid, 1)
    seed = 0x5EED00 + oid
    lib.lib.xmss_mi355x_rng_seed(seed)
    orc.seed(seed)
    rc, pk, sk = lib.keypair(oid, 1, sk_init=bytes(4 + p.sk_bytes))
    rc2, pk2, sk2 = orc.keypair(oid, 1, 1, sk_init=bytes(4 + p.sk_bytes))
    assert rc == rc2 == 0
    assert pk == pk2 and sk == sk2
    rng = random.Random(oid)
    for i in range(40):
        m = bytes(rng.getrandbits(8) for _ in range(int(32 * 64 ** rng.random())))
        rc, sm, sk = lib.sign(sk, m, 1)
        rc2, sm2, sk2 = orc.sign(sk2, m, 1, 1)
        assert rc == rc2 == 0, i
        assert sm == sm2, "sign %d sm" % i
        assert sk == sk2, "sign %d sk" % i
    vr, mm = lib.sign_open(sm, pk, 1)
    assert vr == 0 and mm == m


def test_bds_interleaved_keys_and_an_older_state():
    """A signature finishes the leaves whose chains its key's previous call
    hashed ahead; here two keys alternate (the chain-end slots follow the
    key of each call), and one key then signs again from an older copy of
    its sk (the chains hashed ahead were for another state): every
    signature and sk still equals the reference's."""
    lib = product(1)
    runs = []
    for name, mt in (("bds_xmss_sha2_10_256.json", 0), ("bds_xmssmt_sha2_20_4_256.json", 1)):
        g = golden(name)
        p = lib.params(g["oid"], mt)
        if mt:
            lib.lib.xmss_mi355x_rng_seed(g["rng_seed"])
        rc, pk, sk = lib.keypair(g["oid"], mt, sk_init=bytes(4 + p.sk_bytes))
        assert rc == 0 and sk.hex() == g["sk"]
        runs.append([g, mt, sk, random.Random(g["msg_seed"]), []])

    def step(run, i):
        g, mt, sk, rng, saved = run
        m = bytes(rng.getrandbits(8) for _ in range(32))
        saved.append((sk, m))
        rc, sm, sk = lib.sign(sk, m, mt)
        row = g["signs"][i]
        assert rc == row["rc"] and dg(sm) == row["sm"] and dg(sk) == row["sk"], "%s sign %d" % (g["oid"], i)
        run[2] = sk

    for i in range(40):
        for run in runs:
            step(run, i)
    g, mt, _, _, saved = runs[0]
    for i in (17, 18, 5, 6, 7):
        sk, m = saved[i]
        rc, sm, sk2 = lib.sign(sk, m, mt)
        row = g["signs"][i]
        assert rc == row["rc"] and dg(sm) == row["sm"] and dg(sk2) == row["sk"], "again from state %d" % i


_DAG_CHILD = r'''
import os, random, sys
sys.path.insert(0, sys.argv[1])
from common import dg, golden, product
lib = product(1)
g = golden("bds_xmss_sha2_10_256.json")
p = lib.params(g["oid"], 0)
rc, pk, sk = lib.keypair(g["oid"], 0, sk_init=bytes(4 + p.sk_bytes))
assert rc == 0 and sk.hex() == g["sk"]
rng = random.Random(g["msg_seed"])
for i, row in enumerate(g["signs"][:int(sys.argv[2])]):
    m = bytes(rng.getrandbits(8) for _ in range(32))
    rc, sm, sk = lib.sign(sk, m, 0)
    assert rc == row["rc"] and dg(sm) == row["sm"] and dg(sk) == row["sk"], "sign %d" % i
print("ok")
'''


@pytest.mark.parametrize("spin", ["1", "0"])
def test_bds_merge_dag_fallbacks(spin):
    """A one-call signature hashes its merge DAG in the finishing launch once
    the planned leaves' blocks have counted in, a wait bounded by
    XMSS_MI355X_DAG_SPIN polls; with one poll the wait mostly gives up and the
    update launch hashes the DAG, with 0 the finishing launch never tries.
    Either way the signatures and sk equal the reference's."""
    import os
    import subprocess
    import sys
    tests = os.path.dirname(os.path.abspath(__file__))
    env = dict(os.environ, XMSS_MI355X_DAG_SPIN=spin)
    r = subprocess.run([sys.executable, "-c", _DAG_CHILD, tests, "200"], env=env, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stderr[-2000:]
