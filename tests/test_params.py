"""Parameter sets and OIDs (CPU only): the product library, the oracle and
the reference build agree name by name and field by field; sizes of both
cores.  Parsing touches no GPU."""
import os

import pytest

from common import oracle
from ffi import Api, PRODUCT_DIR, load_ref

FIELDS = ["func", "n", "padding_len", "wots_w", "wots_log_w", "wots_len1", "wots_len2", "wots_len",
          "wots_sig_bytes", "full_height", "tree_height", "d", "index_bytes", "sig_bytes", "pk_bytes",
          "sk_bytes", "bds_k"]

XMSS_NAMES = ["XMSS-SHA2_%d_%d" % (h, b) for b in (256, 512) for h in (10, 16, 20)] + \
             ["XMSS-SHAKE_%d_%d" % (h, b) for b in (256, 512) for h in (10, 16, 20)] + \
             ["XMSS-SHA2_%d_192" % h for h in (10, 16, 20)] + \
             ["XMSS-SHAKE256_%d_%d" % (h, b) for b in (256, 192) for h in (10, 16, 20)]
SHAPES = ["20/2", "20/4", "40/2", "40/4", "40/8", "60/3", "60/6", "60/12"]
MT_NAMES = ["XMSSMT-%s_%s_%s" % (f, s, b) for f, b in (("SHA2", 256), ("SHA2", 512), ("SHAKE", 256),
                                                       ("SHAKE", 512), ("SHA2", 192), ("SHAKE256", 256),
                                                       ("SHAKE256", 192)) for s in SHAPES]


def _lib(core):
    return Api(os.path.join(PRODUCT_DIR, "libxmss_mi355x_fast.so" if core else "libxmss_mi355x.so"), core)


def _fields(p):
    return {f: getattr(p, f) for f in FIELDS}


def test_oids_are_the_reference_oids():
    lib = _lib(0)
    assert len(XMSS_NAMES) == 21 and len(MT_NAMES) == 56
    for i, name in enumerate(XMSS_NAMES):
        assert lib.str_to_oid(name) is not None
    assert sorted(lib.str_to_oid(n) for n in XMSS_NAMES) == list(range(1, 22))
    assert sorted(lib.str_to_oid(n) for n in MT_NAMES) == list(range(1, 57))
    assert lib.str_to_oid("XMSS-SHA2_10_255") is None
    assert lib.str_to_oid("XMSSMT-SHA2_20/3_256") is None


@pytest.mark.parametrize("core", [0, 1])
def test_params_match_oracle_and_reference(core):
    lib = _lib(core)
    o = oracle()
    ref = load_ref(core)
    for name in XMSS_NAMES + MT_NAMES:
        mt = name.startswith("XMSSMT")
        oid = lib.str_to_oid(name)
        assert oid == o.str_to_oid(name)
        mine = _fields(lib.params(oid, mt))
        assert mine == _fields(o.params(oid, mt, core)), name
        if ref is not None:
            assert ref.str_to_oid(name) == oid
            assert mine == _fields(ref.params(oid, mt)), name


def test_known_sizes():
    """Sizes quoted for the reference: XMSS-SHA2_10_256 sig 2500 B, 16_256 2692,
    MT-20/2 4963, 20/4 9251, 10_512 9092; BDS sk 1373 / 2093 / 5998 / 10938 / 15252."""
    plain, fast = _lib(0), _lib(1)
    sig = {"XMSS-SHA2_10_256": 2500, "XMSS-SHA2_16_256": 2692, "XMSSMT-SHA2_20/2_256": 4963,
           "XMSSMT-SHA2_20/4_256": 9251, "XMSS-SHA2_10_512": 9092}
    for name, want in sig.items():
        assert plain.params(plain.str_to_oid(name), name.startswith("XMSSMT")).sig_bytes == want
    bds = {"XMSS-SHA2_10_256": 1373, "XMSS-SHA2_16_256": 2093, "XMSSMT-SHA2_20/2_256": 5998,
           "XMSSMT-SHA2_20/4_256": 10938, "XMSSMT-SHA2_40/4_256": 15252}
    for name, want in bds.items():
        assert fast.params(fast.str_to_oid(name), name.startswith("XMSSMT")).sk_bytes == want
    assert plain.params(1, 0).sk_bytes == 132


def test_unknown_oid_rejected_without_gpu():
    """-1 for an unknown OID from every entry point, before any device work
    (xmss.c:15-17)."""
    import ctypes as C
    from ffi import buf, ull
    for core in (0, 1):
        lib = _lib(core).lib
        pk, sk, sm, m = buf(200), buf(4000), buf(4000), buf(4000)
        smlen, mlen = ull(0), ull(0)
        assert lib.xmss_keypair(pk, sk, 0x16) == -1
        assert lib.xmssmt_keypair(pk, sk, 0x39) == -1
        bad = buf(b"\x00\x00\x00\x00" + bytes(200))
        assert lib.xmss_sign(bad, sm, C.byref(smlen), m, 10) == -1
        assert lib.xmssmt_sign(bad, sm, C.byref(smlen), m, 10) == -1
        assert lib.xmss_sign_open(m, C.byref(mlen), sm, 3000, bad) == -1
        assert lib.xmssmt_sign_open(m, C.byref(mlen), sm, 3000, bad) == -1
