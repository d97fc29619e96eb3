"""ctypes bindings shared by the tests, the fixture generator and bench.py.

Three libraries speak the same C API:
  * the product:  cu-xmss_amd/lib/libxmss_mi355x.so (plain core) and
                  libxmss_mi355x_fast.so (BDS core) -- HIP on the GPU;
  * the oracle:   oracle/_build/liboracle.so -- CPU restatement (checker only);
  * the reference build: oracle/_ref/libxmssref{,_fast}.so -- the reference's
    own C sources compiled by oracle/ref.mk (checker only, where present).
"""
import ctypes as C
import os
import struct

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PRODUCT_DIR = os.path.join(ROOT, "cu-xmss_amd", "lib")
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

u8p = C.POINTER(C.c_ubyte)
ull = C.c_ulonglong


class Params(C.Structure):
    """Layout of xmss_params (cuxmss/params.h:15-33)."""
    _fields_ = [(name, C.c_uint) for name in (
        "func", "n", "padding_len", "wots_w", "wots_log_w", "wots_len1",
        "wots_len2", "wots_len", "wots_sig_bytes", "full_height", "tree_height",
        "d", "index_bytes", "sig_bytes", "pk_bytes")] + [
        ("sk_bytes", C.c_ulonglong), ("bds_k", C.c_uint)]


def buf(data_or_len):
    if isinstance(data_or_len, int):
        return (C.c_ubyte * max(1, data_or_len))()
    b = (C.c_ubyte * max(1, len(data_or_len)))()
    C.memmove(b, bytes(data_or_len), len(data_or_len))
    return b


def raw(b, n):
    return bytes(bytearray(b)[:n])


class Api:
    """The OID-prefixed C API (cuxmss/xmss.h:15-64) of one library."""

    def __init__(self, path, core):
        self.path = path
        self.core = core
        self.lib = C.CDLL(path)
        L = self.lib
        for nm in ("xmss_keypair", "xmssmt_keypair"):
            getattr(L, nm).argtypes = [u8p, u8p, C.c_uint32]
            getattr(L, nm).restype = C.c_int
        for nm in ("xmss_sign", "xmssmt_sign"):
            getattr(L, nm).argtypes = [u8p, u8p, C.POINTER(ull), u8p, ull]
            getattr(L, nm).restype = C.c_int
        for nm in ("xmss_sign_open", "xmssmt_sign_open"):
            getattr(L, nm).argtypes = [u8p, C.POINTER(ull), u8p, ull, u8p]
            getattr(L, nm).restype = C.c_int
        for nm in ("xmss_parse_oid", "xmssmt_parse_oid"):
            getattr(L, nm).argtypes = [C.POINTER(Params), C.c_uint32]
            getattr(L, nm).restype = C.c_int
        for nm in ("xmss_str_to_oid", "xmssmt_str_to_oid"):
            getattr(L, nm).argtypes = [C.POINTER(C.c_uint32), C.c_char_p]
            getattr(L, nm).restype = C.c_int
        if hasattr(L, "xmssmt_core_seed_keypair"):  # plain core only (xmss_core.c:134)
            L.xmssmt_core_seed_keypair.argtypes = [C.POINTER(Params), u8p, u8p, u8p]
            L.xmssmt_core_seed_keypair.restype = C.c_int

    def params(self, oid, mt):
        p = Params()
        rc = (self.lib.xmssmt_parse_oid if mt else self.lib.xmss_parse_oid)(C.byref(p), oid)
        return p if rc == 0 else None

    def str_to_oid(self, name):
        o = C.c_uint32()
        mt = name.startswith("XMSSMT")
        rc = (self.lib.xmssmt_str_to_oid if mt else self.lib.xmss_str_to_oid)(C.byref(o), name.encode())
        return o.value if rc == 0 else None

    def keypair(self, oid, mt, sk_init=None):
        p = self.params(oid, mt)
        pk = buf(4 + p.pk_bytes)
        sk = buf(sk_init) if sk_init is not None else buf(4 + p.sk_bytes)
        rc = (self.lib.xmssmt_keypair if mt else self.lib.xmss_keypair)(pk, sk, oid)
        return rc, raw(pk, 4 + p.pk_bytes), raw(sk, 4 + p.sk_bytes)

    def seed_keypair(self, oid, mt, seed):
        p = self.params(oid, mt)
        pk = buf(p.pk_bytes)
        sk = buf(p.sk_bytes)
        rc = self.lib.xmssmt_core_seed_keypair(C.byref(p), pk, sk, buf(seed))
        return rc, raw(pk, p.pk_bytes), raw(sk, p.sk_bytes)

    def sign(self, sk, msg, mt):
        """Returns (rc, sm, new_sk)."""
        oid = struct.unpack(">I", bytes(sk[:4]))[0]
        p = self.params(oid, mt)
        skb = buf(sk)
        sm = buf(p.sig_bytes + len(msg))
        smlen = ull(0)
        rc = (self.lib.xmssmt_sign if mt else self.lib.xmss_sign)(skb, sm, C.byref(smlen), buf(msg), len(msg))
        return rc, raw(sm, smlen.value if rc == 0 else 0), raw(skb, len(sk))

    def sign_open(self, sm, pk, mt):
        oid = struct.unpack(">I", bytes(pk[:4]))[0]
        p = self.params(oid, mt)
        m = buf(len(sm) + p.sig_bytes)
        mlen = ull(0)
        rc = (self.lib.xmssmt_sign_open if mt else self.lib.xmss_sign_open)(m, C.byref(mlen), buf(sm), len(sm), buf(pk))
        return rc, raw(m, mlen.value)


def load_ref(core):
    """Reference build (oracle/_ref); core 0 = xmss_core.c, 1 = xmss_core_fast.c."""
    path = os.path.join(REF_DIR, "libxmssref_fast.so" if core else "libxmssref.so")
    if not os.path.exists(path):
        return None
    api = Api(path, core)
    api.lib.ref_rng_seed.argtypes = [C.c_uint64]
    api.seed = api.lib.ref_rng_seed
    return api


class Oracle:
    """The CPU oracle (oracle/xo.h)."""

    def __init__(self, path=ORACLE_SO):
        self.lib = C.CDLL(path)
        L = self.lib
        L.xo_rng_seed.argtypes = [C.c_uint64]
        L.xo_randombytes.argtypes = [u8p, ull]
        for nm in ("xo_sha256", "xo_sha512"):
            getattr(L, nm).argtypes = [u8p, u8p, ull]
        for nm in ("xo_shake128", "xo_shake256"):
            getattr(L, nm).argtypes = [u8p, ull, u8p, ull]
        for nm in ("xo_xmss_parse_oid", "xo_xmssmt_parse_oid"):
            getattr(L, nm).argtypes = [C.POINTER(Params), C.c_uint32, C.c_int]
        for nm in ("xo_xmss_str_to_oid", "xo_xmssmt_str_to_oid"):
            getattr(L, nm).argtypes = [C.POINTER(C.c_uint32), C.c_char_p]
        L.xo_prf.argtypes = [C.POINTER(Params), u8p, u8p, u8p]
        L.xo_prf_keygen.argtypes = [C.POINTER(Params), u8p, u8p, u8p]
        L.xo_thash_f.argtypes = [C.POINTER(Params), u8p, u8p, u8p, C.POINTER(C.c_uint32)]
        L.xo_thash_h.argtypes = [C.POINTER(Params), u8p, u8p, u8p, C.POINTER(C.c_uint32)]
        L.xo_hash_message.argtypes = [C.POINTER(Params), u8p, u8p, u8p, ull, u8p, ull]
        L.xo_wots_pkgen.argtypes = [C.POINTER(Params), u8p, u8p, u8p, C.POINTER(C.c_uint32)]
        L.xo_wots_sign.argtypes = [C.POINTER(Params), u8p, u8p, u8p, u8p, C.POINTER(C.c_uint32)]
        L.xo_wots_pk_from_sig.argtypes = [C.POINTER(Params), u8p, u8p, u8p, u8p, C.POINTER(C.c_uint32)]
        L.xo_l_tree.argtypes = [C.POINTER(Params), u8p, u8p, u8p, C.POINTER(C.c_uint32)]
        L.xo_gen_leaf_wots.argtypes = [C.POINTER(Params), u8p, u8p, u8p,
                                       C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.xo_compute_root.argtypes = [C.POINTER(Params), u8p, u8p, C.c_ulong, u8p, u8p,
                                      C.POINTER(C.c_uint32)]
        L.xo_treehash.argtypes = [C.POINTER(Params), u8p, u8p, u8p, u8p, C.c_uint32,
                                  C.POINTER(C.c_uint32)]
        L.xo_plain_seed_keypair.argtypes = [C.POINTER(Params), u8p, u8p, u8p]
        L.xo_plain_sign.argtypes = [C.POINTER(Params), u8p, u8p, C.POINTER(ull), u8p, ull]
        L.xo_sign_open.argtypes = [C.POINTER(Params), u8p, C.POINTER(ull), u8p, ull, u8p]
        L.xo_api_keypair.argtypes = [u8p, u8p, C.c_uint32, C.c_int, C.c_int]
        L.xo_api_sign.argtypes = [u8p, u8p, C.POINTER(ull), u8p, ull, C.c_int, C.c_int]
        L.xo_api_sign_open.argtypes = [u8p, C.POINTER(ull), u8p, ull, u8p, C.c_int]

    # -- primitives --
    def sha256(self, data):
        o = buf(32)
        self.lib.xo_sha256(o, buf(data), len(data))
        return raw(o, 32)

    def sha512(self, data):
        o = buf(64)
        self.lib.xo_sha512(o, buf(data), len(data))
        return raw(o, 64)

    def shake128(self, data, outlen):
        o = buf(outlen)
        self.lib.xo_shake128(o, outlen, buf(data), len(data))
        return raw(o, outlen)

    def shake256(self, data, outlen):
        o = buf(outlen)
        self.lib.xo_shake256(o, outlen, buf(data), len(data))
        return raw(o, outlen)

    # -- params --
    def params(self, oid, mt, core=0):
        p = Params()
        rc = (self.lib.xo_xmssmt_parse_oid if mt else self.lib.xo_xmss_parse_oid)(C.byref(p), oid, core)
        return p if rc == 0 else None

    def str_to_oid(self, name):
        o = C.c_uint32()
        mt = name.startswith("XMSSMT")
        rc = (self.lib.xo_xmssmt_str_to_oid if mt else self.lib.xo_xmss_str_to_oid)(C.byref(o), name.encode())
        return o.value if rc == 0 else None

    # -- API mirror --
    def seed(self, s):
        self.lib.xo_rng_seed(s)

    def keypair(self, oid, mt, core, sk_init=None):
        p = self.params(oid, mt, core)
        pk = buf(4 + p.pk_bytes)
        sk = buf(sk_init) if sk_init is not None else buf(4 + p.sk_bytes)
        rc = self.lib.xo_api_keypair(pk, sk, oid, mt, core)
        return rc, raw(pk, 4 + p.pk_bytes), raw(sk, 4 + p.sk_bytes)

    def seed_keypair(self, oid, mt, seed):
        p = self.params(oid, mt, 0)
        pk = buf(p.pk_bytes)
        sk = buf(p.sk_bytes)
        self.lib.xo_plain_seed_keypair(C.byref(p), pk, sk, buf(seed))
        return 0, raw(pk, p.pk_bytes), raw(sk, p.sk_bytes)

    def sign(self, sk, msg, mt, core):
        oid = struct.unpack(">I", bytes(sk[:4]))[0]
        p = self.params(oid, mt, core)
        skb = buf(sk)
        sm = buf(p.sig_bytes + len(msg))
        smlen = ull(0)
        rc = self.lib.xo_api_sign(skb, sm, C.byref(smlen), buf(msg), len(msg), mt, core)
        return rc, raw(sm, smlen.value if rc == 0 else 0), raw(skb, len(sk))

    def sign_open(self, sm, pk, mt):
        oid = struct.unpack(">I", bytes(pk[:4]))[0]
        p = self.params(oid, mt)
        m = buf(len(sm) + p.sig_bytes)
        mlen = ull(0)
        rc = self.lib.xo_api_sign_open(m, C.byref(mlen), buf(sm), len(sm), buf(pk), mt)
        return rc, raw(m, mlen.value)


def addr(words):
    a = (C.c_uint32 * 8)()
    for i, w in enumerate(words):
        a[i] = w & 0xFFFFFFFF
    return a
