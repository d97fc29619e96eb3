"""Writes the golden fixtures under tests/golden/ from the reference build.

Run here (where /root/reference exists), after `make -C oracle ref`:
    python tests/golden/make_golden.py [--big]

Every fixture is data: inputs and the outputs the reference's own C code
(cuxmss/*.c, compiled unmodified by oracle/ref.mk with a seeded randombytes)
produced for them.  --big also writes the XMSS-SHA2_16_256 BDS keypair
fixture (about two minutes of reference CPU time).
"""
import ctypes as C
import hashlib
import json
import math
import os
import random
import re
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from ffi import Params, addr, buf, load_ref, raw, ull, u8p, REF_DIR  # noqa: E402

SEED = 0x584D5353  # "XMSS"


def h(b):
    return bytes(b).hex()


def dg(b):
    return hashlib.sha256(bytes(b)).hexdigest()


def msg_len(rng):
    # log-uniform over 32 B .. 4 KB
    return int(32 * 128 ** rng.random())


def write(name, obj):
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(obj, f, indent=1, sort_keys=True)
    print("wrote", name, flush=True)


def kat_vectors():
    out = subprocess.run([os.path.join(REF_DIR, "vectors")], capture_output=True, text=True, check=True).stdout
    with open(os.path.join(HERE, "kat_vectors.txt"), "w") as f:
        f.write(out)
    print("wrote kat_vectors.txt", flush=True)


def ref_gpu_keys():
    """Public keys of the zero-seed keypairs the reference ships as data
    (cuxmss/gpu/*_key.h), keyed by parameter-set name."""
    names = {
        ("sha256_key.h", 20): "XMSS-SHA2_20_256", ("sha256_key.h", 40): "XMSSMT-SHA2_40/2_256",
        ("sha256_key.h", 60): "XMSSMT-SHA2_60/3_256",
        ("sha256_192_key.h", 20): "XMSS-SHA2_20_192", ("sha256_192_key.h", 40): "XMSSMT-SHA2_40/2_192",
        ("sha256_192_key.h", 60): "XMSSMT-SHA2_60/3_192",
        ("shake256_key.h", 20): "XMSS-SHAKE256_20_256", ("shake256_key.h", 40): "XMSSMT-SHAKE256_40/2_256",
        ("shake256_key.h", 60): "XMSSMT-SHAKE256_60/3_256",
        ("shake256_192_key.h", 20): "XMSS-SHAKE256_20_192", ("shake256_192_key.h", 40): "XMSSMT-SHAKE256_40/2_192",
        ("shake256_192_key.h", 60): "XMSSMT-SHAKE256_60/3_192",
    }
    res = {}
    for fn in ("sha256_key.h", "sha256_192_key.h", "shake256_key.h", "shake256_192_key.h"):
        text = open(os.path.join("/root/reference/cuxmss/gpu", fn)).read()
        blocks = re.split(r"#(?:el)?if FULL_HEIGHT == ", text)[1:]
        for b in blocks:
            fh = int(b.split()[0])
            m = re.search(r"public_pk\[\]\s*=\s*\{([^}]*)\}", b)
            pk = bytes(int(x, 16) for x in m.group(1).replace("\n", "").split(",") if x.strip())
            res[names[(fn, fh)]] = h(pk)
    write("ref_gpu_keys.json", {"source": "cuxmss/gpu/*_key.h public_pk (all-zero seeds)", "pk": res})


def hash_layer(ref):
    """prf / prf_keygen / thash_f / thash_h / hash_message / WOTS / leaf for
    every (func, n) family, on seeded random inputs."""
    L = ref.lib
    P = C.POINTER(Params)
    L.prf.argtypes = [P, u8p, u8p, u8p]
    L.prf_keygen.argtypes = [P, u8p, u8p, u8p]
    L.thash_f.argtypes = [P, u8p, u8p, u8p, C.POINTER(C.c_uint32)]
    L.thash_h.argtypes = [P, u8p, u8p, u8p, C.POINTER(C.c_uint32)]
    L.hash_message.argtypes = [P, u8p, u8p, u8p, ull, u8p, ull]
    L.wots_pkgen.argtypes = [P, u8p, u8p, u8p, C.POINTER(C.c_uint32)]
    L.wots_sign.argtypes = [P, u8p, u8p, u8p, u8p, C.POINTER(C.c_uint32)]
    L.wots_pk_from_sig.argtypes = [P, u8p, u8p, u8p, u8p, C.POINTER(C.c_uint32)]
    L.gen_leaf_wots.argtypes = [P, u8p, u8p, u8p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    rng = random.Random(SEED)
    fams = {"SHA2_256": 0x01, "SHA2_512": 0x04, "SHA2_192": 0x0d, "SHAKE128_256": 0x07,
            "SHAKE256_512": 0x0a, "SHAKE256_256": 0x10, "SHAKE256_192": 0x13}
    res = {}
    for fam, oid in fams.items():
        p = ref.params(oid, 0)
        n = p.n
        rb = lambda k: bytes(rng.getrandbits(8) for _ in range(k))
        cases = []
        for _ in range(4):
            key, seed2, x, x2, inp = rb(n), rb(n), rb(n), rb(2 * n), rb(32)
            a = [rng.getrandbits(32) for _ in range(8)]
            o = buf(n)
            L.prf(C.byref(p), o, buf(inp), buf(key))
            prf = raw(o, n)
            L.prf_keygen(C.byref(p), o, buf(seed2 + inp), buf(key))
            prfk = raw(o, n)
            aa = addr(a)
            L.thash_f(C.byref(p), o, buf(x), buf(key), aa)
            tf = raw(o, n)
            aa = addr(a)
            L.thash_h(C.byref(p), o, buf(x2), buf(key), aa)
            th = raw(o, n)
            cases.append({"key": h(key), "seed2": h(seed2), "x": h(x), "x2": h(x2), "in32": h(inp),
                          "addr": a, "prf": h(prf), "prf_keygen": h(prfk), "thash_f": h(tf), "thash_h": h(th)})
        hm = []
        for mlen in (0, 1, 32, 55, 100, 1000, 4096):
            R, root, m = rb(n), rb(n), rb(mlen)
            idx = rng.getrandbits(40)
            pre = p.padding_len + 3 * n
            mb = buf(pre + mlen)
            C.memmove(C.byref(mb, pre), m, mlen) if mlen else None
            o = buf(n)
            L.hash_message(C.byref(p), o, buf(R), buf(root), idx, mb, mlen)
            hm.append({"R": h(R), "root": h(root), "idx": idx, "m": h(m), "out": h(raw(o, n))})
        ss, ps, msg = rb(n), rb(n), rb(n)
        a = [rng.getrandbits(32) for _ in range(3)] + [0, rng.getrandbits(10), 0, 0, 0]
        pk = buf(p.wots_sig_bytes)
        sig = buf(p.wots_sig_bytes)
        pk2 = buf(p.wots_sig_bytes)
        L.wots_pkgen(C.byref(p), pk, buf(ss), buf(ps), addr(a))
        L.wots_sign(C.byref(p), sig, buf(msg), buf(ss), buf(ps), addr(a))
        L.wots_pk_from_sig(C.byref(p), pk2, sig, buf(msg), buf(ps), addr(a))
        lt = [a[0], a[1], a[2], 1, a[4], 0, 0, 0]
        leaf = buf(n)
        L.gen_leaf_wots(C.byref(p), leaf, buf(ss), buf(ps), addr(lt), addr(a))
        wots = {"sk_seed": h(ss), "pub_seed": h(ps), "msg": h(msg), "ots_addr": a, "ltree_addr": lt,
                "pk": h(raw(pk, p.wots_sig_bytes)), "sig": h(raw(sig, p.wots_sig_bytes)),
                "pk_from_sig": h(raw(pk2, p.wots_sig_bytes)), "leaf": h(raw(leaf, n))}
        res[fam] = {"oid": oid, "cases": cases, "hash_message": hm, "wots": wots}
    write("hash_layer.json", res)


def plain_config1(ref):
    """XMSS-SHA2_10_256, plain core: seeded keypair, 16 signs of 32-byte
    messages in leaf order 0..15, their verification, then the index edge
    cases (idx 1022, 1023 = last-index quirk, then -2)."""
    ref.seed(SEED)
    oid = 0x01
    rc, pk, sk = ref.keypair(oid, 0)
    assert rc == 0
    rng = random.Random(SEED + 1)
    sk0 = sk
    sigs = []
    for i in range(16):
        m = bytes(rng.getrandbits(8) for _ in range(32))
        rc, sm, sk = ref.sign(sk, m, 0)
        assert rc == 0
        vr, mm = ref.sign_open(sm, pk, 0)
        assert vr == 0 and mm == m
        sigs.append({"msg": h(m), "sm": h(sm)})
    edge = []
    skx = bytearray(sk)
    skx[4:8] = (1022).to_bytes(4, "big")
    skx = bytes(skx)
    for _ in range(3):
        m = bytes(rng.getrandbits(8) for _ in range(32))
        before = skx
        rc, sm, skx = ref.sign(skx, m, 0)
        vr = ref.sign_open(sm, pk, 0)[0] if rc == 0 else None
        edge.append({"sk_in": h(before), "msg": h(m), "rc": rc, "sm": h(sm), "sk_out": h(skx), "verify": vr})
    write("plain_xmss_sha2_10_256.json", {"oid": oid, "rng_seed": SEED, "pk": h(pk), "sk": h(sk0),
                                         "sk_after": h(sk), "signs": sigs, "edge": edge})


SETS = ["XMSS-SHA2_10_192", "XMSS-SHA2_10_512", "XMSS-SHAKE_10_256", "XMSS-SHAKE_10_512",
        "XMSS-SHAKE256_10_256", "XMSS-SHAKE256_10_192",
        "XMSSMT-SHA2_20/2_256", "XMSSMT-SHA2_20/4_256", "XMSSMT-SHA2_40/8_256", "XMSSMT-SHA2_60/12_256",
        "XMSSMT-SHA2_20/4_512", "XMSSMT-SHAKE_20/4_256", "XMSSMT-SHAKE_20/4_512", "XMSSMT-SHA2_20/4_192",
        "XMSSMT-SHAKE256_20/4_256", "XMSSMT-SHAKE256_20/4_192"]


def plain_sets(ref):
    """Every (func, n) family, XMSS and XMSS^MT: seeded keypair, then two
    signs at random indices with log-uniform 32 B - 4 KB messages (SHA-256
    digests of sm and of the updated sk)."""
    res = {}
    rng = random.Random(SEED + 2)
    for k, name in enumerate(SETS):
        t = time.time()
        oid = ref.str_to_oid(name)
        mt = name.startswith("XMSSMT")
        p = ref.params(oid, mt)
        ref.seed(SEED + 100 + k)
        rc, pk, sk = ref.keypair(oid, mt)
        assert rc == 0
        signs = []
        for _ in range(2):
            idx = rng.randrange(0, (1 << p.full_height) - 1)
            skx = bytearray(sk)
            skx[4:4 + p.index_bytes] = idx.to_bytes(p.index_bytes, "big")
            m = bytes(rng.getrandbits(8) for _ in range(msg_len(rng)))
            rc, sm, sko = ref.sign(bytes(skx), m, mt)
            assert rc == 0
            assert ref.sign_open(sm, pk, mt)[0] == 0
            signs.append({"idx": idx, "msg": h(m), "sm": dg(sm), "sk_out": dg(sko)})
        res[name] = {"oid": oid, "mt": mt, "rng_seed": SEED + 100 + k, "pk": h(pk), "sk": h(sk), "signs": signs}
        print("  ", name, "%.1fs" % (time.time() - t), flush=True)
    write("plain_sets.json", res)


def plain_batch(ref):
    """Eight independent XMSS-SHA2_10_256 keys, each at a random leaf with a
    log-uniform 32 B - 4 KB message (the shape of the 4096-key batch)."""
    rng = random.Random(SEED + 3)
    items = []
    for k in range(8):
        ref.seed(SEED + 1000 + k)
        rc, pk, sk = ref.keypair(0x01, 0)
        idx = rng.randrange(0, 1024)
        skx = bytearray(sk)
        skx[4:8] = idx.to_bytes(4, "big")
        m = bytes(rng.getrandbits(8) for _ in range(msg_len(rng)))
        rc, sm, sko = ref.sign(bytes(skx), m, 0)
        assert ref.sign_open(sm, pk, 0)[0] == (0 if idx < 1023 else -1)
        items.append({"rng_seed": SEED + 1000 + k, "pk": h(pk), "sk_in": h(bytes(skx)), "idx": idx,
                      "msg": h(m), "rc": rc, "sm": h(sm), "sk_out": h(sko)})
    write("plain_batch.json", {"oid": 1, "items": items})


def bds_xmss(ref_fast, oid, nsigns, name, full=()):
    """BDS core (xmss_core_fast.c) on an all-zero sk buffer: the fixed-seed
    keypair, then nsigns signatures; digests of every sm and sk."""
    rng = random.Random(SEED + 4)
    p = ref_fast.params(oid, 0)
    rc, pk, sk = ref_fast.keypair(oid, 0, sk_init=bytes(4 + p.sk_bytes))
    sk0 = sk
    rows = []
    for i in range(nsigns):
        m = bytes(rng.getrandbits(8) for _ in range(32))
        rc, sm, sk = ref_fast.sign(sk, m, 0)
        row = {"rc": rc, "sm": dg(sm), "sk": dg(sk)}
        if i in full:
            row["sm_hex"] = h(sm)
            row["sk_hex"] = h(sk)
        rows.append(row)
    write(name, {"oid": oid, "msg_seed": SEED + 4, "pk": h(pk), "sk": h(sk0), "signs": rows})


def bds_xmssmt(ref_fast, name, nsigns, fname):
    rng = random.Random(SEED + 5)
    oid = ref_fast.str_to_oid(name)
    p = ref_fast.params(oid, 1)
    ref_fast.seed(SEED + 5)
    rc, pk, sk = ref_fast.keypair(oid, 1, sk_init=bytes(4 + p.sk_bytes))
    sk0 = sk
    rows = []
    for i in range(nsigns):
        m = bytes(rng.getrandbits(8) for _ in range(32))
        rc, sm, sk = ref_fast.sign(sk, m, 1)
        rows.append({"rc": rc, "sm": dg(sm), "sk": dg(sk)})
    write(fname, {"name": name, "oid": oid, "rng_seed": SEED + 5, "msg_seed": SEED + 5,
                  "pk": h(pk), "sk": h(sk0), "signs": rows})


def main():
    big = "--big" in sys.argv
    only = [a for a in sys.argv[1:] if not a.startswith("--")]
    ref, ref_fast = load_ref(0), load_ref(1)
    assert ref and ref_fast, "build oracle/_ref first: make -C oracle ref"
    jobs = {
        "kat": kat_vectors,
        "gpukeys": ref_gpu_keys,
        "hash": lambda: hash_layer(ref),
        "config1": lambda: plain_config1(ref),
        "sets": lambda: plain_sets(ref),
        "batch": lambda: plain_batch(ref),
        "bds10": lambda: bds_xmss(ref_fast, 0x01, 1025, "bds_xmss_sha2_10_256.json", full=(0, 1, 1022, 1023)),
        "bdsmt": lambda: bds_xmssmt(ref_fast, "XMSSMT-SHA2_20/4_256", 1100, "bds_xmssmt_sha2_20_4_256.json"),
    }
    if big:
        jobs["bds16"] = lambda: bds_xmss(ref_fast, 0x02, 16, "bds_xmss_sha2_16_256.json", full=(0,))
    for k, fn in jobs.items():
        if only and k not in only:
            continue
        t = time.time()
        fn()
        print(k, "%.1fs" % (time.time() - t), flush=True)


if __name__ == "__main__":
    main()
