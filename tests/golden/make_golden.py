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
import workloads as W  # noqa: E402

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


# ---- fixtures of the batch, mixed-set, serving and whole-life workloads ----
# These run many reference signatures (about a second each for a plain-core
# signature of one 1024-leaf tree), so they fan out over worker processes;
# each worker loads its own handle of the reference build.

_REF = {}


def _ref(core):
    if core not in _REF:
        _REF[core] = load_ref(core)
    return _REF[core]


def _pool_map(fn, args):
    import multiprocessing as mp
    with mp.get_context("fork").Pool(min(8, os.cpu_count() or 1)) as pool:
        return pool.map(fn, args, chunksize=1)


def _seed_keypair_task(a):
    oid, mt, seed3 = a
    ref = _ref(0)
    rc, pk, sk = ref.seed_keypair(oid, mt, seed3)
    assert rc == 0
    return pk


def _plain_sign_task(a):
    sk, m, mt, pk = a
    ref = _ref(0)
    rc, sm, sko = ref.sign(sk, m, mt)
    row = {"rc": rc, "sk_out": dg(sko)}
    if rc == 0:
        row["sm"] = dg(sm)
        row["verify"] = ref.sign_open(sm, pk, mt)[0]
    return row


def plain_batch_large(count=320):
    """Independent XMSS-SHA2_10_256 keys from seeds (one keypair per key via
    xmssmt_core_seed_keypair), each signing one log-uniform 32 B - 4 KB
    message at a random index in [0, 1024] (the first four keys at 1023, 1024,
    1022 and 0).  Shape of the 4096-key batch, sized for the reference's CPU
    time; it reaches the throughput kernels of a GPU signing batch."""
    oid, mt = 0x01, 0
    ref = _ref(0)
    p = ref.params(oid, mt)
    kseed, mseed = W.SEED + 600, W.SEED + 601
    seeds = [W.key_seed(kseed, k, p.n) for k in range(count)]
    roots = _pool_map(_seed_keypair_task, [(oid, mt, s) for s in seeds])
    roots = [r[:p.n] for r in roots]
    idx = W.batch_indices(kseed, count, (1 << p.full_height) - 1)
    tasks = []
    for k in range(count):
        sk = W.plain_sk(oid, p, idx[k], seeds[k], roots[k])
        pk = W.with_oid(oid, roots[k] + seeds[k][2 * p.n:])
        tasks.append((sk, W.message(mseed, k), mt, pk))
    rows = _pool_map(_plain_sign_task, tasks)
    items = [dict(root=h(roots[k]), idx=idx[k], **rows[k]) for k in range(count)]
    write("plain_batch_large.json", {"oid": oid, "mt": mt, "count": count, "key_seed": kseed, "msg_seed": mseed,
                                     "items": items})


MIXED4 = [("XMSSMT-SHA2_20/2_256", 1), ("XMSS-SHA2_10_512", 0), ("XMSS-SHAKE_10_256", 0)]


def mixed_sets(requests=260, keys=8):
    """The three parameter sets of the mixed workload: per set `keys` seeded
    keys and `requests` plain signatures, request i by key i % keys at a
    uniform index in [0, 2^h - 2] with a log-uniform 32 B - 4 KB message."""
    ref = _ref(0)
    res = {}
    for s, (name, mt) in enumerate(MIXED4):
        t = time.time()
        oid = ref.str_to_oid(name)
        p = ref.params(oid, mt)
        kseed, mseed = W.SEED + 700 + 10 * s, W.SEED + 701 + 10 * s
        seeds = [W.key_seed(kseed, k, p.n) for k in range(keys)]
        roots = [r[:p.n] for r in _pool_map(_seed_keypair_task, [(oid, mt, x) for x in seeds])]
        rng = random.Random(kseed)
        idx = [rng.randrange(0, (1 << p.full_height) - 1) for _ in range(requests)]
        tasks = []
        for i in range(requests):
            k = i % keys
            sk = W.plain_sk(oid, p, idx[i], seeds[k], roots[k])
            pk = W.with_oid(oid, roots[k] + seeds[k][2 * p.n:])
            tasks.append((sk, W.message(mseed, i), mt, pk))
        rows = _pool_map(_plain_sign_task, tasks)
        res[name] = {"oid": oid, "mt": mt, "keys": keys, "key_seed": kseed, "msg_seed": mseed,
                     "roots": [h(r) for r in roots], "items": [dict(idx=idx[i], **rows[i]) for i in range(requests)]}
        print("  ", name, "%.1fs" % (time.time() - t), flush=True)
    write("mixed_sets.json", res)


def _bds_seq_task(a):
    """BDS signatures of one key in order: digests of every sm and sk."""
    sk, mt, mseed, count = a
    ref = _ref(1)
    rows = []
    for i in range(count):
        rc, sm, sk = ref.sign(sk, W.message(mseed, i), mt)
        rows.append({"rc": rc, "sm": dg(sm), "sk": dg(sk)})
    return rows


SERVE5 = [("XMSS-SHA2_16_256", 0), ("XMSSMT-SHA2_20/4_256", 1), ("XMSSMT-SHA2_40/4_256", 1)]


def serve_sets(keys=3, count=24, long_count=1040):
    """The three BDS parameter sets of the serving workload: per set `keys`
    keys (XMSS: the BDS keypair's fixed seeds, so one key with separate
    message streams; XMSS^MT: seeded keypairs), each signing `count`
    log-uniform 32 B - 4 KB messages in order, the last XMSS^MT key of each
    set `long_count` (past its first layer-0 subtree switch at 1024)."""
    ref_fast = _ref(1)
    res = {}
    for s, (name, mt) in enumerate(SERVE5):
        t = time.time()
        oid = ref_fast.str_to_oid(name)
        p = ref_fast.params(oid, mt)
        tasks, keyrows = [], []
        pk0 = sk0 = None
        for k in range(keys):
            rseed = W.SEED + 800 + 10 * s + k
            if mt or pk0 is None:
                ref_fast.seed(rseed)
                rc, pk0, sk0 = ref_fast.keypair(oid, mt, sk_init=bytes(4 + p.sk_bytes))
                assert rc == 0
            mseed = W.SEED + 900 + 10 * s + k
            cnt = long_count if (mt and k == keys - 1) else count
            tasks.append((sk0, mt, mseed, cnt))
            keyrows.append({"rng_seed": rseed, "msg_seed": mseed, "pk": h(pk0), "sk": dg(sk0), "count": cnt})
        rows = _pool_map(_bds_seq_task, tasks)
        for kr, r in zip(keyrows, rows):
            kr["signs"] = r
        res[name] = {"oid": oid, "mt": mt, "keys": keyrows}
        print("  ", name, "%.1fs" % (time.time() - t), flush=True)
    write("serve_sets.json", res)


def nonzero_sk_keypairs():
    """BDS keypairs written into caller buffers that are not zero: the
    reference leaves some state fields unwritten (xmss_core_fast.c:656-701),
    so they keep the caller's bytes; then four signatures."""
    ref_fast = _ref(1)
    res = {}
    for name, mt in (("XMSS-SHA2_10_256", 0), ("XMSSMT-SHA2_20/4_256", 1)):
        oid = ref_fast.str_to_oid(name)
        p = ref_fast.params(oid, mt)
        init = random.Random(W.SEED + 950 + mt).randbytes(4 + p.sk_bytes)
        ref_fast.seed(W.SEED + 951 + mt)
        rc, pk, sk = ref_fast.keypair(oid, mt, sk_init=init)
        assert rc == 0
        sk0 = sk
        rows = []
        for i in range(4):
            m = W.fixed_message(W.SEED + 952, i)
            rc, sm, sk = ref_fast.sign(sk, m, mt)
            rows.append({"rc": rc, "sm": dg(sm), "sk": dg(sk)})
        res[name] = {"oid": oid, "mt": mt, "rng_seed": W.SEED + 951 + mt, "init_seed": W.SEED + 950 + mt,
                     "pk": h(pk), "sk": h(sk0), "signs": rows}
    write("nonzero_sk_keypairs.json", res)


def whole_life_samples():
    """XMSSMT-SHA2_20/2_256 (cuxmss/test/xmss_max_signatures.c:35-102): a
    seeded key and one random 32-byte message, as that test draws them
    (randombytes(m) before the keypair); reference signatures at sampled
    indices, the last index (wiped key, returns 0) and the two -2 calls after
    it, with the sk after each.  Also the plain-core signature of the last
    index of the XMSS-SHA2_10_256 BDS life fixture's key with its message 1023
    (the BDS core signs that index differently)."""
    ref = _ref(0)
    name, mt = "XMSSMT-SHA2_20/2_256", 1
    oid = ref.str_to_oid(name)
    p = ref.params(oid, mt)
    ref.seed(W.SEED + 960)
    m = buf(32)
    ref.lib.randombytes.argtypes = [u8p, ull]
    ref.lib.randombytes(m, 32)
    msg = raw(m, 32)
    rc, pk, sk = ref.keypair(oid, mt)
    assert rc == 0
    top = (1 << p.full_height) - 1
    samples = [0, 1, 2, 1023, 1024, 1025, 524287, 524288, top - 1]
    tasks = []
    for i in samples:
        skx = bytearray(sk)
        skx[4:4 + p.index_bytes] = i.to_bytes(p.index_bytes, "big")
        tasks.append((bytes(skx), msg, mt, pk))
    rows = _pool_map(_plain_sign_task, tasks)
    skx = bytearray(sk)
    skx[4:4 + p.index_bytes] = top.to_bytes(p.index_bytes, "big")
    tail = []
    cur = bytes(skx)
    for _ in range(3):
        rc, sm, cur = ref.sign(cur, msg, mt)
        tail.append({"rc": rc, "sm": dg(sm) if rc == 0 else None, "sk_out": h(cur)})
    # the BDS life fixture's key in plain-core form at its last index
    bds = json.load(open(os.path.join(HERE, "bds_xmss_sha2_10_256.json")))
    p10 = ref.params(0x01, 0)
    gen = W.bds10_messages()
    msgs = [next(gen) for _ in range(1024)]
    pk10 = bytes.fromhex(bds["pk"])
    seed3 = b"\x02" * 64 + b"\x04" * 32
    sk10 = W.plain_sk(0x01, p10, 1023, seed3, pk10[4:4 + 32])
    rc, sm, sko = ref.sign(sk10, msgs[1023], 0)
    write("whole_life.json", {"name": name, "oid": oid, "rng_seed": W.SEED + 960, "msg": h(msg), "pk": h(pk),
                              "sk": h(sk), "samples": [dict(idx=i, **r) for i, r in zip(samples, rows)],
                              "last": tail,
                              "bds10_last_plain": {"rc": rc, "sm": dg(sm), "sk_out": h(sko)}})


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
        "bdsmt40": lambda: bds_xmssmt(ref_fast, "XMSSMT-SHA2_40/4_256", 1100, "bds_xmssmt_sha2_40_4_256.json"),
        "batch320": plain_batch_large,
        "mixed": mixed_sets,
        "serve": serve_sets,
        "nonzero": nonzero_sk_keypairs,
        "life": whole_life_samples,
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
