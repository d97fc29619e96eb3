"""Deterministic synthetic inputs shared by the fixture generator
(tests/golden/make_golden.py) and the tests that replay its fixtures.

Every input is a pure function of integer seeds, so a fixture stores seeds,
indices and digests instead of the inputs themselves: the test regenerates
the same key seeds and message bytes and compares what the product returns
with what the reference returned for them.
"""
import random

SEED = 0x584D5353  # "XMSS"


def key_seed(seed, k, n):
    """3n seed bytes (SK_SEED || SK_PRF || PUB_SEED) of key k."""
    return random.Random(seed * 7919 + k).randbytes(3 * n)


def message(seed, i):
    """Message i of a stream: length log-uniform over 32 B .. 4 KB
    (L = floor(32 * 128^u), u uniform in [0, 1)), uniform random bytes."""
    rng = random.Random(seed * 1000003 + i)
    mlen = int(32 * 128 ** rng.random())
    return rng.randbytes(mlen)


def fixed_message(seed, i, mlen=32):
    """Message i of a stream of fixed-length messages."""
    return random.Random(seed * 1000003 + i).randbytes(mlen)


def batch_indices(seed, count, top):
    """One signature index per key, uniform in [0, top]; the first four keys
    take the edge cases top (the last index: wiped key, still signs), top + 1
    (exhausted: -2), top - 1 and 0."""
    rng = random.Random(seed * 31 + 7)
    idx = [rng.randrange(0, top + 1) for _ in range(count)]
    for k, v in enumerate((top, top + 1, top - 1, 0)):
        if k < count:
            idx[k] = v
    return idx


def bds10_messages():
    """The 32-byte message stream of the XMSS-SHA2_10_256 BDS life fixture
    (bds_xmss_sha2_10_256.json): getrandbits(8) x 32 per message."""
    rng = random.Random(SEED + 4)
    while True:
        yield bytes(rng.getrandbits(8) for _ in range(32))


def with_oid(oid, b):
    return oid.to_bytes(4, "big") + bytes(b)


def plain_sk(oid, p, idx, seed3, root):
    """OID-prefixed plain-core sk: idx || SK_SEED || SK_PRF || root || PUB_SEED
    (cuxmss/xmss_core.c:146-158)."""
    n = p.n
    body = idx.to_bytes(p.index_bytes, "big") + seed3[:2 * n] + root + seed3[2 * n:3 * n]
    return with_oid(oid, body)
