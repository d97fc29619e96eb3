"""Message lengths at every padding boundary of each hash family's H_msg
(hash_message hashes a 4n-byte prefix and the message: SHA-256 blocks of
64 bytes, SHA-512 of 128, SHAKE128 / SHAKE256 rates of 168 / 136 bytes),
the empty message and a 1 MiB one, on both cores: every signature verifies
under the oracle (so H_msg agrees with the reference's), the GPU's
sign_open returns the message, and a flipped message byte is rejected by
both."""
import random

import pytest

from common import oracle, product

pytestmark = pytest.mark.gpu

SETS = ["XMSS-SHA2_10_256", "XMSS-SHA2_10_192", "XMSS-SHA2_10_512", "XMSS-SHAKE_10_256",
        "XMSS-SHAKE256_10_256", "XMSS-SHAKE256_10_192"]


def _lengths(n):
    pre = 4 * n if n != 24 else 4 + 3 * 24  # padding_len + 3n (params.c: padding_len = n, 4 for n = 24)
    out = {0, 1, 1 << 20}
    for block, tail in ((64, 9), (128, 17), (168, 1), (136, 1)):
        for k in (1, 2, 3, 17):
            edge = k * block - tail - pre  # the last length whose padding fits in k blocks
            for d in (-1, 0, 1):
                if edge + d >= 0:
                    out.add(edge + d)
    return sorted(out)


@pytest.mark.parametrize("core", [0, 1])
@pytest.mark.parametrize("name", SETS)
def test_message_lengths_verify_under_oracle(name, core):
    o = oracle()
    lib = product(core)
    oid = lib.str_to_oid(name)
    p = lib.params(oid, 0)
    lib.lib.xmss_mi355x_rng_seed(oid + 100 * core)
    rc, pk, sk = lib.keypair(oid, 0, sk_init=bytes(4 + p.sk_bytes))
    assert rc == 0
    rng = random.Random(oid * 3 + core)
    for mlen in _lengths(p.n):
        m = bytes(rng.getrandbits(8) for _ in range(mlen)) if mlen < 4096 else rng.randbytes(mlen)
        rc, sm, sk = lib.sign(sk, m, 0)
        assert rc == 0, (name, mlen)
        assert o.sign_open(sm, pk, 0) == (0, m), (name, core, mlen)
        assert lib.sign_open(sm, pk, 0) == (0, m), (name, core, mlen)
        if mlen:
            bad = bytearray(sm)
            bad[p.sig_bytes + rng.randrange(mlen)] ^= 0x10
            assert lib.sign_open(bytes(bad), pk, 0)[0] == -1 == o.sign_open(bytes(bad), pk, 0)[0], (name, mlen)
