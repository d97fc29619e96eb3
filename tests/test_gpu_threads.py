"""The C API from several threads at once: each thread signs with its own key
and verifies signatures of the others' keys, on both cores; every signature,
secret key and verdict equals the same calls made one after another
(distinct keys are safe from any thread, as with the reference)."""
import random
import threading

import pytest

from common import product

pytestmark = pytest.mark.gpu

SETS = {0: ("XMSS-SHA2_10_256", 0, 0x01), 1: ("XMSSMT-SHA2_20/4_256", 1, 0x02)}


@pytest.mark.parametrize("core", [0, 1])
def test_threads_sign_and_verify_like_sequential_calls(core):
    lib = product(core)
    name, mt, oid = SETS[core]
    p = lib.params(oid, mt)
    nkeys, nsig = 6, 5
    keys = []
    for k in range(nkeys):
        lib.lib.xmss_mi355x_rng_seed(9000 + 17 * k)
        rc, pk, sk = lib.keypair(oid, mt, sk_init=bytes(4 + p.sk_bytes))
        assert rc == 0
        keys.append((pk, sk))
    if core == 0:  # distinct starting indices, away from the first leaf
        keys = [(pk, sk[:4] + (37 * k + 3).to_bytes(p.index_bytes, "big") + sk[4 + p.index_bytes:])
                for k, (pk, sk) in enumerate(keys)]
    rng = random.Random(core)
    msgs = [[bytes(rng.getrandbits(8) for _ in range(rng.randrange(1, 600))) for _ in range(nsig)]
            for _ in range(nkeys)]

    # one after another
    want_sm, want_sk = [], []
    for k in range(nkeys):
        sk, sms = keys[k][1], []
        for m in msgs[k]:
            rc, sm, sk = lib.sign(sk, m, mt)
            assert rc == 0
            sms.append(sm)
        want_sm.append(sms)
        want_sk.append(sk)

    # all at once: thread k signs with key k, then verifies key (k+1)'s
    # expected signatures and a corrupted one
    got_sm = [None] * nkeys
    got_sk = [None] * nkeys
    verdicts = [None] * nkeys
    errors = []

    def worker(k):
        try:
            sk, sms = keys[k][1], []
            for m in msgs[k]:
                rc, sm, sk = lib.sign(sk, m, mt)
                assert rc == 0
                sms.append(sm)
            got_sm[k], got_sk[k] = sms, sk
            o = (k + 1) % nkeys
            v = [lib.sign_open(sm, keys[o][0], mt)[0] for sm in want_sm[o]]
            bad = bytearray(want_sm[o][0])
            bad[p.index_bytes + 1] ^= 1
            v.append(lib.sign_open(bytes(bad), keys[o][0], mt)[0])
            verdicts[k] = v
        except Exception as e:  # reported below
            errors.append((k, e))

    th = [threading.Thread(target=worker, args=(k,)) for k in range(nkeys)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for k in range(nkeys):
        assert got_sm[k] == want_sm[k], "key %d signatures" % k
        assert got_sk[k] == want_sk[k], "key %d sk" % k
        assert verdicts[k] == [0] * nsig + [-1], "key %d verdicts" % k
