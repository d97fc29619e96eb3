"""Serving loop (continuous batching of BDS requests across keys): every
signature and every secret-key update must be byte-identical to the same
request run alone through xmss[mt]_sign, whatever else shares its launches;
verifications accept exactly the valid signatures."""
import ctypes as C
import random

import pytest

from common import product
from ffi import buf, raw

pytestmark = pytest.mark.gpu

KEYS = [("XMSS-SHA2_10_256", 0, 0x01), ("XMSS-SHA2_10_256", 0, 0x01),
        ("XMSSMT-SHA2_20/4_256", 1, 0x02), ("XMSSMT-SHA2_20/4_256", 1, 0x02),
        ("XMSSMT-SHA2_40/8_256", 1, 0x05), ("XMSSMT-SHA2_20/2_192", 1, 0x21)]


def _server_api(L):
    L.xmss_mi355x_server_create.restype = C.c_void_p
    L.xmss_mi355x_server_create.argtypes = [C.c_size_t, C.c_int]
    L.xmss_mi355x_server_add_key.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    for nm in ("xmss_mi355x_server_sign", "xmss_mi355x_server_verify"):
        getattr(L, nm).restype = C.c_uint64
        getattr(L, nm).argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_ulonglong]
    L.xmss_mi355x_server_result.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.POINTER(C.c_ulonglong),
                                            C.POINTER(C.c_double)]
    L.xmss_mi355x_server_get_sk.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    L.xmss_mi355x_server_stats.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                           C.POINTER(C.c_size_t), C.POINTER(C.c_uint64)]
    L.xmss_mi355x_server_destroy.argtypes = [C.c_void_p]
    L.xmss_mi355x_server_memory.argtypes = [C.c_void_p] + [C.POINTER(C.c_size_t)] * 4
    L.xmss_mi355x_device_bytes.restype = C.c_size_t


def _memory(L, srv):
    b, r, u, p = C.c_size_t(), C.c_size_t(), C.c_size_t(), C.c_size_t()
    L.xmss_mi355x_server_memory(srv, C.byref(b), C.byref(r), C.byref(u), C.byref(p))
    return b.value, r.value, u.value, p.value


def _tight_budget(lib, keys, msgs):
    """The keys' resident bytes plus 1.5 x the largest working set of a
    single request of theirs (each key's first messages signed, and one
    signature verified, alone on a probe server), so that two large requests
    never fit at once."""
    L = lib.lib
    srv = L.xmss_mi355x_server_create(1 << 30, 4)
    assert srv
    for (mt, oid, p, pk, sk) in keys:
        assert L.xmss_mi355x_server_add_key(srv, mt, buf(sk), buf(pk)) >= 0
    _, resident, _, _ = _memory(L, srv)
    most = 0
    for k, (mt, oid, p, pk, sk) in enumerate(keys):
        for m in msgs[k][:2]:
            rid = L.xmss_mi355x_server_sign(srv, k, m, len(m))
            out, olen = buf(p.sig_bytes + len(m)), C.c_ulonglong(0)
            assert L.xmss_mi355x_server_result(srv, rid, out, C.byref(olen), None) == 0
            _, _, _, peak = _memory(L, srv)
            most = max(most, peak - resident)
        sm = raw(out, olen.value)
        rid = L.xmss_mi355x_server_verify(srv, k, sm, len(sm))
        assert L.xmss_mi355x_server_result(srv, rid, buf(len(sm)), None, None) == 0
        _, _, _, peak = _memory(L, srv)
        most = max(most, peak - resident)
    L.xmss_mi355x_server_destroy(srv)
    return resident + most + most // 2


def _keys(lib):
    keys = []
    for i, (name, mt, oid) in enumerate(KEYS):
        lib.lib.xmss_mi355x_rng_seed(1000 + i)
        p = lib.params(oid, mt)
        rc, pk, sk = lib.keypair(oid, mt, sk_init=bytes(4 + p.sk_bytes))
        assert rc == 0
        keys.append((mt, oid, p, pk, sk))
    return keys


def _alone(lib, keys, msgs):
    """Each key's messages signed in order through the C API."""
    out = []
    for (mt, oid, p, pk, sk), ms in zip(keys, msgs):
        sms = []
        for m in ms:
            rc, sm, sk = lib.sign(sk, m, mt)
            assert rc == 0
            sms.append(sm)
        out.append((sms, sk))
    return out


# the tight budget holds the six keys' resident state and one and a half
# times the largest single request: requests wait for each other's memory
@pytest.mark.parametrize("tight", [False, True])
def test_serve_mix_equals_alone(tight):
    lib = product(1)
    L = lib.lib
    _server_api(L)
    keys = _keys(lib)
    rng = random.Random(77)
    nsig = 40  # crosses layer-0 subtree switches of the th = 5 keys (idx 31 -> 32)
    msgs = [[bytes(rng.getrandbits(8) for _ in range(int(32 * 128 ** rng.random()))) for _ in range(nsig)]
            for _ in keys]
    want = _alone(lib, keys, msgs)
    budget = _tight_budget(lib, keys, msgs) if tight else 1 << 30

    srv = L.xmss_mi355x_server_create(budget, 64)
    assert srv
    held = L.xmss_mi355x_device_bytes()  # the library's device memory, the arena included
    ids = [L.xmss_mi355x_server_add_key(srv, mt, buf(sk), buf(pk)) for (mt, oid, p, pk, sk) in keys]
    assert ids == list(range(len(keys)))
    # interleave: round-robin over keys, with verifications of signatures
    # already produced mixed in
    reqs = []
    order = [(k, j) for j in range(nsig) for k in range(len(keys))]
    for k, j in order:
        reqs.append(("s", k, j, L.xmss_mi355x_server_sign(srv, k, msgs[k][j], len(msgs[k][j]))))
        if j % 5 == 4:
            sm = want[k][0][j - 3]
            reqs.append(("v", k, j - 3, L.xmss_mi355x_server_verify(srv, k, sm, len(sm))))
    bad = bytearray(want[2][0][7])
    bad[-1] ^= 1
    reqs.append(("bad", 2, 7, L.xmss_mi355x_server_verify(srv, 2, bytes(bad), len(bad))))
    for kind, k, j, rid in reqs:
        p = keys[k][2]
        cap = p.sig_bytes + 5000
        out = buf(cap)
        olen = C.c_ulonglong(0)
        lat = C.c_double(0)
        rc = L.xmss_mi355x_server_result(srv, rid, out, C.byref(olen), C.byref(lat))
        if kind == "s":
            assert rc == 0
            assert raw(out, olen.value) == want[k][0][j], "key %d sign %d" % (k, j)
        elif kind == "v":
            assert rc == 0 and raw(out, olen.value) == msgs[k][j]
        else:
            assert rc == -1 and olen.value == 0
    for k, (mt, oid, p, pk, sk) in enumerate(keys):
        skb = buf(4 + p.sk_bytes)
        assert L.xmss_mi355x_server_get_sk(srv, k, skb) == 0
        assert raw(skb, 4 + p.sk_bytes) == want[k][1], "key %d final sk" % k
    launches, calls, peak, waits = C.c_uint64(), C.c_uint64(), C.c_size_t(), C.c_uint64()
    L.xmss_mi355x_server_stats(srv, C.byref(launches), C.byref(calls), C.byref(peak), C.byref(waits))
    size, resident, in_use, peak2 = _memory(L, srv)
    # no device memory outside the arena grew while the server ran
    assert L.xmss_mi355x_device_bytes() == held
    L.xmss_mi355x_server_destroy(srv)
    assert calls.value > launches.value  # calls of different keys shared launches
    assert size >= budget and size - budget < 256 and peak.value == peak2 <= size and in_use == resident
    if tight:
        assert waits.value > 0


def test_serve_key_to_its_end():
    """A resident XMSS-SHA2_10_256 key signed through the server across its
    last indices: 1020-1022 take the device path, 1023 (the wiped key, still
    returning 0) and the exhausted calls after it (-2) take the host path,
    which re-uploads the resident state; signatures, return codes and the
    final sk equal the key signing alone."""
    lib = product(1)
    L = lib.lib
    _server_api(L)
    lib.lib.xmss_mi355x_rng_seed(4242)
    p = lib.params(0x01, 0)
    rc, pk, sk = lib.keypair(0x01, 0, sk_init=bytes(4 + p.sk_bytes))
    assert rc == 0
    for i in range(1020):
        rc, _, sk = lib.sign(sk, b"%d" % i, 0)
        assert rc == 0
    rng = random.Random(5)
    msgs = [bytes(rng.getrandbits(8) for _ in range(40 + 7 * j)) for j in range(6)]
    want, ska = [], sk
    for m in msgs:
        rc, sm, ska = lib.sign(ska, m, 0)
        want.append((rc, sm))
    assert [w[0] for w in want] == [0, 0, 0, 0, -2, -2]

    srv = L.xmss_mi355x_server_create(1 << 24, 8)
    assert srv
    k = L.xmss_mi355x_server_add_key(srv, 0, buf(sk), buf(pk))
    assert k == 0
    for j, m in enumerate(msgs):
        rid = L.xmss_mi355x_server_sign(srv, k, m, len(m))
        out = buf(p.sig_bytes + len(m))
        olen = C.c_ulonglong(0)
        rc = L.xmss_mi355x_server_result(srv, rid, out, C.byref(olen), None)
        assert rc == want[j][0], "sign %d" % j
        if rc == 0:
            assert raw(out, olen.value) == want[j][1], "sign %d" % j
        if j == 2:  # device path so far: the mirror matches the key signed alone
            skb = buf(4 + p.sk_bytes)
            assert L.xmss_mi355x_server_get_sk(srv, k, skb) == 0
            _, _, sk3 = lib.sign(sk, msgs[0], 0)
            _, _, sk3 = lib.sign(sk3, msgs[1], 0)
            _, _, sk3 = lib.sign(sk3, msgs[2], 0)
            assert raw(skb, 4 + p.sk_bytes) == sk3
    skb = buf(4 + p.sk_bytes)
    assert L.xmss_mi355x_server_get_sk(srv, k, skb) == 0
    assert raw(skb, 4 + p.sk_bytes) == ska
    L.xmss_mi355x_server_destroy(srv)


def test_serve_budget_counts_resident_keys():
    """Adding keys reserves their resident state (key contexts and sk bytes)
    in the server's arena before any request runs; a key the budget cannot
    hold is refused."""
    lib = product(1)
    L = lib.lib
    _server_api(L)
    keys = _keys(lib)
    srv = L.xmss_mi355x_server_create(1 << 30, 4)
    assert srv
    res = []
    for (mt, oid, p, pk, sk) in keys:
        assert L.xmss_mi355x_server_add_key(srv, mt, buf(sk), buf(pk)) >= 0
        res.append(_memory(L, srv)[1])
    launches, calls, peak, waits = C.c_uint64(), C.c_uint64(), C.c_size_t(), C.c_uint64()
    L.xmss_mi355x_server_stats(srv, C.byref(launches), C.byref(calls), C.byref(peak), C.byref(waits))
    L.xmss_mi355x_server_destroy(srv)
    assert res[-1] >= sum(p.sk_bytes for (_, _, p, _, _) in keys)
    assert calls.value == 8  # a key context each, and the zero-seed context of each of the two families
    # a budget just below the six keys' resident bytes refuses the last key
    srv = L.xmss_mi355x_server_create(res[-1] - 256, 4)
    assert srv
    got = [L.xmss_mi355x_server_add_key(srv, mt, buf(sk), buf(pk)) for (mt, oid, p, pk, sk) in keys]
    L.xmss_mi355x_server_destroy(srv)
    assert got == [0, 1, 2, 3, 4, -1]


def test_serve_config5_sets_equal_reference():
    """The serving workload's three parameter sets -- XMSS-SHA2_16_256,
    XMSSMT-SHA2_20/4_256, XMSSMT-SHA2_40/4_256 -- three keys each, signing
    log-uniform 32 B - 4 KB messages in order (one key of each XMSS^MT set
    past its first layer-0 subtree switch), interleaved with verifications.
    Every served signature and every final sk equals the reference's BDS core
    (tests/golden/serve_sets.json, written by xmss_core_fast.c).  The last
    key of each set is added while the others' requests are in flight."""
    import workloads as W
    from common import dg, golden
    g = golden("serve_sets.json")
    lib = product(1)
    L = lib.lib
    _server_api(L)
    keys = []  # (mt, p, pk, sk, row)
    for name, s in g.items():
        mt, oid = s["mt"], s["oid"]
        p = lib.params(oid, mt)
        for row in s["keys"]:
            lib.lib.xmss_mi355x_rng_seed(row["rng_seed"])
            rc, pk, sk = lib.keypair(oid, mt, sk_init=bytes(4 + p.sk_bytes))
            assert rc == 0
            assert pk.hex() == row["pk"] and dg(sk) == row["sk"], name
            keys.append((mt, p, pk, sk, row))
    srv = L.xmss_mi355x_server_create(1 << 30, 64)
    assert srv
    late = [i for i in range(len(keys)) if i % 3 == 2]
    kid = {}
    for i, (mt, p, pk, sk, row) in enumerate(keys):
        if i not in late:
            kid[i] = L.xmss_mi355x_server_add_key(srv, mt, buf(sk), buf(pk))
            assert kid[i] >= 0
    reqs = []

    def submit(i, j):
        m = W.message(keys[i][4]["msg_seed"], j)
        reqs.append((i, j, m, L.xmss_mi355x_server_sign(srv, kid[i], m, len(m))))

    early = [i for i in kid]
    for j in range(8):
        for i in early:
            submit(i, j)
    for i in late:  # added while the requests above run
        mt, p, pk, sk, row = keys[i]
        kid[i] = L.xmss_mi355x_server_add_key(srv, mt, buf(sk), buf(pk))
        assert kid[i] >= 0
    for i in late:
        for j in range(8):
            submit(i, j)
    jmax = max(k[4]["count"] for k in keys)
    for j in range(8, jmax):
        for i in range(len(keys)):
            if j < keys[i][4]["count"]:
                submit(i, j)
    done = {}
    for i, j, m, rid in reqs:
        p = keys[i][1]
        out = buf(p.sig_bytes + len(m))
        olen = C.c_ulonglong(0)
        rc = L.xmss_mi355x_server_result(srv, rid, out, C.byref(olen), None)
        row = keys[i][4]["signs"][j]
        assert rc == row["rc"] == 0, "key %d sign %d" % (i, j)
        sm = raw(out, olen.value)
        assert dg(sm) == row["sm"], "key %d sign %d" % (i, j)
        done[(i, j)] = sm
    # verifications of served signatures (and one corrupted), as requests
    vreqs = []
    for i in range(len(keys)):
        for j in (0, 5, keys[i][4]["count"] - 1):
            sm = done[(i, j)]
            vreqs.append((i, j, True, L.xmss_mi355x_server_verify(srv, kid[i], sm, len(sm))))
        bad = bytearray(done[(i, 3)])
        bad[keys[i][1].index_bytes + 9] ^= 2
        vreqs.append((i, 3, False, L.xmss_mi355x_server_verify(srv, kid[i], bytes(bad), len(bad))))
    for i, j, ok, rid in vreqs:
        out = buf(len(done[(i, j)]))
        olen = C.c_ulonglong(0)
        rc = L.xmss_mi355x_server_result(srv, rid, out, C.byref(olen), None)
        if ok:
            assert rc == 0 and raw(out, olen.value) == W.message(keys[i][4]["msg_seed"], j)
        else:
            assert rc == -1 and olen.value == 0
    for i, (mt, p, pk, sk, row) in enumerate(keys):
        skb = buf(4 + p.sk_bytes)
        assert L.xmss_mi355x_server_get_sk(srv, kid[i], skb) == 0
        assert dg(raw(skb, 4 + p.sk_bytes)) == row["signs"][-1]["sk"], "key %d final sk" % i
    L.xmss_mi355x_server_destroy(srv)
