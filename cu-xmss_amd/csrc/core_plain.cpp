// core_plain.cpp -- the plain core (drop-in for cuxmss/xmss_core.c): the
// secret key is idx || SK_SEED || SK_PRF || root || PUB_SEED and every
// signature rebuilds its subtrees, here as one GPU tree build per call.
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/randombytes.h"
#include "../../include/xmss_core.h"
#include "../../include/xmss_mi355x.h"
#include "core_common.h"

using namespace xm;

extern "C" {

// xmss_core.c:93-96
unsigned long long xmss_xmssmt_core_sk_bytes(const xmss_params* params) {
    return params->index_bytes + 4 * params->n;
}

// xmss_core.c:134-161: only the top subtree (layer d-1, tree 0) is built.
int xmssmt_core_seed_keypair(const xmss_params* params, unsigned char* pk, unsigned char* sk,
                             unsigned char* seed) {
    const unsigned n = params->n;
    const int fam = fam_of(params->func, n);
    if (fam < 0) return -1;
    return guarded([&] {
        std::lock_guard<std::mutex> lk(engine_mutex());
        Engine* e = Engine::get();
        if (!e) return XMSS_ENODEV;
        memset(sk, 0, params->index_bytes);
        unsigned char* s = sk + params->index_bytes;
        memcpy(s, seed, 2 * n);
        memcpy(s + 3 * n, seed + 2 * n, n);
        memcpy(pk + n, s + 3 * n, n);
        KeySeeds ks{s, s + n, pk + n};
        e->set_keys(fam, 1, &ks);
        LeafRun run{0, params->d - 1, 0, 0};
        e->build_trees(fam, params->tree_height, 1, &run);
        NodeRef root{0, params->tree_height, 0};
        e->fetch_nodes(fam, params->tree_height, 1, &root, pk);
        memcpy(s + 2 * n, pk, n);
        e->scrub_keys();
        return 0;
    });
}

// xmss_core.c:168-177
int xmssmt_core_keypair(const xmss_params* params, unsigned char* pk, unsigned char* sk) {
    std::vector<unsigned char> seed(3 * params->n);
    randombytes(seed.data(), seed.size());
    return xmssmt_core_seed_keypair(params, pk, sk, seed.data());
}

// xmss_core.c:103-110
int xmss_core_keypair(const xmss_params* params, unsigned char* pk, unsigned char* sk) {
    return xmssmt_core_keypair(params, pk, sk);
}

// xmss_core.c:183-273
int xmssmt_core_sign(const xmss_params* params, unsigned char* sk, unsigned char* sm,
                     unsigned long long* smlen, const unsigned char* m, unsigned long long mlen) {
    const unsigned n = params->n, ib = params->index_bytes;
    const int fam = fam_of(params->func, n);
    if (fam < 0 || mlen > kMaxMsgBytes) return -1;
    // a device failure leaves the sk as it was (no signature left the call)
    std::vector<unsigned char> before(sk, sk + params->sk_bytes);
    const int rc = guarded([&] {
        std::lock_guard<std::mutex> lk(engine_mutex());
        Engine* e = Engine::get();
        if (!e) return XMSS_ENODEV;
        memmove(sm + params->sig_bytes, m, mlen);
        *smlen = params->sig_bytes + mlen;
        const unsigned long long idx = (unsigned long)bytes_to_ull(sk, ib);
        if (index_check(params, sk, idx)) return -2;
        memcpy(sm, sk, ib);
        // the index is advanced before the signature exists (xmss_core.c:235-239)
        ull_to_bytes(sk, ib, idx + 1);
        Engine::SignItem it{sk + ib, sk + ib + n, sk + ib + 3 * n, sk + ib + 2 * n, idx, sm + params->sig_bytes, mlen};
        // one stream, one wait: the staging copy, the launches and the result
        // copy are ordered on it, and collect waits for the last
        Engine::Batch* b = e->sign_prepare(fam, params->tree_height, params->d, 1, &it, false, false);
        struct Free {
            Engine* e;
            Engine::Batch* b;
            ~Free() {
                try {
                    e->free_batch(b);
                } catch (const DeviceError&) {
                }
            }
        } fb{e, b};
        e->run(b, false);
        unsigned char* out = sm + ib;
        e->collect_sign(b, &out);
        return 0;
    });
    if (rc == XMSS_ENODEV) memcpy(sk, before.data(), before.size());
    volatile unsigned char* v = before.data();
    for (size_t i = 0; i < before.size(); ++i) v[i] = 0;
    return rc;
}

// Consecutive signatures of one key (the reference's GPU "opk" mode,
// gpu/gpu_sign_origin.h:23-26): the same bytes and final sk as `count`
// successive xmss[mt]_sign calls, with every subtree the signatures share
// built once.  The last index and anything after it go through
// xmssmt_core_sign one at a time (wiped key, then -2).
int xmss_mi355x_sign_many(int mt, unsigned char* sk, unsigned char* const* sms, unsigned long long* smlens,
                          const unsigned char* const* ms, const unsigned long long* mlens, size_t count,
                          int* rets) {
    xmss_params p;
    const uint32_t oid = (uint32_t)bytes_to_ull(sk, XMSS_OID_LEN);
    if ((mt ? xmssmt_parse_oid(&p, oid) : xmss_parse_oid(&p, oid)) != 0) {
        for (size_t i = 0; i < count; ++i) rets[i] = -1;
        return -1;
    }
    const unsigned n = p.n, ib = p.index_bytes;
    const int fam = fam_of(p.func, n);
    unsigned char* k = sk + XMSS_OID_LEN;
    const unsigned long long idx0 = bytes_to_ull(k, ib);
    const unsigned long long last = (1ULL << p.full_height) - 1;
    size_t shared = (fam >= 0 && idx0 < last) ? (size_t)std::min<unsigned long long>(count, last - idx0) : 0;
    // an over-long message is refused without using an index: from there on
    // the messages go one at a time, as successive calls would
    for (size_t i = 0; i < shared; ++i)
        if (mlens[i] > kMaxMsgBytes) shared = i;
    if (shared) {
        std::vector<unsigned char> before(k, k + p.sk_bytes);
        const int rc = guarded([&] {
            std::lock_guard<std::mutex> lk(engine_mutex());
            Engine* e = Engine::get();
            if (!e) return XMSS_ENODEV;
            std::vector<Engine::SignItem> items(shared);
            std::vector<unsigned char*> outs(shared);
            for (size_t i = 0; i < shared; ++i) {
                unsigned char* sm = sms[i];
                memmove(sm + p.sig_bytes, ms[i], mlens[i]);
                smlens[i] = p.sig_bytes + mlens[i];
                ull_to_bytes(sm, ib, idx0 + i);
                items[i] = {k + ib, k + ib + n, k + ib + 3 * n, k + ib + 2 * n, idx0 + i, sm + p.sig_bytes, mlens[i]};
                outs[i] = sm + ib;
            }
            // the index is past all of them before any signature exists
            ull_to_bytes(k, ib, idx0 + shared);
            Engine::Batch* b = e->sign_prepare(fam, p.tree_height, p.d, (int)shared, items.data(), true, false);
            struct Free {
                Engine* e;
                Engine::Batch* b;
                ~Free() {
                    try {
                        e->free_batch(b);
                    } catch (const DeviceError&) {
                    }
                }
            } fb{e, b};
            e->run(b, false);
            e->collect_sign(b, outs.data());
            return 0;
        });
        if (rc != 0) {  // no signature left the call: the sk as it was
            memcpy(k, before.data(), before.size());
            for (size_t i = 0; i < count; ++i) rets[i] = rc;
            return rc;
        }
        for (size_t i = 0; i < shared; ++i) rets[i] = 0;
    }
    for (size_t i = shared; i < count; ++i) rets[i] = xmssmt_core_sign(&p, k, sms[i], &smlens[i], ms[i], mlens[i]);
    return 0;
}

// xmss_core.c:116-126
int xmss_core_sign(const xmss_params* params, unsigned char* sk, unsigned char* sm, unsigned long long* smlen,
                   const unsigned char* m, unsigned long long mlen) {
    return xmssmt_core_sign(params, sk, sm, smlen, m, mlen);
}

}  // extern "C"
