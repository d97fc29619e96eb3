// core_plain.cpp -- the plain core (drop-in for cuxmss/xmss_core.c): the
// secret key is idx || SK_SEED || SK_PRF || root || PUB_SEED and every
// signature rebuilds its subtrees, here as one GPU tree build per call.
#include <string.h>

#include <vector>

#include "../../include/randombytes.h"
#include "../../include/xmss_core.h"
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
    std::lock_guard<std::mutex> lk(engine_mutex());
    Engine* e = Engine::get();
    if (!e) return XMSS_ENODEV;
    memset(sk, 0, params->index_bytes);
    sk += params->index_bytes;
    memcpy(sk, seed, 2 * n);
    memcpy(sk + 3 * n, seed + 2 * n, n);
    memcpy(pk + n, sk + 3 * n, n);
    KeySeeds ks{sk, sk + n, pk + n};
    e->set_keys(fam, 1, &ks);
    LeafRun run{0, params->d - 1, 0, 0};
    e->build_trees(fam, params->tree_height, 1, &run);
    NodeRef root{0, params->tree_height, 0};
    e->fetch_nodes(fam, params->tree_height, 1, &root, pk);
    memcpy(sk + 2 * n, pk, n);
    return 0;
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
    if (fam < 0) return -1;
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
    Engine::Batch* b = e->sign_prepare(fam, params->tree_height, params->d, 1, &it);
    e->run(b);
    unsigned char* out = sm + ib;
    e->collect_sign(b, &out);
    e->free_batch(b);
    return 0;
}

// xmss_core.c:116-126
int xmss_core_sign(const xmss_params* params, unsigned char* sk, unsigned char* sm, unsigned long long* smlen,
                   const unsigned char* m, unsigned long long mlen) {
    return xmssmt_core_sign(params, sk, sm, smlen, m, mlen);
}

}  // extern "C"
