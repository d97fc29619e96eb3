// core_bds.cpp -- the BDS core (drop-in for cuxmss/xmss_core_fast.c): the
// secret key carries the serialized traversal state, so a signature costs at
// most (h - k)/2 leaf computations instead of a full tree.
//
// The traversal state machine runs on the host over the sk bytes, exactly as
// the reference maps its state onto them (xmss_core_fast.c:38-136): node
// arrays are views into sk, integer fields are read on load and written back
// on store, so fields the reference never assigns keep the caller's bytes.
// Every hash it needs -- leaves (gen_leaf_wots), node merges (thash_h), WOTS
// signatures, R and H_msg, and the keygen trees -- is computed on the GPU,
// through a BdsOps backend (bds.h).
//
// Control flow never depends on hash values, only on the integer state, so
// a signature first runs on a copy of sk with a recording backend: that dry
// run lists exactly the leaves the signature will consume, which are then
// computed in a single launch before the real run.
#include <string.h>

#include <algorithm>
#include <array>
#include <map>
#include <vector>

#include "../../include/randombytes.h"
#include "../../include/xmss_core.h"
#include "bds.h"
#include "core_common.h"

using namespace xm;

namespace {

constexpr uint32_t ADDR_OTS = 0, ADDR_LTREE = 1, ADDR_HASHTREE = 2;

struct TreehashInst {
    unsigned char h;
    unsigned long next_idx;
    unsigned char stackusage;
    unsigned char completed;
    unsigned char* node;
};

struct BdsState {
    unsigned char* stack;
    unsigned int stackoffset;
    unsigned char* stacklevels;
    unsigned char* auth;
    unsigned char* keep;
    TreehashInst* th;
    unsigned char* retain;
    unsigned int next_leaf;
};

void subtree_addrs(const uint32_t addr[8], uint32_t nd[8]) {
    memset(nd, 0, 32);
    nd[0] = addr[0];
    nd[1] = addr[1];
    nd[2] = addr[2];
    nd[3] = ADDR_HASHTREE;
}

// xmssmt_deserialize_state (xmss_core_fast.c:80-136)
void load_states(const xmss_params* p, BdsState* st, TreehashInst* ths, unsigned char** wots_sigs,
                 unsigned char* sk) {
    const unsigned nth = p->tree_height - p->bds_k;
    sk += p->index_bytes + 4 * p->n;
    for (unsigned i = 0; i < 2 * p->d - 1; ++i) {
        st[i].th = ths + i * nth;
        st[i].stack = sk;
        sk += (p->tree_height + 1) * p->n;
        st[i].stackoffset = (unsigned)bytes_to_ull(sk, 4);
        sk += 4;
        st[i].stacklevels = sk;
        sk += p->tree_height + 1;
        st[i].auth = sk;
        sk += p->tree_height * p->n;
        st[i].keep = sk;
        sk += (p->tree_height >> 1) * p->n;
        for (unsigned j = 0; j < nth; ++j) {
            st[i].th[j].h = (unsigned char)bytes_to_ull(sk, 1);
            st[i].th[j].next_idx = (unsigned long)bytes_to_ull(sk + 1, 4);
            st[i].th[j].stackusage = (unsigned char)bytes_to_ull(sk + 5, 1);
            st[i].th[j].completed = (unsigned char)bytes_to_ull(sk + 6, 1);
            st[i].th[j].node = sk + 7;
            sk += 7 + p->n;
        }
        st[i].retain = sk;
        sk += ((1 << p->bds_k) - p->bds_k - 1) * p->n;
        st[i].next_leaf = (unsigned)bytes_to_ull(sk, 4);
        sk += 4;
    }
    if (wots_sigs) *wots_sigs = sk;
}

// xmssmt_serialize_state (xmss_core_fast.c:38-78)
void store_states(const xmss_params* p, unsigned char* sk, const BdsState* st) {
    const unsigned nth = p->tree_height - p->bds_k;
    sk += p->index_bytes + 4 * p->n;
    for (unsigned i = 0; i < 2 * p->d - 1; ++i) {
        sk += (p->tree_height + 1) * p->n;
        ull_to_bytes(sk, 4, st[i].stackoffset);
        sk += 4 + p->tree_height + 1 + p->tree_height * p->n + (p->tree_height >> 1) * p->n;
        for (unsigned j = 0; j < nth; ++j) {
            ull_to_bytes(sk, 1, st[i].th[j].h);
            ull_to_bytes(sk + 1, 4, st[i].th[j].next_idx);
            ull_to_bytes(sk + 5, 1, st[i].th[j].stackusage);
            ull_to_bytes(sk + 6, 1, st[i].th[j].completed);
            sk += 7 + p->n;
        }
        sk += ((1 << p->bds_k) - p->bds_k - 1) * p->n;
        ull_to_bytes(sk, 4, st[i].next_leaf);
        sk += 4;
    }
}

void swap_bytes(unsigned char* a, unsigned char* b, size_t len) {
    for (size_t i = 0; i < len; ++i) std::swap(a[i], b[i]);
}

// deep_state_swap (xmss_core_fast.c:165-193): swaps contents, not pointers
void swap_states(const xmss_params* p, BdsState* a, BdsState* b) {
    swap_bytes(a->stack, b->stack, (p->tree_height + 1) * p->n);
    std::swap(a->stackoffset, b->stackoffset);
    swap_bytes(a->stacklevels, b->stacklevels, p->tree_height + 1);
    swap_bytes(a->auth, b->auth, p->tree_height * p->n);
    swap_bytes(a->keep, b->keep, (p->tree_height >> 1) * p->n);
    for (unsigned i = 0; i < p->tree_height - p->bds_k; ++i) {
        std::swap(a->th[i].h, b->th[i].h);
        std::swap(a->th[i].next_idx, b->th[i].next_idx);
        std::swap(a->th[i].stackusage, b->th[i].stackusage);
        std::swap(a->th[i].completed, b->th[i].completed);
        swap_bytes(a->th[i].node, b->th[i].node, p->n);
    }
    swap_bytes(a->retain, b->retain, ((1 << p->bds_k) - p->bds_k - 1) * p->n);
    std::swap(a->next_leaf, b->next_leaf);
}

// Where a right child of height nodeh and index j goes when the stack merges
// it (xmss_core_fast.c:262-269 and :534-541).
unsigned char* record_slot(const xmss_params* p, BdsState* st, unsigned nodeh, unsigned long j) {
    if (j == 1) return st->auth + nodeh * p->n;
    if (nodeh < p->tree_height - p->bds_k && j == 3) return st->th[nodeh].node;
    if (nodeh >= p->tree_height - p->bds_k)
        return st->retain + ((1 << (p->tree_height - 1 - nodeh)) + nodeh - p->tree_height + ((j - 3) >> 1)) * p->n;
    return nullptr;
}

// treehash_init (xmss_core_fast.c:216-285) for several subtrees at once: the
// whole subtree is built on the GPU and the nodes the stack walk would have
// recorded (auth: index 1, treehash nodes: index 3, retain: odd indices >= 3
// of the top bds_k heights) are read back.  Returns the roots.
void treehash_init_many(const xmss_params* p, Engine* e, int fam, unsigned nt, const uint32_t* layers,
                        BdsState* const* states, unsigned char* roots) {
    const unsigned th = p->tree_height, n = p->n;
    std::vector<LeafRun> runs(nt);
    for (unsigned t = 0; t < nt; ++t) runs[t] = {0, layers[t], 0, 0};
    e->build_trees(fam, th, (int)nt, runs.data());
    std::vector<NodeRef> refs;
    std::vector<unsigned char*> dst;
    for (unsigned t = 0; t < nt; ++t) {
        BdsState* st = states[t];
        for (unsigned i = 0; i < th - p->bds_k; ++i) {
            st->th[i].h = (unsigned char)i;
            st->th[i].completed = 1;
            st->th[i].stackusage = 0;
        }
        for (unsigned h = 0; h < th; ++h) {
            for (unsigned long j = 1; j < (1ul << (th - h)); j += 2) {
                unsigned char* slot = record_slot(p, st, h, j);
                if (slot) {
                    refs.push_back({t, h, (uint32_t)j});
                    dst.push_back(slot);
                }
                if (h < th - p->bds_k && j >= 3) break;  // only index 1 and 3 below the retain rows
            }
        }
        refs.push_back({t, th, 0});
        dst.push_back(roots + (size_t)t * n);
    }
    std::vector<unsigned char> got(refs.size() * n);
    e->fetch_nodes(fam, th, (int)refs.size(), refs.data(), got.data());
    for (size_t i = 0; i < refs.size(); ++i) memcpy(dst[i], &got[i * n], n);
}

unsigned treehash_minheight_on_stack(const xmss_params* p, const BdsState* st, const TreehashInst* t) {
    unsigned r = p->tree_height;
    for (unsigned i = 0; i < t->stackusage; ++i) {
        if (st->stacklevels[st->stackoffset - i - 1] < r) r = st->stacklevels[st->stackoffset - i - 1];
    }
    return r;
}

// treehash_update (xmss_core_fast.c:406-452)
void treehash_update(const xmss_params* p, BdsOps& H, TreehashInst* t, BdsState* st, const uint32_t addr[8]) {
    const unsigned n = p->n;
    uint32_t nd[8];
    subtree_addrs(addr, nd);
    unsigned char nb[2 * 64];
    unsigned nodeheight = 0;
    H.leaf(nb, {addr[0], ((uint64_t)addr[1] << 32) | addr[2], (uint32_t)t->next_idx});
    while (t->stackusage > 0 && st->stacklevels[st->stackoffset - 1] == nodeheight) {
        memcpy(nb + n, nb, n);
        memcpy(nb, st->stack + (st->stackoffset - 1) * n, n);
        nd[5] = nodeheight;
        nd[6] = (uint32_t)(t->next_idx >> (nodeheight + 1));
        H.thash_h(nb, nb, nd);
        ++nodeheight;
        --t->stackusage;
        --st->stackoffset;
    }
    if (nodeheight == t->h) {
        memcpy(t->node, nb, n);
        t->completed = 1;
    } else {
        memcpy(st->stack + st->stackoffset * n, nb, n);
        ++t->stackusage;
        st->stacklevels[st->stackoffset] = (unsigned char)nodeheight;
        ++st->stackoffset;
        ++t->next_idx;
    }
}

// bds_treehash_update (xmss_core_fast.c:458-491).  The reference returns
// `char`; the result is narrowed and sign-extended the same way.
unsigned bds_treehash_update(const xmss_params* p, BdsOps& H, BdsState* st, unsigned updates,
                             const uint32_t addr[8]) {
    const unsigned nth = p->tree_height - p->bds_k;
    unsigned used = 0;
    for (unsigned j = 0; j < updates; ++j) {
        unsigned l_min = p->tree_height, level = nth;
        for (unsigned i = 0; i < nth; ++i) {
            unsigned low;
            if (st->th[i].completed) low = p->tree_height;
            else if (st->th[i].stackusage == 0) low = i;
            else low = treehash_minheight_on_stack(p, st, &st->th[i]);
            if (low < l_min) {
                level = i;
                l_min = low;
            }
        }
        if (level == nth) break;
        treehash_update(p, H, &st->th[level], st, addr);
        ++used;
    }
    return (unsigned)(int)(signed char)(updates - used);
}

// bds_state_update (xmss_core_fast.c:497-552)
int bds_state_update(const xmss_params* p, BdsOps& H, BdsState* st, const uint32_t addr[8]) {
    const unsigned n = p->n;
    const int idx = (int)st->next_leaf;
    if (idx == 1 << p->tree_height) return -1;
    uint32_t nd[8];
    subtree_addrs(addr, nd);
    H.leaf(st->stack + st->stackoffset * n, {addr[0], ((uint64_t)addr[1] << 32) | addr[2], (uint32_t)idx});
    st->stacklevels[st->stackoffset] = 0;
    ++st->stackoffset;
    while (st->stackoffset > 1 && st->stacklevels[st->stackoffset - 1] == st->stacklevels[st->stackoffset - 2]) {
        const unsigned nodeh = st->stacklevels[st->stackoffset - 1];
        unsigned char* slot = record_slot(p, st, nodeh, (unsigned long)(idx >> nodeh));
        if (slot) memcpy(slot, st->stack + (st->stackoffset - 1) * n, n);
        nd[5] = st->stacklevels[st->stackoffset - 1];
        nd[6] = (uint32_t)(idx >> (st->stacklevels[st->stackoffset - 1] + 1));
        H.thash_h(st->stack + (st->stackoffset - 2) * n, st->stack + (st->stackoffset - 2) * n, nd);
        ++st->stacklevels[st->stackoffset - 2];
        --st->stackoffset;
    }
    ++st->next_leaf;
    return 0;
}

// bds_round (xmss_core_fast.c:559-626)
void bds_round(const xmss_params* p, BdsOps& H, BdsState* st, unsigned long leaf_idx, const uint32_t addr[8]) {
    const unsigned n = p->n, th = p->tree_height, nth = th - p->bds_k;
    unsigned tau = th;
    unsigned char buf[2 * 64];
    uint32_t nd[8];
    subtree_addrs(addr, nd);
    for (unsigned i = 0; i < th; ++i) {
        if (!((leaf_idx >> i) & 1)) {
            tau = i;
            break;
        }
    }
    if (tau > 0) {
        memcpy(buf, st->auth + (tau - 1) * n, n);
        memcpy(buf + n, st->keep + ((tau - 1) >> 1) * n, n);
    }
    if (!((leaf_idx >> (tau + 1)) & 1) && (tau < th - 1)) {
        memcpy(st->keep + (tau >> 1) * n, st->auth + tau * n, n);
    }
    if (tau == 0) {
        H.leaf(st->auth, {addr[0], ((uint64_t)addr[1] << 32) | addr[2], (uint32_t)leaf_idx});
    } else {
        nd[5] = tau - 1;
        nd[6] = (uint32_t)(leaf_idx >> tau);
        H.thash_h(st->auth + tau * n, buf, nd);
        for (unsigned i = 0; i < tau; ++i) {
            if (i < nth) {
                memcpy(st->auth + i * n, st->th[i].node, n);
            } else {
                const unsigned offset = (1 << (th - 1 - i)) + i - th;
                const unsigned rowidx = (unsigned)(((leaf_idx >> i) - 1) >> 1);
                memcpy(st->auth + i * n, st->retain + (offset + rowidx) * n, n);
            }
        }
        const unsigned lim = tau < nth ? tau : nth;
        for (unsigned i = 0; i < lim; ++i) {
            const unsigned startidx = (unsigned)(leaf_idx + 1 + 3 * (1u << i));
            if (startidx < 1u << th) {
                st->th[i].h = (unsigned char)i;
                st->th[i].next_idx = startidx;
                st->th[i].completed = 0;
                st->th[i].stackusage = 0;
            }
        }
    }
}

int prepare_engine(const xmss_params* p, Engine** e, int* fam) {
    *fam = fam_of(p->func, p->n);
    if (*fam < 0) return -1;
    *e = Engine::get();
    return *e ? 0 : XMSS_ENODEV;
}

}  // namespace

namespace xm {

// The tag pattern: bytes no hash output or unwritten sk region is expected
// to hold (a chance of 2^-(8(n-5)) per node value).
static const unsigned char* tag_pattern() {
    // dry runs run on many threads at once (serving workers): initialized once, thread-safe
    static const std::array<unsigned char, 64> t = [] {
        std::array<unsigned char, 64> a{};
        uint64_t x = 0x5852a7e3c11d94b5ull;  // fixed: the tags never leave the dry run
        for (unsigned char& b : a) {
            x = x * 6364136223846793005ull + 1442695040888963407ull;
            b = (unsigned char)(x >> 56) | 0x01;
        }
        return a;
    }();
    return t.data();
}
static void write_tag(unsigned char* out, unsigned n, uint32_t kind, size_t index) {
    memcpy(out, tag_pattern(), n - 5);
    out[n - 5] = (unsigned char)(kind >> 28);
    for (int b = 0; b < 4; ++b) out[n - 4 + b] = (unsigned char)(index >> (24 - 8 * b));
}

void RecordingOps::leaf(unsigned char* out, const LeafId& id) {
    write_tag(out, n, Engine::DAG_LEAF, leaves.size());
    leaves.push_back(id);
}
void RecordingOps::thash_h(unsigned char* out, const unsigned char* in, const uint32_t* adrs) {
    uint32_t src[2], lev = 1;
    for (int h = 0; h < 2; ++h) {
        const unsigned char* x = in + h * n;
        const uint32_t kind = (uint32_t)x[n - 5] << 28;
        const uint32_t idx = ((uint32_t)x[n - 4] << 24) | ((uint32_t)x[n - 3] << 16) | ((uint32_t)x[n - 2] << 8) | x[n - 1];
        if (memcmp(x, tag_pattern(), n - 5) == 0 && (kind == Engine::DAG_LEAF || kind == Engine::DAG_MERGE)) {
            src[h] = kind | idx;
            if (kind == Engine::DAG_MERGE) lev = std::max(lev, dag[3 * idx + 2] + 1);
        } else {
            src[h] = Engine::DAG_LIT | (uint32_t)(lits.size() / n);
            lits.insert(lits.end(), x, x + n);
        }
    }
    const size_t k = merges.size() / 8;
    merges.insert(merges.end(), adrs, adrs + 8);
    dag.push_back(src[0]);
    dag.push_back(src[1]);
    dag.push_back(lev);
    write_tag(out, n, Engine::DAG_MERGE, k);
}
void RecordingOps::wots_sign(unsigned char* sig, const unsigned char*, uint32_t layer, uint64_t tree, uint32_t ots) {
    if (layer > 0) {  // an XMSS^MT subtree root signed by the layer above
        const uint32_t w[4] = {layer, (uint32_t)(tree >> 32), (uint32_t)tree, ots};
        wots_up.insert(wots_up.end(), w, w + 4);
    }
    memset(sig, 0, (size_t)(2 * n + 3) * n);
}
void RecordingOps::hmsg_sign(unsigned char* R, unsigned char* digest, uint64_t, const unsigned char*,
                             const unsigned char*, uint64_t) {
    memset(R, 0, n);
    memset(digest, 0, n);
}

// xmss_core_sign (xmss_core_fast.c:710-851)
static int sign_xmss(const xmss_params* p, unsigned char* sk, unsigned char* sm, unsigned long long* smlen,
                     const unsigned char* m, unsigned long long mlen, BdsOps& ops) {
    const unsigned n = p->n, th = p->tree_height;
    std::vector<TreehashInst> ths(th - p->bds_k + 1);
    BdsState st;
    load_states(p, &st, ths.data(), nullptr, sk);
    const unsigned long idx = ((unsigned long)sk[0] << 24) | ((unsigned long)sk[1] << 16) |
                              ((unsigned long)sk[2] << 8) | sk[3];
    if (index_check(p, sk, idx)) return -2;
    unsigned char sk_seed[64], sk_prf[64], pub_seed[64], pub_root[64];
    memcpy(sk_seed, sk + p->index_bytes, n);
    memcpy(sk_prf, sk + p->index_bytes + n, n);
    memcpy(pub_seed, sk + p->index_bytes + 3 * n, n);
    memcpy(pub_root, sk + p->index_bytes + 2 * n, n);
    sk[0] = ((idx + 1) >> 24) & 255;
    sk[1] = ((idx + 1) >> 16) & 255;
    sk[2] = ((idx + 1) >> 8) & 255;
    sk[3] = (idx + 1) & 255;
    ops.bind_key(sk_seed, sk_prf, pub_seed);
    memmove(sm + p->sig_bytes, m, mlen);
    unsigned char R[64], mh[64];
    ops.hmsg_sign(R, mh, idx, pub_root, sm + p->sig_bytes, mlen);
    ops.wots_sign(sm + 4 + n, mh, 0, 0, (uint32_t)idx);
    sm[0] = (idx >> 24) & 255;
    sm[1] = (idx >> 16) & 255;
    sm[2] = (idx >> 8) & 255;
    sm[3] = idx & 255;
    memcpy(sm + 4, R, n);
    memcpy(sm + 4 + n + p->wots_sig_bytes, st.auth, th * n);
    const uint32_t ots_addr[8] = {0, 0, 0, 0, (uint32_t)idx, 0, 0, 0};
    if (idx < (1UL << th) - 1) {
        bds_round(p, ops, &st, idx, ots_addr);
        bds_treehash_update(p, ops, &st, (th - p->bds_k) >> 1, ots_addr);
    }
    *smlen = p->sig_bytes + mlen;
    store_states(p, sk, &st);
    return 0;
}

// xmssmt_core_sign (xmss_core_fast.c:917-1121)
static int sign_xmssmt(const xmss_params* p, unsigned char* sk, unsigned char* sm, unsigned long long* smlen,
                       const unsigned char* m, unsigned long long mlen, BdsOps& ops) {
    const unsigned n = p->n, th = p->tree_height, d = p->d, nth = th - p->bds_k;
    std::vector<BdsState> st(2 * d - 1);
    std::vector<TreehashInst> ths((2 * d - 1) * nth + 1);
    unsigned char* wots_sigs;
    load_states(p, st.data(), ths.data(), &wots_sigs, sk);
    unsigned long long idx = 0;
    for (unsigned i = 0; i < p->index_bytes; ++i) idx |= ((unsigned long long)sk[i]) << 8 * (p->index_bytes - 1 - i);
    if (index_check(p, sk, idx)) return -2;
    unsigned char sk_seed[64], sk_prf[64], pub_seed[64], pub_root[64];
    memcpy(sk_seed, sk + p->index_bytes, n);
    memcpy(sk_prf, sk + p->index_bytes + n, n);
    memcpy(pub_seed, sk + p->index_bytes + 3 * n, n);
    memcpy(pub_root, sk + p->index_bytes + 2 * n, n);
    for (unsigned i = 0; i < p->index_bytes; ++i) sk[i] = ((idx + 1) >> 8 * (p->index_bytes - 1 - i)) & 255;
    ops.bind_key(sk_seed, sk_prf, pub_seed);
    memmove(sm + p->sig_bytes, m, mlen);
    uint64_t idx_tree = idx >> th;
    uint32_t idx_leaf = (uint32_t)(idx & ((1u << th) - 1));
    unsigned char R[64], mh[64];
    ops.hmsg_sign(R, mh, idx, pub_root, sm + p->sig_bytes, mlen);
    ops.wots_sign(sm + p->index_bytes + n, mh, 0, idx_tree, idx_leaf);
    for (unsigned i = 0; i < p->index_bytes; ++i) sm[i] = (idx >> 8 * (p->index_bytes - 1 - i)) & 255;
    memcpy(sm + p->index_bytes, R, n);
    unsigned char* s = sm + p->index_bytes + n + p->wots_sig_bytes;
    memcpy(s, st[0].auth, th * n);
    s += th * n;
    for (unsigned i = 1; i < d; ++i) {
        memcpy(s, wots_sigs + (i - 1) * p->wots_sig_bytes, p->wots_sig_bytes);
        s += p->wots_sig_bytes;
        memcpy(s, st[i].auth, th * n);
        s += th * n;
    }

    uint32_t addr[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int needswap_upto = -1;
    unsigned updates = (th - p->bds_k) >> 1;
    addr[1] = (uint32_t)((idx_tree + 1) >> 32);
    addr[2] = (uint32_t)(idx_tree + 1);
    if ((1 + idx_tree) * (1ULL << th) + idx_leaf < (1ULL << p->full_height)) {
        bds_state_update(p, ops, &st[d], addr);
    }
    for (unsigned i = 0; i < d; ++i) {
        if (!(((idx + 1) & ((1ULL << ((i + 1) * th)) - 1)) == 0)) {
            idx_leaf = (uint32_t)((idx >> (th * i)) & ((1u << th) - 1));
            idx_tree = idx >> (th * (i + 1));
            addr[0] = i;
            addr[1] = (uint32_t)(idx_tree >> 32);
            addr[2] = (uint32_t)idx_tree;
            if ((int)i == needswap_upto + 1) bds_round(p, ops, &st[i], idx_leaf, addr);
            updates = bds_treehash_update(p, ops, &st[i], updates, addr);
            addr[1] = (uint32_t)((idx_tree + 1) >> 32);
            addr[2] = (uint32_t)(idx_tree + 1);
            if ((1 + idx_tree) * (1ULL << th) + idx_leaf < (1ULL << (p->full_height - th * i))) {
                if (i > 0 && updates > 0 && st[d + i].next_leaf < (1ULL << p->full_height)) {
                    bds_state_update(p, ops, &st[d + i], addr);
                    --updates;
                }
            }
        } else if (idx < (1ULL << p->full_height) - 1) {
            swap_states(p, &st[d + i], &st[i]);
            const uint64_t t = (idx + 1) >> ((i + 2) * th);
            const uint32_t ots = (uint32_t)(((idx >> ((i + 1) * th)) + 1) & ((1u << th) - 1));
            ops.wots_sign(wots_sigs + i * p->wots_sig_bytes, st[i].stack, i + 1, t, ots);
            st[d + i].stackoffset = 0;
            st[d + i].next_leaf = 0;
            --updates;  // WOTS signing counts as one update (:1107)
            needswap_upto = (int)i;
            for (unsigned j = 0; j < nth; ++j) st[i].th[j].completed = 1;
        }
    }
    *smlen = p->sig_bytes + mlen;
    store_states(p, sk, st.data());
    return 0;
}

int bds_sign(const xmss_params* p, unsigned char* sk, unsigned char* sm, unsigned long long* smlen,
             const unsigned char* m, unsigned long long mlen, BdsOps& ops) {
    return p->d == 1 ? sign_xmss(p, sk, sm, smlen, m, mlen, ops) : sign_xmssmt(p, sk, sm, smlen, m, mlen, ops);
}

std::vector<LeafId> bds_plan_leaves(const xmss_params* p, const unsigned char* sk, unsigned long long mlen,
                                    bool in_order) {
    std::vector<unsigned char> skc(sk, sk + p->sk_bytes);
    std::vector<unsigned char> smc(p->sig_bytes + mlen + 1), mc(mlen + 1, 0);
    unsigned long long l = 0;
    RecordingOps rec(p->n);
    bds_sign(p, skc.data(), smc.data(), &l, mc.data(), mlen, rec);
    std::vector<LeafId> v = rec.leaves;
    if (!in_order) {
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    return v;
}

void bds_plan(const xmss_params* p, const unsigned char* sk, unsigned long long mlen, std::vector<LeafId>& leaves,
              std::vector<uint32_t>& merges, std::vector<uint32_t>* dag, std::vector<unsigned char>* lits,
              std::vector<uint32_t>* wots_up, std::vector<unsigned char>* post) {
    std::vector<unsigned char> skc(sk, sk + p->sk_bytes);
    std::vector<unsigned char> smc(p->sig_bytes + mlen + 1), mc(mlen + 1, 0);
    unsigned long long l = 0;
    RecordingOps rec(p->n);
    bds_sign(p, skc.data(), smc.data(), &l, mc.data(), mlen, rec);
    leaves.swap(rec.leaves);
    merges.swap(rec.merges);
    if (dag) dag->swap(rec.dag);
    if (lits) lits->swap(rec.lits);
    if (wots_up) wots_up->swap(rec.wots_up);
    if (post) post->swap(skc);
}

namespace {
// Backend that only lists the leaves asked for; every hash result is zeros.
struct LeafListOps final : BdsOps {
    unsigned n;
    std::vector<LeafId> leaves;
    explicit LeafListOps(unsigned n_) : n(n_) {}
    void bind_key(const unsigned char*, const unsigned char*, const unsigned char*) override {}
    void leaf(unsigned char* out, const LeafId& id) override {
        memset(out, 0, n);
        leaves.push_back(id);
    }
    void thash_h(unsigned char* out, const unsigned char*, const uint32_t*) override { memset(out, 0, n); }
    void wots_sign(unsigned char* sig, const unsigned char*, uint32_t, uint64_t, uint32_t) override {
        memset(sig, 0, (size_t)(2 * n + 3) * n);
    }
    void hmsg_sign(unsigned char* R, unsigned char* digest, uint64_t, const unsigned char*, const unsigned char*,
                   uint64_t) override {
        memset(R, 0, n);
        memset(digest, 0, n);
    }
};
}  // namespace

std::vector<LeafId> bds_leaves_of(const xmss_params* p, const unsigned char* sk) {
    unsigned long long idx = 0;
    for (unsigned i = 0; i < p->index_bytes; ++i) idx |= ((unsigned long long)sk[i]) << 8 * (p->index_bytes - 1 - i);
    if (idx >= (1ULL << p->full_height) - 1) return {};
    std::vector<unsigned char> skc(sk, sk + p->sk_bytes), smc(p->sig_bytes + 1);
    unsigned char m = 0;
    unsigned long long l = 0;
    LeafListOps ops(p->n);
    bds_sign(p, skc.data(), smc.data(), &l, &m, 0, ops);
    std::sort(ops.leaves.begin(), ops.leaves.end());
    ops.leaves.erase(std::unique(ops.leaves.begin(), ops.leaves.end()), ops.leaves.end());
    return ops.leaves;
}

bool bds_device_prepare(const xmss_params* p, unsigned char* sk, unsigned char* sm, const unsigned char* m,
                        unsigned long long mlen, BdsDeviceSign& ds) {
    const unsigned n = p->n, ib = p->index_bytes;
    unsigned long long idx = 0;
    for (unsigned i = 0; i < ib; ++i) idx |= ((unsigned long long)sk[i]) << 8 * (ib - 1 - i);
    if (idx >= (1ULL << p->full_height) - 1) return false;  // wipe / exhausted: host path
    std::vector<LeafId> plan;
    std::vector<unsigned char> post;
    bds_plan(p, sk, mlen, plan, ds.merges, &ds.dag, &ds.lits, &ds.wots_up, &post);
    ds.next = bds_leaves_of(p, post.data());
    ds.layer.resize(plan.size());
    ds.tree.resize(plan.size());
    ds.leaf.resize(plan.size());
    for (size_t i = 0; i < plan.size(); ++i) {
        ds.layer[i] = plan[i].layer;
        ds.tree[i] = plan[i].tree;
        ds.leaf[i] = plan[i].leaf;
    }
    ds.seeds.assign(4 * n, 0);
    memcpy(&ds.seeds[0], sk + ib, n);           // SK_SEED
    memcpy(&ds.seeds[n], sk + ib + n, n);       // SK_PRF
    memcpy(&ds.seeds[2 * n], sk + ib + 3 * n, n);  // PUB_SEED
    memcpy(&ds.seeds[3 * n], sk + ib + 2 * n, n);  // root
    // the index advances before the signature exists (xmss_core_fast.c:769-772, :986-988)
    ull_to_bytes(sk, ib, idx + 1);
    memmove(sm + p->sig_bytes, m, mlen);
    ds.idx = idx;
    ds.job = {};
    ds.job.sk_seed = &ds.seeds[0];
    ds.job.sk_prf = &ds.seeds[n];
    ds.job.pub_seed = &ds.seeds[2 * n];
    ds.job.pub_root = &ds.seeds[3 * n];
    ds.job.sk = sk;
    ds.job.idx = idx;
    ds.job.m = sm + p->sig_bytes;
    ds.job.mlen = mlen;
    ds.job.plan_layer = ds.layer.data();
    ds.job.plan_tree = ds.tree.data();
    ds.job.plan_leaf = ds.leaf.data();
    ds.job.nplan = (uint32_t)plan.size();
    ds.job.plan_merge = ds.merges.data();
    ds.job.nmerge = (uint32_t)(ds.merges.size() / 8);
    ds.job.plan_dag = ds.dag.data();
    ds.job.plan_lit = ds.lits.data();
    ds.job.nlit = (uint32_t)(ds.lits.size() / n);
    ds.job.plan_wots = ds.wots_up.data();
    ds.job.nwots = (uint32_t)(ds.wots_up.size() / 4);
    ds.job.R_out = sm + ib;
    ds.job.wots0_out = sm + ib + n;
    ds.job.tail_out = sm + ib + n + p->wots_sig_bytes;
    ds.job.shape = bds_shape(p);
    return true;
}

void bds_device_finish(const xmss_params* p, unsigned char* sm, unsigned long long* smlen, unsigned long long mlen,
                       const BdsDeviceSign& ds) {
    ull_to_bytes(sm, p->index_bytes, ds.idx);
    *smlen = p->sig_bytes + mlen;
}

Engine::BdsShape bds_shape(const xmss_params* p) {
    return {p->tree_height, p->d, p->full_height, p->index_bytes, p->bds_k, p->wots_sig_bytes, p->sig_bytes,
            p->sk_bytes};
}

// Backend of the C API: one key, launches straight on the engine; the
// planned leaves are computed in one launch when the key is bound.
struct DirectOps final : BdsOps {
    Engine* e;
    int fam;
    unsigned n;
    std::vector<LeafId> plan;
    std::map<LeafId, std::vector<unsigned char>> cache;
    DirectOps(Engine* e_, int fam_, unsigned n_, std::vector<LeafId> plan_)
        : e(e_), fam(fam_), n(n_), plan(std::move(plan_)) {}
    void bind_key(const unsigned char* sk_seed, const unsigned char* sk_prf, const unsigned char* pub_seed) override {
        KeySeeds ks{sk_seed, sk_prf, pub_seed};
        e->set_keys(fam, 1, &ks);
        fetch(plan);
    }
    void fetch(const std::vector<LeafId>& ids) {
        if (ids.empty()) return;
        std::vector<LeafRun> runs(ids.size());
        for (size_t i = 0; i < ids.size(); ++i) runs[i] = {0, ids[i].layer, ids[i].tree, ids[i].leaf};
        std::vector<unsigned char> out(ids.size() * n);
        e->leaves(fam, (int)runs.size(), runs.data(), 1, out.data());
        for (size_t i = 0; i < ids.size(); ++i)
            cache[ids[i]] = std::vector<unsigned char>(out.begin() + i * n, out.begin() + (i + 1) * n);
    }
    void leaf(unsigned char* out, const LeafId& id) override {
        auto it = cache.find(id);
        if (it == cache.end()) {
            fetch({id});
            it = cache.find(id);
        }
        memcpy(out, it->second.data(), n);
    }
    void thash_h(unsigned char* out, const unsigned char* in, const uint32_t adrs[8]) override {
        ThashReq q;
        q.key = 0;
        memcpy(q.adrs, adrs, 32);
        e->thash_h(fam, 1, &q, in, out);
    }
    void wots_sign(unsigned char* sig, const unsigned char* msg, uint32_t layer, uint64_t tree,
                   uint32_t ots) override {
        WotsReq w{0, layer, tree, ots, WOTS_SIGN};
        e->wots(fam, 1, &w, msg, nullptr, sig);
    }
    void hmsg_sign(unsigned char* R, unsigned char* digest, uint64_t idx, const unsigned char* root,
                   const unsigned char* m, uint64_t mlen) override {
        MsgReq mr{0, 1, idx, nullptr, root, m, mlen};
        e->hmsg(fam, 1, &mr, R, digest);
    }
};

}  // namespace xm

extern "C" {

// xmss_core_fast.c:635-649
unsigned long long xmss_xmssmt_core_sk_bytes(const xmss_params* p) {
    return p->index_bytes + 4 * p->n +
           (2 * p->d - 1) * ((p->tree_height + 1) * p->n + 4 + p->tree_height + 1 + p->tree_height * p->n +
                             (p->tree_height >> 1) * p->n + (p->tree_height - p->bds_k) * (7 + p->n) +
                             ((1 << p->bds_k) - p->bds_k - 1) * p->n + 4) +
           (p->d - 1) * p->wots_sig_bytes;
}

// xmss_core_fast.c:656-701.  As in the reference, randombytes is not used:
// SK_SEED = SK_PRF = 0x02.., PUB_SEED = 0x04.. (:678-687).
int xmss_core_keypair(const xmss_params* params, unsigned char* pk, unsigned char* sk) {
    const xmss_params* p = params;
    if (fam_of(p->func, p->n) < 0) return -1;
    return guarded([&] {
        std::lock_guard<std::mutex> lk(engine_mutex());
        Engine* e;
        int fam;
        if (int rc = prepare_engine(p, &e, &fam)) return rc;
        std::vector<TreehashInst> ths(p->tree_height - p->bds_k + 1);
        BdsState st;
        load_states(p, &st, ths.data(), nullptr, sk);
        st.stackoffset = 0;
        st.next_leaf = 0;
        sk[0] = sk[1] = sk[2] = sk[3] = 0;
        memset(sk + p->index_bytes, 2, 2 * p->n);
        memset(sk + p->index_bytes + 3 * p->n, 4, p->n);
        memcpy(pk + p->n, sk + p->index_bytes + 3 * p->n, p->n);
        KeySeeds ks{sk + p->index_bytes, sk + p->index_bytes + p->n, sk + p->index_bytes + 3 * p->n};
        e->set_keys(fam, 1, &ks);
        uint32_t layer = 0;
        BdsState* sp = &st;
        treehash_init_many(p, e, fam, 1, &layer, &sp, pk);
        memcpy(sk + p->index_bytes + 2 * p->n, pk, p->n);
        store_states(p, sk, &st);
        e->scrub_keys();
        return 0;
    });
}

// xmss_core_fast.c:858-908
int xmssmt_core_keypair(const xmss_params* params, unsigned char* pk, unsigned char* sk) {
    const xmss_params* p = params;
    if (fam_of(p->func, p->n) < 0) return -1;
    return guarded([&] {
        const unsigned n = p->n, d = p->d;
        std::lock_guard<std::mutex> lk(engine_mutex());
        Engine* e;
        int fam;
        if (int rc = prepare_engine(p, &e, &fam)) return rc;
        const unsigned nth = p->tree_height - p->bds_k;
        std::vector<BdsState> st(2 * d - 1);
        std::vector<TreehashInst> ths((2 * d - 1) * nth + 1);
        unsigned char* wots_sigs;
        load_states(p, st.data(), ths.data(), &wots_sigs, sk);
        for (unsigned i = 0; i < 2 * d - 1; ++i) {
            st[i].stackoffset = 0;
            st[i].next_leaf = 0;
        }
        memset(sk, 0, p->index_bytes);
        randombytes(sk + p->index_bytes, 2 * n);
        randombytes(sk + p->index_bytes + 3 * n, n);
        memcpy(pk + n, sk + p->index_bytes + 3 * n, n);
        KeySeeds ks{sk + p->index_bytes, sk + p->index_bytes + n, sk + p->index_bytes + 3 * n};
        e->set_keys(fam, 1, &ks);
        // the d layer-0..d-1 trees (tree address 0) are independent: one build
        std::vector<uint32_t> layers(d);
        std::vector<BdsState*> sps(d);
        for (unsigned i = 0; i < d; ++i) {
            layers[i] = i;
            sps[i] = &st[i];
        }
        std::vector<unsigned char> roots((size_t)d * n);
        treehash_init_many(p, e, fam, d, layers.data(), sps.data(), roots.data());
        // layer i+1 signs the root of layer i (leaf 0 of tree 0)
        if (d > 1) {
            std::vector<WotsReq> w(d - 1);
            for (unsigned i = 0; i + 1 < d; ++i) w[i] = {0, i + 1, 0, 0, WOTS_SIGN};
            e->wots(fam, (int)(d - 1), w.data(), roots.data(), nullptr, wots_sigs);
        }
        memcpy(pk, &roots[(size_t)(d - 1) * n], n);
        memcpy(sk + p->index_bytes + 2 * n, pk, n);
        store_states(p, sk, st.data());
        e->scrub_keys();
        return 0;
    });
}

static int bds_sign_api(const xmss_params* p, unsigned char* sk, unsigned char* sm, unsigned long long* smlen,
                        const unsigned char* m, unsigned long long mlen) {
    if (fam_of(p->func, p->n) < 0 || mlen > kMaxMsgBytes) return -1;
    // a device failure leaves the sk (index and traversal state) as it was:
    // no signature left the call
    std::vector<unsigned char> before(sk, sk + p->sk_bytes);
    const int rc = guarded([&] {
        std::lock_guard<std::mutex> lk(engine_mutex());
        Engine* e;
        int fam;
        if (int rc = prepare_engine(p, &e, &fam)) return rc;
        BdsDeviceSign ds;
        if (bds_device_prepare(p, sk, sm, m, mlen, ds)) {
            // the key context stays in the device pool across signatures
            e->pool_begin();
            const KeySeeds ks{ds.job.sk_seed, ds.job.sk_prf, ds.job.pub_seed};
            ds.job.key = (uint32_t)e->pool_slot(fam, ks);
            // the chains of the leaves this key's next signature will consume
            // are hashed beside this signature's WOTS signature; this one
            // finishes the leaves whose chains the previous call hashed
            std::vector<LeafRun> next(ds.next.size());
            for (size_t i = 0; i < next.size(); ++i) next[i] = {0, ds.next[i].layer, ds.next[i].tree, ds.next[i].leaf};
            e->ends_ahead(fam, ks, next.data(), (int)next.size());
            ds.job.sig_values = true;
            e->bds_sign_batch(fam, 1, &ds.job);
            bds_device_finish(p, sm, smlen, mlen, ds);
            return 0;
        }
        // last index (the key is wiped now) or exhausted key: the pool keeps no
        // context of the wiped seeds; no traversal work, host control
        const unsigned n = p->n, ib = p->index_bytes;
        e->ends_ahead_drop();
        e->pool_forget(fam, {sk + ib, sk + ib + n, sk + ib + 3 * n});
        DirectOps ops(e, fam, n, bds_plan_leaves(p, sk, mlen, false));
        const int r = bds_sign(p, sk, sm, smlen, m, mlen, ops);
        e->scrub_keys();
        return r;
    });
    if (rc == XMSS_ENODEV) memcpy(sk, before.data(), before.size());
    volatile unsigned char* v = before.data();
    for (size_t i = 0; i < before.size(); ++i) v[i] = 0;
    return rc;
}

// xmss_core_fast.c:710-851
int xmss_core_sign(const xmss_params* params, unsigned char* sk, unsigned char* sm, unsigned long long* smlen,
                   const unsigned char* m, unsigned long long mlen) {
    return bds_sign_api(params, sk, sm, smlen, m, mlen);
}

// xmss_core_fast.c:917-1121
int xmssmt_core_sign(const xmss_params* params, unsigned char* sk, unsigned char* sm, unsigned long long* smlen,
                     const unsigned char* m, unsigned long long mlen) {
    return bds_sign_api(params, sk, sm, smlen, m, mlen);
}

}  // extern "C"
