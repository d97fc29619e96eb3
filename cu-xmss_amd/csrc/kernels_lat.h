// kernels_lat.h -- latency-shaped leaf, WOTS-signature and verify-layer
// kernels (gen_leaf_wots, wots_sign, and one layer of sign_open:
// wots_pk_from_sig + l_tree + compute_root; xmss_commons.c:16-113,
// wots.c:117-175).
//
// A single leaf is a chain of dependent hashes: 15 thash_f per WOTS chain,
// then the L-tree heights, then (verify) the auth-path walk.  But every PRF
// key and mask, and for SHA-2 the first block of every F / H outer hash,
// depends only on the address, never on a chain value.  So each request runs
// in two launches:
//   k_lat_pre  -- wide: one lane per (chain, step), per L-tree merge and per
//                 auth level computes F::pre_f / F::pre_h (and the chain
//                 starts, PRF_keygen) into a per-request scratch area;
//   k_lat_fin  -- one workgroup per request: the chains run F::fin_f only
//                 (one compression per step for SHA-2 n = 32), the L-tree
//                 and the auth walk F::fin_h (two).
// The hash depth of a leaf drops from about 100 compressions to about 30;
// the work is the same.  Batches of thousands of requests use the one-launch
// kernels of kernels.h instead, where the GPU is full either way.
#pragma once
#include "kernels.h"

namespace xm {

constexpr uint32_t LAT_LEAF = 0, LAT_SIGN = 1, LAT_VERIFY = 2;
// Request fields beyond the address (DLeaf p0..p2) for LAT_VERIFY: p0 = the
// tree height (auth levels) of this request, p1 bit 0 = precompute all 15
// steps of every chain (the message is not known yet when k_lat_pre runs,
// as for the upper XMSS^MT layers), p2 = the offset of its auth path, in
// nodes, from the auth pointer of the k_lat_fin launch.  LAT_SIGN requests
// use p1 bit 0 the same way (then k_lat_pre reads no message).  A BDS
// signature's LAT_LEAF rows may be split over two signatures of a key:
// LAT_ENDS_OUT rows hash only the chains and store the LEN chain ends in
// slot p2 of the ends buffer; LAT_ENDS_IN rows (a later launch) take the
// chain ends from slot p2 and hash only the L-tree.
constexpr uint32_t LAT_ALL_STEPS = 1, LAT_ENDS_OUT = 2, LAT_ENDS_IN = 4;

// Per-request scratch: chain starts [LEN], F pre-images [LEN][15], H
// pre-images [LEN - 1 L-tree merges, in height order][th auth levels].
template <class F>
struct LatLayout {
    using FPre = typename F::FPre;
    using HPre = typename F::HPre;
    static constexpr size_t align(size_t x) { return (x + 255) & ~(size_t)255; }
    static constexpr size_t starts = 0;
    static constexpr size_t fpre = align(sizeof(typename F::Node) * F::LEN);
    static constexpr size_t hpre = fpre + align(sizeof(FPre) * F::LEN * 15);
    static size_t bytes(uint32_t th) { return align(hpre + sizeof(HPre) * (F::LEN - 1 + th)); }
};


// L-tree merge t (0 .. LEN-2, height-major) -> (height, index).
template <class F>
__device__ __forceinline__ void ltree_merge(uint32_t t, uint32_t& height, uint32_t& index) {
    uint32_t l = F::LEN, off = 0, h = 0;
    while (t >= off + (l >> 1)) {
        off += l >> 1;
        l = (l + 1) >> 1;
        ++h;
    }
    height = h;
    index = t - off;
}

// The H pre-image u of a request: L-tree merge u (u < LEN - 1), else auth level u - (LEN - 1).
template <class F>
__device__ __forceinline__ void lat_pre_h_item(const typename F::Key& k, const DLeaf& q, uint32_t u, uint8_t* S) {
    constexpr uint32_t LEN = F::LEN;
    uint32_t a[8];
    if (u < LEN - 1) {  // L-tree merge (xmss_commons.c:16-49)
        uint32_t h, i;
        ltree_merge<F>(u, h, i);
        const uint32_t b[8] = {q.layer, q.tree_hi, q.tree_lo, ADDR_LTREE, q.leaf, h, i, 0};
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] = b[j];
    } else {  // auth level (xmss_commons.c:54-96)
        const uint32_t i = u - (LEN - 1);
        const uint32_t b[8] = {q.layer, q.tree_hi, q.tree_lo, ADDR_HASHTREE, 0, i, q.leaf >> (i + 1), 0};
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] = b[j];
    }
    typename F::HPre p;
    F::pre_h(p, k, a);
    reinterpret_cast<typename F::HPre*>(S + LatLayout<F>::hpre)[u] = p;
}

// Item t of request r (one lane each).
template <class F, uint32_t MODE>
__device__ __forceinline__ void lat_pre_one(const typename F::Key* __restrict__ keys,
                                            const DLeaf* __restrict__ reqs, uint32_t r, uint32_t t,
                                            const uint8_t* __restrict__ msgs, uint8_t* __restrict__ scratch,
                                            size_t stride) {
    using Lay = LatLayout<F>;
    constexpr uint32_t LEN = F::LEN;
    const DLeaf q = reqs[r];
    const uint32_t nf = LEN * 15, nh = (MODE == LAT_SIGN) ? 0 : LEN - 1 + (MODE == LAT_VERIFY ? q.p0 : 0);
    if (t >= LEN + nf + nh) return;
    if (MODE == LAT_LEAF && (q.p1 & (t < LEN + nf ? LAT_ENDS_IN : LAT_ENDS_OUT))) return;  // the other half's work
    const typename F::Key& k = keys[q.key];
    uint8_t* S = scratch + (size_t)r * stride;
    if (t < LEN) {  // chain start: PRF_keygen (wots.c:14-29)
        if (MODE == LAT_VERIFY) return;
        uint32_t a[8] = {q.layer, q.tree_hi, q.tree_lo, ADDR_OTS, q.leaf, t, 0, 0};
        typename F::Node x;
        F::prf_keygen(x, k, a);
        reinterpret_cast<typename F::Node*>(S + Lay::starts)[t] = x;
        return;
    }
    if (t < LEN + nf) {  // chain c, step s
        const uint32_t c = (t - LEN) / 15, s = (t - LEN) % 15;
        if (!(q.p1 & LAT_ALL_STEPS) && (MODE == LAT_SIGN || MODE == LAT_VERIFY)) {
            const uint32_t len = chain_len<F>(msgs + (size_t)r * F::kN, c);
            if (MODE == LAT_VERIFY ? s < len : s >= len) return;
        }
        uint32_t a[8] = {q.layer, q.tree_hi, q.tree_lo, ADDR_OTS, q.leaf, c, s, 0};
        typename F::FPre p;
        F::pre_f(p, k, a);
        reinterpret_cast<typename F::FPre*>(S + Lay::fpre)[c * 15 + s] = p;
        return;
    }
    lat_pre_h_item<F>(k, q, t - LEN - nf, S);
}

template <class F, uint32_t MODE>
__global__ __launch_bounds__(WG) void k_lat_pre(const typename F::Key* __restrict__ keys,
                                                 const DLeaf* __restrict__ reqs,
                                                 const uint8_t* __restrict__ msgs,
                                                 uint8_t* __restrict__ scratch, size_t stride) {
    lat_pre_one<F, MODE>(keys, reqs, blockIdx.y, blockIdx.x * WG + threadIdx.x, msgs, scratch, stride);
}

// H_msg of nmsg verification requests (the first rows, one message per lane:
// the longest-running lanes start first) and the address-only work of their
// verify layers (LAT_VERIFY requests with LAT_ALL_STEPS: the upper layers'
// messages are not known yet) in one launch.
template <class F>
__global__ __launch_bounds__(WG) void k_ver_pre(const typename F::Key* __restrict__ keys,
                                                 const DMsg* __restrict__ msgs, uint32_t nmsg, uint32_t pad,
                                                 uint8_t* __restrict__ blob, uint8_t* __restrict__ digests,
                                                 uint32_t hrows, const DLeaf* __restrict__ reqs,
                                                 uint8_t* __restrict__ scratch, size_t stride) {
    const uint32_t t = blockIdx.x * WG + threadIdx.x;
    if (blockIdx.y < hrows) {
        hmsg_mirrored<F>(keys, msgs, blockIdx.y * gridDim.x * WG + t, nmsg, pad, blob, nullptr, digests);
        return;
    }
    lat_pre_one<F, LAT_VERIFY>(keys, reqs, blockIdx.y - hrows, t, nullptr, scratch, stride);
}

// Only the H pre-images of LAT_VERIFY requests that all have tree height th
// (their F pre-images then come from k_pre_need, once the message is known),
// flat over (request, merge) items.
template <class F>
__global__ __launch_bounds__(WG) void k_lat_pre_h(const typename F::Key* __restrict__ keys,
                                                   const DLeaf* __restrict__ reqs, uint32_t n, uint32_t th,
                                                   uint8_t* __restrict__ scratch, size_t stride, uint32_t prio) {
    wave_priority(prio);
    const uint32_t nh = F::LEN - 1 + th;
    const uint32_t t = blockIdx.x * WG + threadIdx.x;
    if (t >= n * nh) return;
    const uint32_t r = t / nh;
    const DLeaf q = reqs[r];
    lat_pre_h_item<F>(keys[q.key], q, t - r * nh, scratch + (size_t)r * stride);
}

// One workgroup per request, ceil(LEN/64) waves.
//   LAT_LEAF:   out[r] = leaf (n bytes)
//   LAT_SIGN:   out[r] = the WOTS signature (LEN x n bytes) of msgs[r]
//   LAT_VERIFY: out[r] = the root implied by sigs[r], msgs[r] and auth[r]
// L-tree (xmss_commons.c:16-49) of the LEN nodes in pk (LDS), height by
// height from H pre-images hp; thread c hashes pair c (at least LEN / 2
// threads); the root ends in pk[0].
template <class F>
__device__ __forceinline__ void ltree_lds(typename F::Node* pk, const typename F::HPre* hp, uint32_t c) {
    uint32_t l = F::LEN, off = 0;
    while (l > 1) {
        const uint32_t parents = l >> 1;
        typename F::Node o;
        // the idle lanes of a wave that holds merges hash a copy of one of
        // them (mirror_lanes); only the merge's own lane writes below
        const uint32_t m = mirror_lanes(c, parents);
        if (m < parents) F::fin_h(o, hp[off + m], pk[2 * m], pk[2 * m + 1]);
        __syncthreads();
        if (c < parents) pk[c] = o;
        if ((l & 1) && c == 0) pk[l >> 1] = pk[l - 1];
        __syncthreads();
        off += parents;
        l = (l + 1) >> 1;
    }
}

// Request r on the calling workgroup; pk: LEN nodes of LDS.  LAT_VERIFY
// items r >= spec_first take their chain ends from spec (16 per chain, by
// start digit: k_lat_verify_layers' speculative chains) instead of hashing.
template <class F, uint32_t MODE>
__device__ __forceinline__ void lat_fin_one(uint32_t r, typename F::Node* pk, const DLeaf* __restrict__ reqs,
                                            const uint8_t* __restrict__ msgs, const uint8_t* __restrict__ sigs,
                                            const uint8_t* auth, const uint8_t* scratch,
                                            size_t stride, uint8_t* __restrict__ out,
                                            const uint8_t* __restrict__ spec = nullptr, uint32_t spec_first = 0) {
    using Lay = LatLayout<F>;
    using FPre = typename F::FPre;
    using HPre = typename F::HPre;
    const uint32_t own = threadIdx.x;
    const uint8_t* S = scratch + (size_t)r * stride;
    const HPre* hp = reinterpret_cast<const HPre*>(S + Lay::hpre);
    // lanes past the last chain in the last chain's wave walk a copy of a
    // chain (mirror_lanes); only the chain's own lane stores it
    const uint32_t c = mirror_lanes(own, F::LEN);
    if (c < (uint32_t)F::LEN) {
        typename F::Node x;
        uint32_t s0 = 0, s1 = 15;
        if (MODE == LAT_VERIFY) {
            s0 = chain_len<F>(msgs + (size_t)r * F::kN, c);
            if (spec && r >= spec_first) {
                F::load(x, spec + (((size_t)(r - spec_first) * F::LEN + c) * 16 + s0) * F::kN);
                s0 = 15;  // already at the chain's end
            } else {
                F::load(x, sigs + ((size_t)r * F::LEN + c) * F::kN);
            }
        } else {
            if (MODE == LAT_SIGN) s1 = chain_len<F>(msgs + (size_t)r * F::kN, c);
            x = reinterpret_cast<const typename F::Node*>(S + Lay::starts)[c];
        }
        const FPre* fp = reinterpret_cast<const FPre*>(S + Lay::fpre) + c * 15;
        // the next step's pre-image is loaded while this step hashes
        FPre cur = fp[s0 < 15 ? s0 : 14];
        for (uint32_t s = s0; s < s1; ++s) {
            const FPre nxt = fp[s + 1 < 15 ? s + 1 : 14];
            keep_loads_ahead();
            F::fin_f(x, cur, x);
            cur = nxt;
        }
        if (MODE == LAT_SIGN) {
            if (own == c) F::store(x, out + ((size_t)r * F::LEN + c) * F::kN);
            return;
        }
        if (own == c) pk[c] = x;
    }
    if (MODE == LAT_SIGN) return;
    __syncthreads();
    ltree_lds<F>(pk, hp, own);
    // the auth walk: lane 0, with the rest of its wave walking copies
    if (own >= 64) return;
    typename F::Node node = pk[0];
    keep_in_vgprs(node);
    if (MODE == LAT_VERIFY) {
        const DLeaf q = reqs[r];
        uint32_t idx = q.leaf;
        const uint8_t* ap = auth + (size_t)q.p2 * F::kN;
        // each level's sibling and pre-image are loaded one level ahead (a
        // level's loads waited for in place cost about as much as its two
        // compressions: 8.6 us per level measured)
        const uint32_t levels = q.p0;
        typename F::Node sib, sib_next;
        HPre h, h_next;
        if (levels) {
            F::load(sib, ap);
            h = hp[F::LEN - 1];
        }
        for (uint32_t i = 0; i < levels; ++i) {
            const uint32_t j = i + 1 < levels ? i + 1 : i;
            F::load(sib_next, ap + (size_t)j * F::kN);
            h_next = hp[F::LEN - 1 + j];
            keep_loads_ahead();
            typename F::Node l, rr;
            auth_pair<F>(l, rr, node, sib, idx & 1);
            F::fin_h(node, h, l, rr);
            idx >>= 1;
            sib = sib_next;
            h = h_next;
        }
    }
    if (own == 0) F::store(node, out + (size_t)r * F::kN);
}

// The two halves of a LAT_LEAF request split over two launches (the flags
// LAT_ENDS_OUT / LAT_ENDS_IN): the LEN chains of request r to their ends,
// stored at ends; and the L-tree of LEN chain ends read from ends, its root
// stored at out[r].  Every thread of the workgroup calls them.
template <class F>
__device__ __forceinline__ void lat_leaf_chains(uint32_t r, const uint8_t* scratch, size_t stride,
                                                uint8_t* __restrict__ ends) {
    using Lay = LatLayout<F>;
    using FPre = typename F::FPre;
    const uint32_t own = threadIdx.x, c = mirror_lanes(own, F::LEN);
    if (c >= (uint32_t)F::LEN) return;
    const uint8_t* S = scratch + (size_t)r * stride;
    typename F::Node x = reinterpret_cast<const typename F::Node*>(S + Lay::starts)[c];
    const FPre* fp = reinterpret_cast<const FPre*>(S + Lay::fpre) + c * 15;
    FPre cur = fp[0];
    for (uint32_t s = 0; s < 15; ++s) {
        const FPre nxt = fp[s + 1 < 15 ? s + 1 : 14];
        keep_loads_ahead();
        F::fin_f(x, cur, x);
        cur = nxt;
    }
    if (own == c) F::store(x, ends + (size_t)c * F::kN);
}

// Every value along the LEN chains of WOTS signature request r (scratch row r
// holds its chain starts and all 15 steps' F pre-images): values[c * 16 + s]
// = chain c after s steps.  A signature whose message digest is still being
// hashed walks its chains this way beside H_msg; the signature is then the
// value at each chain's digit (lat_sig_select).
template <class F>
__device__ __forceinline__ void lat_sig_values(uint32_t r, const uint8_t* scratch, size_t stride,
                                               uint8_t* __restrict__ values) {
    using Lay = LatLayout<F>;
    using FPre = typename F::FPre;
    const uint32_t own = threadIdx.x, c = mirror_lanes(own, F::LEN);
    if (c >= (uint32_t)F::LEN) return;
    const uint8_t* S = scratch + (size_t)r * stride;
    typename F::Node x = reinterpret_cast<const typename F::Node*>(S + Lay::starts)[c];
    const FPre* fp = reinterpret_cast<const FPre*>(S + Lay::fpre) + c * 15;
    uint8_t* v = values + (size_t)c * 16 * F::kN;
    if (own == c) F::store(x, v);
    FPre cur = fp[0];
    for (uint32_t s = 0; s < 15; ++s) {
        const FPre nxt = fp[s + 1 < 15 ? s + 1 : 14];
        keep_loads_ahead();
        F::fin_f(x, cur, x);
        cur = nxt;
        if (own == c) F::store(x, v + (size_t)(s + 1) * F::kN);
    }
}

// The WOTS signature of message digest msg from the chain values
// lat_sig_values left (values: LEN x 16 nodes): chain c at its digit.
template <class F>
__device__ __forceinline__ void lat_sig_select(const uint8_t* __restrict__ msg, const uint8_t* __restrict__ values,
                                               uint8_t* __restrict__ out) {
    for (uint32_t c = threadIdx.x; c < (uint32_t)F::LEN; c += blockDim.x) {
        typename F::Node x;
        F::load(x, values + ((size_t)c * 16 + chain_len<F>(msg, c)) * F::kN);
        F::store(x, out + (size_t)c * F::kN);
    }
}

template <class F>
__device__ __forceinline__ void lat_leaf_ltree(uint32_t r, typename F::Node* pk, const uint8_t* __restrict__ scratch,
                                               size_t stride, const uint8_t* __restrict__ ends,
                                               uint8_t* __restrict__ out) {
    for (uint32_t c = threadIdx.x; c < (uint32_t)F::LEN; c += blockDim.x) F::load(pk[c], ends + (size_t)c * F::kN);
    __syncthreads();
    ltree_lds<F>(pk, reinterpret_cast<const typename F::HPre*>(scratch + (size_t)r * stride + LatLayout<F>::hpre),
                 threadIdx.x);
    if (threadIdx.x == 0) F::store(pk[0], out + (size_t)r * F::kN);
}

template <class F, uint32_t MODE>
__global__ __launch_bounds__(192) void k_lat_fin(const DLeaf* __restrict__ reqs, const uint8_t* __restrict__ msgs,
                                                  const uint8_t* __restrict__ sigs,
                                                  const uint8_t* auth,  // not restrict: the auth walk's
                                                  const uint8_t* scratch,  // loads stay one level ahead
                                                  size_t stride, uint8_t* __restrict__ out) {
    __shared__ typename F::Node pk[F::LEN];
    lat_fin_one<F, MODE>(blockIdx.x, pk, reqs, msgs, sigs, auth, scratch, stride, out);
}

// Every verify layer of a request in one workgroup, bottom up (one launch
// for any XMSS^MT depth): request j has layer L (item base.b[L] + j of reqs,
// sigs and scratch) while j < base.b[L + 1] - base.b[L]; layer 0 verifies
// digests[j], layer L the root layer L - 1 left in roots.  The roots of every
// item are written to roots; auth offsets (DLeaf p2) count from auth.
struct LayerBase {
    uint32_t b[17];
};
//
// Layers L0 .. L1-1 are walked here (the layers below were walked by an
// earlier launch and left their roots).  Speculation for few XMSS^MT
// requests: an upper layer's chain lengths follow the root of the layer
// below, but the chain from each possible start digit d (16 of them) to the
// end does not, so blocks past the nwalk walkers hash all of them for the
// nspec items from spec_first on (one lane per item, chain and digit, into
// spec) while the walkers verify layer 0; a second launch then walks the
// upper layers with their chain ends looked up (L-tree and auth path only).
template <class F>
__global__ __launch_bounds__(192) void k_lat_verify_layers(const DLeaf* __restrict__ reqs, LayerBase base,
                                                            uint32_t L0, uint32_t L1,
                                                            const uint8_t* __restrict__ digests,
                                                            const uint8_t* __restrict__ sigs,
                                                            const uint8_t* auth,  // not restrict: see k_lat_fin
                                                            const uint8_t* scratch, size_t stride,
                                                            uint8_t* __restrict__ roots, uint32_t nwalk,
                                                            uint8_t* __restrict__ spec, uint32_t spec_first,
                                                            uint32_t nspec) {
    using FPre = typename F::FPre;
    __shared__ typename F::Node pk[F::LEN];
    if (blockIdx.x >= nwalk) {
        const uint32_t g = (blockIdx.x - nwalk) * blockDim.x + threadIdx.x;
        if (g >= nspec * F::LEN * 16) return;
        const uint32_t t = spec_first + g / (F::LEN * 16), c = (g / 16) % F::LEN, d = g % 16;
        typename F::Node x;
        F::load(x, sigs + ((size_t)t * F::LEN + c) * F::kN);
        const FPre* fp = reinterpret_cast<const FPre*>(scratch + (size_t)t * stride + LatLayout<F>::fpre) + c * 15;
        for (uint32_t s = d; s < 15; ++s) F::fin_f(x, fp[s], x);
        F::store(x, spec + (size_t)g * F::kN);
        return;
    }
    const uint32_t j = blockIdx.x;
    for (uint32_t L = L0; L < L1 && j < base.b[L + 1] - base.b[L]; ++L) {
        const uint32_t t = base.b[L] + j;
        // the message of item t is at msgs + t * n: the digest, or the root of item base.b[L - 1] + j
        const uint8_t* msgs = L == 0 ? digests - (size_t)base.b[0] * F::kN
                                     : roots - (size_t)(base.b[L] - base.b[L - 1]) * F::kN;
        lat_fin_one<F, LAT_VERIFY>(t, pk, reqs, msgs, sigs, auth, scratch, stride, roots,
                                   nspec ? nullptr : spec, spec_first);
        __threadfence_block();
        __syncthreads();
    }
}

// Request descriptors of the latency kernels from those of the throughput
// path: every leaf of nruns subtrees of height th (run-major), and WOTS
// signature requests (defined in engine.hip).
__global__ __launch_bounds__(WG) void k_runs_to_leaves(const DRun* __restrict__ runs, uint32_t nruns, uint32_t th,
                                                       DLeaf* __restrict__ out);
__global__ __launch_bounds__(WG) void k_wots_to_leaves(const DWots* __restrict__ w, uint32_t n,
                                                       DLeaf* __restrict__ out);

// ---- subtree levels of a few trees -------------------------------------
// Internal node (level, i) of a 2^th-leaf subtree -- the parent of children
// 2i, 2i+1 at `level` -- has pre-image slot 2^th - 2^(th - level) + i.
__host__ __device__ inline uint64_t tree_pre_slot(uint32_t th, uint32_t level, uint32_t i) {
    return (1ull << th) - (1ull << (th - level)) + i;
}

// Address-only part of every internal node of each run's subtree.
template <class F>
__global__ __launch_bounds__(WG) void k_pre_tree(const typename F::Key* __restrict__ keys,
                                                  const DRun* __restrict__ runs, uint32_t th,
                                                  typename F::HPre* __restrict__ out) {
    const uint32_t r = blockIdx.y;
    const uint64_t h = (uint64_t)blockIdx.x * WG + threadIdx.x;
    const uint64_t per = (1ull << th) - 1;
    if (h >= per) return;
    uint32_t level = 0;
    while (h >= tree_pre_slot(th, level + 1, 0)) ++level;
    const uint32_t i = (uint32_t)(h - tree_pre_slot(th, level, 0));
    const DRun run = runs[r];
    uint32_t a[8] = {run.layer, run.tree_hi, run.tree_lo, ADDR_HASHTREE, 0, level, i, 0};
    typename F::HPre p;
    F::pre_h(p, keys[run.key], a);
    out[(size_t)r * per + h] = p;
}

// One subtree level (parents at level + 1) from pre-images (k_level's split form).
template <class F>
__global__ __launch_bounds__(WG) void k_level_fin(uint32_t th, uint32_t level, uint8_t* __restrict__ nodes,
                                                   const typename F::HPre* __restrict__ pre) {
    const uint32_t r = blockIdx.y;
    const uint32_t i = blockIdx.x * WG + threadIdx.x;
    if (i >= (1u << (th - level - 1))) return;
    const uint64_t stride = (2ull << th) - 1;
    const uint64_t base_c = (2ull << th) - (2ull << (th - level));
    const uint64_t base_p = (2ull << th) - (2ull << (th - level - 1));
    uint8_t* T = nodes + (size_t)r * stride * F::kN;
    typename F::Node L, R, o;
    F::load(L, T + (base_c + 2 * i) * F::kN);
    F::load(R, T + (base_c + 2 * i + 1) * F::kN);
    F::fin_h(o, pre[(size_t)r * ((1ull << th) - 1) + tree_pre_slot(th, level, i)], L, R);
    F::store(o, T + (base_p + i) * F::kN);
}

// The top levels from `level0` (at most 64 children there) up to the root,
// one workgroup of 64 lanes per run, the level in between kept in LDS.
template <class F>
__global__ __launch_bounds__(64) void k_top_fin(uint32_t th, uint32_t level0, uint8_t* __restrict__ nodes,
                                                 const typename F::HPre* __restrict__ pre) {
    __shared__ typename F::Node buf[64];
    const uint32_t r = blockIdx.x;
    const uint32_t t = threadIdx.x;
    const uint64_t stride = (2ull << th) - 1;
    const typename F::HPre* hp = pre + (size_t)r * ((1ull << th) - 1);
    uint8_t* T = nodes + (size_t)r * stride * F::kN;
    const uint32_t m0 = 1u << (th - level0);  // children at level0
    if (t < m0) F::load(buf[t], T + ((2ull << th) - (2ull << (th - level0)) + t) * F::kN);
    __syncthreads();
    for (uint32_t level = level0; level < th; ++level) {
        const uint32_t parents = 1u << (th - level - 1);
        typename F::Node o;
        const uint32_t m = mirror_lanes(t, parents);  // lanes past the parents hash a copy
        F::fin_h(o, hp[tree_pre_slot(th, level, m)], buf[2 * m], buf[2 * m + 1]);
        __syncthreads();
        if (t < parents) {
            buf[t] = o;
            F::store(o, T + ((2ull << th) - (2ull << (th - level - 1)) + t) * F::kN);
        }
        __syncthreads();
    }
}

// F pre-images of the chain steps a verification needs (steps len_c .. 14 of
// chain c, len_c the message's base-w digit), once the message is known: one
// workgroup per request walks the needed (chain, step) items densely (about
// half of the 15 x LEN), so no lane idles on a step the chain skips.  Item
// offsets come from a prefix sum across the lanes (one thread per chain) and
// an item -> chain map in LDS.
template <class F>
__global__ __launch_bounds__(256) void k_pre_need(const typename F::Key* __restrict__ keys,
                                                   const DLeaf* __restrict__ reqs,
                                                   const uint8_t* __restrict__ msgs,
                                                   uint8_t* __restrict__ scratch, size_t stride, uint32_t prio) {
    constexpr uint32_t LEN = F::LEN;
    static_assert(LEN <= 256, "one thread per chain");
    __shared__ uint32_t first[LEN + 1];  // first needed item of each chain; first[LEN] = all
    __shared__ uint32_t lens[LEN];
    __shared__ uint32_t wsum[4];        // needed steps per wave of chains
    __shared__ uint8_t owner[15 * LEN];  // item -> its chain
    wave_priority(prio);
    const uint32_t r = blockIdx.x;
    const DLeaf q = reqs[r];
    const uint8_t* msg = msgs + (size_t)r * F::kN;
    // the items' offsets: an exclusive prefix sum of the needed steps over the
    // chains, scanned across the lanes of each wave (shuffles), then the
    // earlier waves' totals added
    const uint32_t c = threadIdx.x, lane = c & 63;
    uint32_t need = 0;
    if (c < LEN) {
        lens[c] = chain_len<F>(msg, c);
        need = 15 - lens[c];
    }
    uint32_t incl = need;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t v = __shfl_up(incl, d, 64);
        if (lane >= d) incl += v;
    }
    if (lane == 63) wsum[c >> 6] = incl;
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t w = 0; w < (c >> 6); ++w) base += wsum[w];
    if (c < LEN) {
        const uint32_t f = base + incl - need;
        first[c] = f;
        for (uint32_t i = 0; i < need; ++i) owner[f + i] = (uint8_t)c;
        if (c == LEN - 1) first[LEN] = base + incl;
    }
    __syncthreads();
    const typename F::Key& k = keys[q.key];
    auto* fp = reinterpret_cast<typename F::FPre*>(scratch + (size_t)r * stride + LatLayout<F>::fpre);
    const uint32_t total = first[LEN];
    for (uint32_t t = threadIdx.x; t < total; t += blockDim.x) {
        const uint32_t ch = owner[t];
        const uint32_t s = lens[ch] + (t - first[ch]);
        uint32_t a[8] = {q.layer, q.tree_hi, q.tree_lo, ADDR_OTS, q.leaf, ch, s, 0};
        typename F::FPre p;
        F::pre_f(p, k, a);
        fp[ch * 15 + s] = p;
    }
}

// ---- LAT_LEAF finish for hundreds of leaves --------------------------
// k_lat_fin gives each leaf a workgroup of ceil(LEN/64) waves, the last one
// nearly empty; with many leaves the finish runs instead as one lane per
// chain (all 15 steps from the pre-images, full waves) and then one workgroup
// of ceil(LEN/2 / 64) waves per leaf for the L-tree through LDS.
template <class F>
__global__ __launch_bounds__(WG) void k_lat_chains_all(uint32_t n, const uint8_t* scratch, size_t stride,
                                                        uint8_t* __restrict__ ends) {
    using FPre = typename F::FPre;
    const uint32_t g = blockIdx.x * WG + threadIdx.x;
    if (g >= n * (uint32_t)F::LEN) return;
    const uint32_t r = g / F::LEN, c = g % F::LEN;
    const uint8_t* S = scratch + (size_t)r * stride;
    typename F::Node x = reinterpret_cast<const typename F::Node*>(S + LatLayout<F>::starts)[c];
    const FPre* fp = reinterpret_cast<const FPre*>(S + LatLayout<F>::fpre) + c * 15;
    FPre cur = fp[0];
    for (uint32_t s = 0; s < 15; ++s) {
        const FPre nxt = fp[s + 1 < 15 ? s + 1 : 14];
        keep_loads_ahead();  // (scratch is not restrict, so the load stays here)
        F::fin_f(x, cur, x);
        cur = nxt;
    }
    F::store(x, ends + (size_t)g * F::kN);
}

template <class F>
struct LtreeThreads {
    static constexpr uint32_t value = ((F::LEN / 2 + 63) / 64) * 64;
};

template <class F>
__global__ __launch_bounds__(LtreeThreads<F>::value) void k_lat_ltree(const uint8_t* __restrict__ ends,
                                                                      const uint8_t* __restrict__ scratch,
                                                                      size_t stride, uint8_t* __restrict__ out) {
    __shared__ typename F::Node pk[F::LEN];
    const uint32_t r = blockIdx.x;
    for (uint32_t c = threadIdx.x; c < (uint32_t)F::LEN; c += blockDim.x)
        F::load(pk[c], ends + ((size_t)r * F::LEN + c) * F::kN);
    __syncthreads();
    ltree_lds<F>(pk, reinterpret_cast<const typename F::HPre*>(scratch + (size_t)r * stride + LatLayout<F>::hpre),
                 threadIdx.x);
    if (threadIdx.x == 0) F::store(pk[0], out + (size_t)r * F::kN);
}

// ---- the same finish for large verification batches, lane-dense --------
// k_lat_fin gives each request a workgroup, whose L-tree heights and auth
// walk leave most lanes idle; for thousands of requests the finish runs as
// one lane per (request, chain), then per (request, L-tree pair) and height,
// then per request for the auth walk.

// WOTS chains of verify layers from their signatures: ends[r][c].  A chain
// runs 15 - d steps (d its message digit, uniform over 0..15), so a wave of
// arbitrary chains runs as long as its longest one, twice the average.  Each
// workgroup therefore takes 256 consecutive chains, sorts them by step count
// in LDS (counting sort) and hands them to its lanes in that order: each wave
// gets chains of about equal length.  Odd workgroups hand them out in
// reverse, so that the long waves do not all land on the same SIMD.
constexpr uint32_t LATB_CHAINS_WG = 256;
template <class F>
// (scratch is not __restrict__: restrict-qualified loads count as invariant
// and would move past keep_loads_ahead)
__global__ __launch_bounds__(LATB_CHAINS_WG) void k_latb_chains(uint32_t n, const uint8_t* __restrict__ msgs,
                                                                 const uint8_t* __restrict__ sigs,
                                                                 const uint8_t* scratch, size_t stride,
                                                                 uint8_t* __restrict__ ends, uint32_t prio) {
    using FPre = typename F::FPre;
    __shared__ uint32_t hist[16];
    wave_priority(prio);
    __shared__ uint32_t order[LATB_CHAINS_WG];
    __shared__ uint8_t first[LATB_CHAINS_WG];
    const uint32_t total = n * (uint32_t)F::LEN;
    const uint32_t base = blockIdx.x * LATB_CHAINS_WG;
    const uint32_t tid = threadIdx.x;
    // first step of this thread's own chain (15: none, also past the end)
    uint32_t d0 = 15;
    if (base + tid < total) {
        const uint32_t r0 = (base + tid) / F::LEN, c0 = (base + tid) % F::LEN;
        d0 = chain_len<F>(msgs + (size_t)r0 * F::kN, c0);
    }
    first[tid] = (uint8_t)d0;
    if (tid < 16) hist[tid] = 0;
    __syncthreads();
    const uint32_t rank = atomicAdd(&hist[d0], 1u);
    __syncthreads();
    uint32_t pos = rank;  // longest first: after every chain that starts earlier
    for (uint32_t b = 0; b < d0; ++b) pos += hist[b];
    order[pos] = tid;
    __syncthreads();
    const uint32_t mine = order[(blockIdx.x & 1) ? LATB_CHAINS_WG - 1 - tid : tid];
    const uint32_t g = base + mine;
    if (g >= total) return;
    const uint32_t r = g / F::LEN, c = g % F::LEN;
    const uint32_t s0 = first[mine];
    typename F::Node x;
    F::load(x, sigs + (size_t)g * F::kN);
    const FPre* fp = reinterpret_cast<const FPre*>(scratch + (size_t)r * stride + LatLayout<F>::fpre) + c * 15;
    // two pre-image registers in turn: the next step's pre-image is loaded
    // before the current step hashes and is waited for only at its use
    FPre a = fp[s0 < 15 ? s0 : 14], b;
    for (uint32_t s = s0; s < 15; s += 2) {
        b = fp[s + 1 < 15 ? s + 1 : 14];
        keep_loads_ahead();
        F::fin_f(x, a, x);
        if (s + 1 >= 15) break;
        a = fp[s + 2 < 15 ? s + 2 : 14];
        keep_loads_ahead();
        F::fin_f(x, b, x);
    }
    F::store(x, ends + (size_t)g * F::kN);
}

// The same chains, two per lane: each workgroup takes 2 x LATB_CHAINS_WG
// consecutive chains, counting-sorts them by step count and gives lane i the
// i-th longest and the i-th shortest, walked one after the other: every lane
// walks about 15 steps, and half as many waves as one chain per lane carry
// the same steps.
template <class F>
__device__ __forceinline__ void latb_walk(typename F::Node& x, const typename F::FPre* fp, uint32_t s0) {
    using FPre = typename F::FPre;
    FPre a = fp[s0 < 15 ? s0 : 14], b;
    for (uint32_t s = s0; s < 15; s += 2) {
        b = fp[s + 1 < 15 ? s + 1 : 14];
        keep_loads_ahead();
        F::fin_f(x, a, x);
        if (s + 1 >= 15) break;
        a = fp[s + 2 < 15 ? s + 2 : 14];
        keep_loads_ahead();
        F::fin_f(x, b, x);
    }
}

template <class F>
__global__ __launch_bounds__(LATB_CHAINS_WG) void k_latb_chains2(uint32_t n, const uint8_t* __restrict__ msgs,
                                                                  const uint8_t* __restrict__ sigs,
                                                                  const uint8_t* scratch, size_t stride,
                                                                  uint8_t* __restrict__ ends, uint32_t prio) {
    using FPre = typename F::FPre;
    constexpr uint32_t W = LATB_CHAINS_WG, C2 = 2 * LATB_CHAINS_WG;
    __shared__ uint32_t hist[16];
    __shared__ uint16_t order[C2];
    __shared__ uint8_t first[C2];
    wave_priority(prio);
    const uint32_t total = n * (uint32_t)F::LEN;
    const uint32_t base = blockIdx.x * C2;
    const uint32_t tid = threadIdx.x;
    if (tid < 16) hist[tid] = 0;
    __syncthreads();
    uint32_t d[2], rank[2];
    for (uint32_t h = 0; h < 2; ++h) {
        const uint32_t k = tid + h * W;
        d[h] = 15;  // past the end: no steps
        if (base + k < total) {
            const uint32_t r = (base + k) / F::LEN, c = (base + k) % F::LEN;
            d[h] = chain_len<F>(msgs + (size_t)r * F::kN, c);
        }
        first[k] = (uint8_t)d[h];
        rank[h] = atomicAdd(&hist[d[h]], 1u);
    }
    __syncthreads();
    for (uint32_t h = 0; h < 2; ++h) {
        uint32_t pos = rank[h];  // longest first
        for (uint32_t b = 0; b < d[h]; ++b) pos += hist[b];
        order[pos] = (uint16_t)(tid + h * W);
    }
    __syncthreads();
    for (uint32_t h = 0; h < 2; ++h) {
        const uint32_t k = order[h ? C2 - 1 - tid : tid];
        const uint32_t g = base + k;
        if (g >= total) continue;
        const uint32_t r = g / F::LEN, c = g % F::LEN;
        typename F::Node x;
        F::load(x, sigs + (size_t)g * F::kN);
        latb_walk<F>(x, reinterpret_cast<const FPre*>(scratch + (size_t)r * stride + LatLayout<F>::fpre) + c * 15,
                     first[k]);
        F::store(x, ends + (size_t)g * F::kN);
    }
}

// One L-tree height for every request: l nodes per request in src (src_stride
// nodes apart) -> ceil(l/2) in dst; merge `off + i` of the request's H
// pre-images hashes pair i; an odd last node is carried up.
template <class F>
__global__ __launch_bounds__(WG) void k_latb_ltree(uint32_t n, uint32_t l, uint32_t off,
                                                    const uint8_t* __restrict__ src, uint32_t src_stride,
                                                    uint8_t* __restrict__ dst, uint32_t dst_stride,
                                                    const uint8_t* __restrict__ scratch, size_t stride,
                                                    uint32_t prio) {
    using HPre = typename F::HPre;
    wave_priority(prio);
    const uint32_t half = (l + 1) >> 1;
    const uint32_t g = blockIdx.x * WG + threadIdx.x;
    if (g >= n * half) return;
    const uint32_t r = g / half, i = g % half;
    const uint8_t* s = src + (size_t)r * src_stride * F::kN;
    typename F::Node L, R, o;
    F::load(L, s + (size_t)(2 * i) * F::kN);
    if (2 * i + 1 < l) {
        F::load(R, s + (size_t)(2 * i + 1) * F::kN);
        const HPre* hp = reinterpret_cast<const HPre*>(scratch + (size_t)r * stride + LatLayout<F>::hpre);
        F::fin_h(o, hp[off + i], L, R);
    } else {
        o = L;
    }
    F::store(o, dst + ((size_t)r * dst_stride + i) * F::kN);
}

// The whole L-tree (xmss_commons.c:16-49) of LtreeBatch<F>::R verification
// requests per workgroup through LDS, in one launch instead of one per
// height: the chain ends are loaded once, and at each height the threads are
// packed densely over (request, pair), so that the waves of the shrinking
// heights drop out instead of idling.  With AUTH the first R lanes then walk
// their requests' auth paths (compute_root, xmss_commons.c:54-96) and write
// the roots; without, they write the L-tree roots (leaves) for k_latb_auth.
template <class F>
struct LtreeBatch {
    static constexpr uint32_t THREADS = 256;
    static constexpr uint32_t R = THREADS / (F::LEN / 2);  // every first-height pair has a lane
};

template <class F, bool AUTH>
__global__ __launch_bounds__(256) void k_latb_tree(const DLeaf* __restrict__ reqs, uint32_t n,
                                                   const uint8_t* __restrict__ ends, const uint8_t* auth,
                                                   const uint8_t* scratch, size_t stride,
                                                   uint8_t* __restrict__ out, uint32_t prio) {
    using Node = typename F::Node;
    using HPre = typename F::HPre;
    constexpr uint32_t LEN = F::LEN, R = LtreeBatch<F>::R, T = LtreeBatch<F>::THREADS;
    __shared__ Node pk[R * LEN];
    wave_priority(prio);
    const uint32_t r0 = blockIdx.x * R;
    const uint32_t nr = n - r0 < R ? n - r0 : R;
    const uint32_t t = threadIdx.x;
    for (uint32_t i = t; i < nr * LEN; i += T) F::load(pk[i], ends + ((size_t)r0 * LEN + i) * F::kN);
    __syncthreads();
    uint32_t l = LEN, off = 0;
    while (l > 1) {
        const uint32_t parents = l >> 1;
        // lane t: request j of the group, pair c; the idle lanes of a wave
        // that holds merges hash a copy of one of them (mirror_lanes)
        const uint32_t m = mirror_lanes(t, nr * parents);
        const uint32_t j = m / parents, c = m % parents;
        const bool act = t < nr * parents;  // then m == t
        Node o;
        if (m < nr * parents) {
            const HPre* hp = reinterpret_cast<const HPre*>(scratch + (size_t)(r0 + j) * stride + LatLayout<F>::hpre);
            F::fin_h(o, hp[off + c], pk[j * LEN + 2 * c], pk[j * LEN + 2 * c + 1]);
        }
        __syncthreads();
        if (act) pk[j * LEN + c] = o;
        if ((l & 1) && t < nr) pk[t * LEN + parents] = pk[t * LEN + l - 1];  // odd last node carried up
        __syncthreads();
        off += parents;
        l = (l + 1) >> 1;
    }
    // the auth walks: lane t < nr walks request t, the rest of its wave copies
    const uint32_t w = mirror_lanes(t, nr);
    if (w >= nr) return;
    const uint32_t r = r0 + w;
    Node node = pk[w * LEN];
    if (AUTH) {
        const DLeaf q = reqs[r];
        const HPre* hp = reinterpret_cast<const HPre*>(scratch + (size_t)r * stride + LatLayout<F>::hpre);
        uint32_t idx = q.leaf;
        const uint8_t* ap = auth + (size_t)q.p2 * F::kN;
        for (uint32_t i = 0; i < q.p0; ++i) {
            Node sib, l, rr;
            F::load(sib, ap + (size_t)i * F::kN);
            auth_pair<F>(l, rr, node, sib, idx & 1);
            F::fin_h(node, hp[LEN - 1 + i], l, rr);
            idx >>= 1;
        }
    }
    if (w == t) F::store(node, out + (size_t)r * F::kN);
}

// Auth walk (compute_root, xmss_commons.c:54-96) of every request, from its
// L-tree root at leaves[r * leaf_stride].
template <class F>
__global__ __launch_bounds__(WG) void k_latb_auth(const DLeaf* __restrict__ reqs, uint32_t n,
                                                   const uint8_t* __restrict__ leaves, uint32_t leaf_stride,
                                                   const uint8_t* auth, const uint8_t* scratch, size_t stride,
                                                   uint8_t* __restrict__ out, uint32_t prio) {
    using HPre = typename F::HPre;
    wave_priority(prio);
    const uint32_t r = blockIdx.x * WG + threadIdx.x;
    if (r >= n) return;
    const DLeaf q = reqs[r];
    const HPre* hp = reinterpret_cast<const HPre*>(scratch + (size_t)r * stride + LatLayout<F>::hpre);
    typename F::Node node;
    F::load(node, leaves + (size_t)r * leaf_stride * F::kN);
    uint32_t idx = q.leaf;
    const uint8_t* ap = auth + (size_t)q.p2 * F::kN;
    // each level's sibling and pre-image are loaded one level ahead
    const uint32_t levels = q.p0;
    typename F::Node sib, sib_next;
    HPre h, h_next;
    if (levels) {
        F::load(sib, ap);
        h = hp[F::LEN - 1];
    }
    for (uint32_t i = 0; i < levels; ++i) {
        const uint32_t j = i + 1 < levels ? i + 1 : i;
        F::load(sib_next, ap + (size_t)j * F::kN);
        h_next = hp[F::LEN - 1 + j];
        keep_loads_ahead();
        typename F::Node l, rr;
        auth_pair<F>(l, rr, node, sib, idx & 1);
        F::fin_h(node, h, l, rr);
        idx >>= 1;
        sib = sib_next;
        h = h_next;
    }
    F::store(node, out + (size_t)r * F::kN);
}

}  // namespace xm
