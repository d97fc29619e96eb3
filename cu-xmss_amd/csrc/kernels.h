// kernels.h -- XMSS hot-path kernels for gfx950, templated on the hash family
// (dev_family.h).  Each family is instantiated in its own translation unit
// (fam_*.hip) so the heavy, fully unrolled code compiles in parallel.
//
// Work mapping (wave64):
//   k_chains       one lane per (leaf, chain): PRF_keygen + 15 thash_f
//                  (gen_chain, cuxmss/wots.c:38-53, fed by expand_seed :14-29);
//                  grid.y = leaf run (key, layer, tree), so the key context is
//                  wave-uniform and is read through the scalar cache.
//   k_ltree_level  one lane per (leaf, node pair) of one L-tree height
//                  (xmss_commons.c:16-49), ping-ponging between two buffers.
//   k_level        one lane per parent node of one subtree height (the merges
//                  of treehash, xmss_core.c:62-83).
//   k_wots         one lane per (request, chain): wots_pkgen / wots_sign /
//                  wots_pk_from_sig (wots.c:117-175).
//   k_leaf_wg      one workgroup per leaf or per signature layer: chains in
//                  parallel, L-tree through LDS, auth-path walk (verify).
//   k_thash_h, k_hmsg: one lane per request.
// Node layers of a subtree are kept level-major in HBM: height j starts at node
// 2^(th+1) - 2^(th+1-j); one subtree spans 2^(th+1) - 1 nodes.
#pragma once
#include <hip/hip_runtime.h>

#include "dev_family.h"

namespace xm {

// device descriptors, 32 or 64 bytes, naturally aligned
struct DRun { uint32_t key, layer, tree_hi, tree_lo, leaf0, p0, p1, p2; };
struct DWots { uint32_t key, layer, tree_hi, tree_lo, ots, mode, p0, p1; };
struct DLeaf { uint32_t key, layer, tree_hi, tree_lo, leaf, p0, p1, p2; };
struct DThash { uint32_t key, adrs[8], p[7]; };
struct DMsg { uint32_t key, sign; uint64_t idx; uint64_t off; uint32_t len, p0; };

constexpr uint32_t ADDR_OTS = 0, ADDR_LTREE = 1, ADDR_HASHTREE = 2;
// hash_bytes reads whole blocks: every message blob a kernel hashes is
// followed by at least this many readable bytes (at most one SHAKE128 block
// of 168 bytes is read past a message's end).
constexpr size_t kMsgSlack = 256;
constexpr uint32_t MODE_PKGEN = 0, MODE_SIGN = 1, MODE_VERIFY = 2;
constexpr int WG = 64;

// Issue priority of the calling wave on its SIMD (s_setprio): kernels on a
// batch's critical path run at 2 so that, sharing a SIMD with the waves of
// wider work launched beside them, they are not slowed by it.  0 leaves the
// default.
__device__ __forceinline__ void wave_priority(uint32_t prio) {
    if (prio) __builtin_amdgcn_s_setprio(2);
}

// Lane mirroring.  A wave whose first lanes hold the last of n serial work
// items (a chain walk, an H_msg, a tree merge) and whose other lanes have
// none runs measurably slower and less evenly than a full wave running the
// same instructions, so the idle lanes redo an item of their wave, and only
// the item's own lane (i < n) writes the result.  i is the lane's item; the
// items of a wave are 64 consecutive indices starting at a multiple of 64.
// Returns the item to compute, or n when the whole wave is past the end.
__device__ __forceinline__ uint32_t mirror_lanes(uint32_t i, uint32_t n) {
    if (i < n) return i;
    const uint32_t base = i & ~63u;
    return base < n ? base + (i - base) % (n - base) : n;
}

// Keeps a node in vector registers.  When every lane of a wave holds the
// same node (a lone request's auth walk, its idle lanes mirroring lane 0),
// the compiler proves the hashing wave-uniform and moves much of it onto the
// scalar unit, which runs it at about half the speed (an auth walk measured
// 8.2 against 4.3 us per level); values made opaque here stay per lane.
template <class N>
__device__ __forceinline__ void keep_in_vgprs(N& x) {
#pragma unroll
    for (int i = 0; i < (int)(sizeof(x.w) / sizeof(x.w[0])); ++i) asm volatile("" : "+v"(x.w[i]));
}

// The children of an auth-path merge in the order the index bit gives:
// (node, sib) for a left child, (sib, node) for a right one, selected word by
// word.  Written as two calls behind a branch on the bit, a wave whose lanes
// disagree -- or whose bit the compiler cannot prove uniform -- runs the hash
// twice, once per side.
template <class F>
__device__ __forceinline__ void auth_pair(typename F::Node& l, typename F::Node& r, const typename F::Node& node,
                                          const typename F::Node& sib, uint32_t right) {
    constexpr int W = sizeof(node.w) / sizeof(node.w[0]);
#pragma unroll
    for (int i = 0; i < W; ++i) {
        l.w[i] = right ? sib.w[i] : node.w[i];
        r.w[i] = right ? node.w[i] : sib.w[i];
    }
}

// A compiler barrier for memory operations: loads issued before it stay
// before it, so a prefetch cannot be sunk next to its first use (the
// hardware waits for a load only where its registers are read).  Loads
// through __restrict__ pointers count as invariant and may still move.
__device__ __forceinline__ void keep_loads_ahead() { asm volatile("" ::: "memory"); }

template <class F>
__global__ __launch_bounds__(WG) void k_prepare(const uint8_t* __restrict__ seeds, int nkeys,
                                                 typename F::Key* __restrict__ keys) {
    const int i = blockIdx.x * WG + threadIdx.x;
    if (i >= nkeys) return;
    typename F::Key k;
    typename F::Node t;
    const uint8_t* s = seeds + (size_t)i * 3 * F::kN;
    F::load(t, s);
#pragma unroll
    for (int j = 0; j < F::NW; ++j) k.sks[j] = t.w[j];
    F::load(t, s + F::kN);
#pragma unroll
    for (int j = 0; j < F::NW; ++j) k.prfk[j] = t.w[j];
    F::load(t, s + 2 * F::kN);
#pragma unroll
    for (int j = 0; j < F::NW; ++j) k.pub[j] = t.w[j];
    F::prepare(k);
    keys[i] = k;
}

// gen_chain (wots.c:38-53): `steps` applications of thash_f from position
// `start`; a[0..5] hold the chain's ADRS.  Families that can share the
// ADRS-invariant PRF rounds along the chain do so.
template <class F>
__device__ __forceinline__ void run_chain(typename F::Node& x, const typename F::Key& k, uint32_t (&a)[8],
                                          uint32_t start, uint32_t steps) {
    if constexpr (F::kChainPrf) {
        typename F::ChainPrf cp;
        F::chain_prf_init(cp, k, a);
        for (uint32_t i = start; i < start + steps; ++i) {
            typename F::Node key, mask;
            F::chain_prf_step(key, mask, k, cp, i);
            F::f_outer(x, key, mask, x);
        }
    } else {
        for (uint32_t i = start; i < start + steps; ++i) {
            a[6] = i;
            F::thash_f(x, k, a, x);
        }
    }
}

// Chain ends of `leaves` consecutive leaves per run, from run.leaf0 + leaf_off.
template <class F>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(F::kChainWaves)))
void k_chains(const typename F::Key* __restrict__ keys,
                                                const DRun* __restrict__ runs, uint32_t leaf_off,
                                                uint32_t leaves, uint8_t* __restrict__ ends) {
    const uint32_t r = blockIdx.y;
    const uint32_t item = blockIdx.x * WG + threadIdx.x;
    if (item >= leaves * (uint32_t)F::LEN) return;
    const uint32_t leaf = item / F::LEN, chain = item % F::LEN;
    const DRun run = runs[r];
    const typename F::Key& k = keys[run.key];
    uint32_t a[8] = {run.layer, run.tree_hi, run.tree_lo, ADDR_OTS, run.leaf0 + leaf_off + leaf,
                     chain, 0, 0};
    typename F::Node x;
    F::prf_keygen(x, k, a);
    run_chain<F>(x, k, a, 0, 15);
    F::store(x, ends + ((size_t)r * leaves * F::LEN + item) * F::kN);
}

// One height of the L-tree (xmss_commons.c:16-49) for `leaves` leaves per run,
// flat over (leaf, parent) items.  The `l` nodes per leaf at this height are
// read from src (src_stride nodes per leaf) and ceil(l/2) are written to dst;
// the item holding an odd last node only carries it up.  The last height
// (l == 2) writes the leaf into `out` (out_stride nodes per run, at out_base).
template <class F>
__global__ __launch_bounds__(WG) void k_ltree_level(const typename F::Key* __restrict__ keys,
                                                     const DRun* __restrict__ runs, uint32_t leaf_off,
                                                     uint32_t leaves, uint32_t l, uint32_t height,
                                                     const uint8_t* __restrict__ src, uint32_t src_stride,
                                                     uint8_t* __restrict__ dst, uint32_t dst_stride,
                                                     uint8_t* __restrict__ out, uint64_t out_stride,
                                                     uint64_t out_base) {
    const uint32_t r = blockIdx.y;
    const uint32_t half = (l + 1) >> 1;
    const uint32_t item = blockIdx.x * WG + threadIdx.x;
    if (item >= leaves * half) return;
    const uint32_t leaf = item / half, i = item % half;
    const uint8_t* s = src + ((size_t)r * leaves + leaf) * src_stride * F::kN;
    typename F::Node L, R, o;
    F::load(L, s + (size_t)(2 * i) * F::kN);
    if (2 * i + 1 < l) {
        const DRun run = runs[r];
        F::load(R, s + (size_t)(2 * i + 1) * F::kN);
        uint32_t a[8] = {run.layer, run.tree_hi, run.tree_lo, ADDR_LTREE, run.leaf0 + leaf_off + leaf,
                         height, i, 0};
        F::thash_h(o, keys[run.key], a, L, R);
    } else {
        o = L;  // odd last node: carried up unhashed
    }
    if (half == 1) {
        F::store(o, out + ((size_t)r * out_stride + out_base + leaf) * F::kN);
    } else {
        F::store(o, dst + (((size_t)r * leaves + leaf) * dst_stride + i) * F::kN);
    }
}

// Parents at height `level` + 1 from children at `level`, all runs.
template <class F>
__global__ __launch_bounds__(WG) void k_level(const typename F::Key* __restrict__ keys,
                                               const DRun* __restrict__ runs, uint32_t th,
                                               uint32_t level, uint8_t* __restrict__ nodes) {
    const uint32_t r = blockIdx.y;
    const uint32_t i = blockIdx.x * WG + threadIdx.x;
    if (i >= (1u << (th - level - 1))) return;
    const DRun run = runs[r];
    const uint64_t stride = (2ull << th) - 1;
    const uint64_t base_c = (2ull << th) - (2ull << (th - level));
    const uint64_t base_p = (2ull << th) - (2ull << (th - level - 1));
    uint8_t* T = nodes + (size_t)r * stride * F::kN;
    typename F::Node L, R, o;
    F::load(L, T + (base_c + 2 * i) * F::kN);
    F::load(R, T + (base_c + 2 * i + 1) * F::kN);
    uint32_t a[8] = {run.layer, run.tree_hi, run.tree_lo, ADDR_HASHTREE, 0, level, i, 0};
    F::thash_h(o, keys[run.key], a, L, R);
    F::store(o, T + (base_p + i) * F::kN);
}

// base-w digit c of a WOTS message: base_w + wots_checksum (wots.c:60-107), w = 16.
template <class F>
__device__ uint32_t chain_len(const uint8_t* msg, uint32_t c) {
    constexpr uint32_t len1 = 2 * F::kN;
    if (c < len1) {
        const uint32_t byte = msg[c >> 1];
        return (c & 1) ? (byte & 15u) : (byte >> 4);
    }
    uint32_t csum = 0;
    for (uint32_t i = 0; i < (uint32_t)F::kN; ++i) {
        const uint32_t byte = msg[i];
        csum += 30u - (byte >> 4) - (byte & 15u);
    }
    // (csum << 4) as two big-endian bytes, read back as three base-16 digits
    const uint32_t d = c - len1;
    return (csum >> (8 - 4 * d)) & 15u;
}

template <class F>
__global__ __launch_bounds__(WG) void k_wots(const typename F::Key* __restrict__ keys,
                                              const DWots* __restrict__ reqs, uint32_t n,
                                              const uint8_t* __restrict__ msgs,
                                              const uint8_t* __restrict__ in,
                                              uint8_t* __restrict__ out) {
    const uint32_t g = blockIdx.x * WG + threadIdx.x;
    if (g >= n * (uint32_t)F::LEN) return;
    const uint32_t r = g / F::LEN, c = g % F::LEN;
    const DWots q = reqs[r];
    const typename F::Key& k = keys[q.key];
    uint32_t a[8] = {q.layer, q.tree_hi, q.tree_lo, ADDR_OTS, q.ots, c, 0, 0};
    typename F::Node x;
    uint32_t start = 0, steps = 15;
    if (q.mode == MODE_VERIFY) {
        const uint32_t len = chain_len<F>(msgs + (size_t)r * F::kN, c);
        start = len;
        steps = 15 - len;
        F::load(x, in + (size_t)g * F::kN);
    } else {
        if (q.mode == MODE_SIGN) steps = chain_len<F>(msgs + (size_t)r * F::kN, c);
        F::prf_keygen(x, k, a);
    }
    run_chain<F>(x, k, a, start, steps);
    F::store(x, out + (size_t)g * F::kN);
}

// One workgroup per request, ceil(LEN/64) waves: the LEN chains in parallel
// (one lane each), the L-tree height by height through LDS, and for VERIFY the
// auth-path walk (compute_root, xmss_commons.c:54-96) in lane 0.
//   VERIFY == false: gen_leaf_wots (xmss_commons.c:104-113)          -> leaf
//   VERIFY == true:  one layer of xmssmt_core_sign_open (:362-375):
//                    wots_pk_from_sig, l_tree, compute_root          -> root
template <class F, bool VERIFY>
__global__ __launch_bounds__(192) void k_leaf_wg(const typename F::Key* __restrict__ keys, const DLeaf* __restrict__ reqs,
                          const uint8_t* __restrict__ msgs, const uint8_t* __restrict__ sigs,
                          const uint8_t* __restrict__ auth, uint32_t th, uint8_t* __restrict__ out) {
    __shared__ typename F::Node pk[F::LEN];
    const uint32_t r = blockIdx.x;
    const uint32_t c = threadIdx.x;
    const DLeaf q = reqs[r];
    const typename F::Key& k = keys[q.key];
    if (c < (uint32_t)F::LEN) {
        uint32_t a[8] = {q.layer, q.tree_hi, q.tree_lo, ADDR_OTS, q.leaf, c, 0, 0};
        typename F::Node x;
        uint32_t start = 0, steps = 15;
        if (VERIFY) {
            const uint32_t len = chain_len<F>(msgs + (size_t)r * F::kN, c);
            start = len;
            steps = 15 - len;
            F::load(x, sigs + ((size_t)r * F::LEN + c) * F::kN);
        } else {
            F::prf_keygen(x, k, a);
        }
        run_chain<F>(x, k, a, start, steps);
        pk[c] = x;
    }
    __syncthreads();
    uint32_t l = F::LEN, height = 0;
    while (l > 1) {
        const uint32_t parents = l >> 1;
        typename F::Node o;
        if (c < parents) {
            uint32_t a[8] = {q.layer, q.tree_hi, q.tree_lo, ADDR_LTREE, q.leaf, height, c, 0};
            F::thash_h(o, k, a, pk[2 * c], pk[2 * c + 1]);
        }
        __syncthreads();
        if (c < parents) pk[c] = o;
        if ((l & 1) && c == 0) pk[l >> 1] = pk[l - 1];
        __syncthreads();
        l = (l & 1) ? (l >> 1) + 1 : (l >> 1);
        ++height;
    }
    if (c != 0) return;
    typename F::Node node = pk[0];
    if (VERIFY) {
        uint32_t a[8] = {q.layer, q.tree_hi, q.tree_lo, ADDR_HASHTREE, 0, 0, 0, 0};
        uint32_t idx = q.leaf;
        const uint8_t* ap = auth + (size_t)r * th * F::kN;
        for (uint32_t i = 0; i < th; ++i) {
            typename F::Node sib;
            F::load(sib, ap + (size_t)i * F::kN);
            a[5] = i;
            a[6] = idx >> 1;
            typename F::Node l, rr;
            auth_pair<F>(l, rr, node, sib, idx & 1);
            F::thash_h(node, k, a, l, rr);
            idx >>= 1;
        }
    }
    F::store(node, out + (size_t)r * F::kN);
}

template <class F>
__global__ __launch_bounds__(WG) void k_thash_h(const typename F::Key* __restrict__ keys,
                                                 const DThash* __restrict__ reqs, uint32_t n,
                                                 const uint8_t* __restrict__ in,
                                                 uint8_t* __restrict__ out) {
    const uint32_t r = blockIdx.x * WG + threadIdx.x;
    if (r >= n) return;
    const DThash q = reqs[r];
    uint32_t a[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = q.adrs[i];
    typename F::Node L, R, o;
    F::load(L, in + (size_t)r * 2 * F::kN);
    F::load(R, in + (size_t)r * 2 * F::kN + F::kN);
    F::thash_h(o, keys[q.key], a, L, R);
    F::store(o, out + (size_t)r * F::kN);
}

// R (sign requests) and H_msg over the prefixed message in `blob`
// (hash.c:107-120 with R = PRF(SK_PRF, toByte(idx, 32)), xmss_core.c:242-248).
template <class F>
__device__ __forceinline__ void hmsg_one(const typename F::Key* __restrict__ keys, const DMsg* __restrict__ reqs,
                                         uint32_t r, uint32_t pad, uint8_t* __restrict__ blob,
                                         uint8_t* __restrict__ R_out, uint8_t* __restrict__ dig_out,
                                         bool own = true) {
    const DMsg q = reqs[r];
    uint8_t* p = blob + q.off;
    if (q.sign) {
        typename F::Node R;
        F::prf_idx(R, keys[q.key], q.idx);
        if (own) {
            F::store(R, p + pad);
            F::store(R, R_out + (size_t)r * F::kN);
        }
    }
    typename F::Node d;
    F::hash_bytes(d, p, q.len);
    if (own) F::store(d, dig_out + (size_t)r * F::kN);
}

// hmsg_one for item i of n, with the idle lanes of the last items' wave
// mirroring them (mirror_lanes).  A mirroring lane may read the message
// before its owner has written R into it: its digest is discarded.
template <class F>
__device__ __forceinline__ void hmsg_mirrored(const typename F::Key* __restrict__ keys,
                                              const DMsg* __restrict__ reqs, uint32_t i, uint32_t n,
                                              uint32_t pad, uint8_t* __restrict__ blob,
                                              uint8_t* __restrict__ R_out, uint8_t* __restrict__ dig_out) {
    const uint32_t r = mirror_lanes(i, n);
    if (r < n) hmsg_one<F>(keys, reqs, r, pad, blob, R_out, dig_out, r == i);
}

template <class F>
__global__ __launch_bounds__(WG) void k_hmsg(const typename F::Key* __restrict__ keys,
                                              const DMsg* __restrict__ reqs, uint32_t n, uint32_t pad,
                                              uint8_t* __restrict__ blob, uint8_t* __restrict__ R_out,
                                              uint8_t* __restrict__ dig_out, uint32_t prio) {
    wave_priority(prio);
    hmsg_mirrored<F>(keys, reqs, blockIdx.x * WG + threadIdx.x, n, pad, blob, R_out, dig_out);
}

// H_msg of up to 64 requests [first, min(first + 64, n)) by the two waves of
// a 128-thread workgroup, for the SHA-256 families.  H_msg is a chain of
// compressions, one per 64 bytes, but the message schedule of every block
// depends only on the message: wave 0 computes block j+1's W[t] + K[t] into
// LDS (its message words loaded one block ahead) while wave 1 runs the 64
// rounds of block j from LDS.  The serial path per block drops from a whole
// compression (about 1400 VALU instructions) to its rounds (about 900).
// Sign requests get R = PRF(SK_PRF, idx) first (wave 0, written into the
// prefix as hmsg_one does).  Every thread of the workgroup must call it.
// With mid set, each request resumes from the chaining value hmsg_head left
// (its R is in the prefix already).
struct HmsgPairLds {
    uint32_t wk[2][64][64];  // [buffer][round][message]: one word per lane, conflict-free
    uint32_t nbmax, b0min;
};

// The chaining value of an H_msg after its first nb blocks.  A launch
// sequence of one signature of a long message hashes R and the first
// kHmsgHeadBlocks blocks in its pre-image launch (hmsg_head, beside the
// address-only work) and the finishing launch resumes from there, so that
// this much of H_msg's serial chain is done before the finishing launch
// starts.  The head is sized to end within the pre-image launch's own
// length (one wave at a whole compression per block: 6 blocks made that
// launch longer and 1.6 KB signatures 5 us slower).
struct HmsgMid {
    uint32_t st[8];
    uint32_t nb;
};
constexpr uint32_t kHmsgHeadBlocks = 3;

// R and the first `head` blocks of H_msg of request r (SHA-256 families), one
// lane per request; only the owning lane writes (own: see hmsg_mirrored).
// A request whose blocks all fit also gets its digest.
template <class F>
__device__ __forceinline__ void hmsg_head(const typename F::Key* __restrict__ keys, const DMsg* __restrict__ reqs,
                                          uint32_t r, uint32_t pad, uint8_t* __restrict__ blob,
                                          uint8_t* __restrict__ R_out, uint8_t* __restrict__ dig_out,
                                          HmsgMid* __restrict__ mid, uint32_t head, bool own) {
    const DMsg q = reqs[r];
    uint8_t* p = blob + q.off;
    if (q.sign) {
        typename F::Node R;
        F::prf_idx(R, keys[q.key], q.idx);
        if (own) {
            F::store(R, p + pad);
            F::store(R, R_out + (size_t)r * F::kN);
        }
    }
    uint32_t st[8];
    sha256_iv(st);
    const uint32_t nb = (q.len + 9 + 63) / 64, b1 = head < nb ? head : nb;
    for (uint32_t b = 0; b < b1; ++b) {
        uint4 raw[4];
        sha256_block_load(raw, p, b);
        uint32_t w[16];
        sha256_block_words(w, raw, q.len, b);
        sha256_compress(st, w);
    }
    if (!own) return;
#pragma unroll
    for (int i = 0; i < 8; ++i) mid[r].st[i] = st[i];
    mid[r].nb = b1;
    if (b1 == nb) {
        typename F::Node d;
#pragma unroll
        for (int i = 0; i < F::NW; ++i) d.w[i] = st[i];
        F::store(d, dig_out + (size_t)r * F::kN);
    }
}

template <class F>
__device__ __forceinline__ void hmsg_pair_block(const typename F::Key* __restrict__ keys,
                                                const DMsg* __restrict__ reqs, uint32_t first, uint32_t n,
                                                uint32_t pad, uint8_t* blob, uint8_t* __restrict__ R_out,
                                                uint8_t* __restrict__ dig_out, HmsgPairLds& sh,
                                                const HmsgMid* mid = nullptr) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // Lanes past the last request hash a copy of request lane % count: with
    // few messages (a single BDS signature's) the waves run at the speed they
    // reach with every lane active, measured 1.5x the speed with one active
    // lane; only the lanes of the requests themselves write.
    const uint32_t count = first < n ? (n - first < 64 ? n - first : 64) : 0;
    const bool own = lane < count;
    const uint32_t r = count ? first + lane % count : n;
    uint32_t len = 0, nb = 0, b0 = 0;
    uint8_t* p = nullptr;
    uint32_t st[8];
    sha256_iv(st);
    if (r < n) {
        const DMsg q = reqs[r];
        p = blob + q.off;
        len = q.len;
        nb = (len + 9 + 63) / 64;
        if (mid) {  // resumed after hmsg_head
            b0 = mid[r].nb;
#pragma unroll
            for (int i = 0; i < 8; ++i) st[i] = mid[r].st[i];
        } else if (q.sign && wave == 0 && own) {  // R, then the prefix holds it (xmss_core.c:242-248)
            typename F::Node R;
            F::prf_idx(R, keys[q.key], q.idx);
            F::store(R, p + pad);
            F::store(R, R_out + (size_t)r * F::kN);
        }
    }
    if (threadIdx.x == 0) {
        sh.nbmax = 0;
        sh.b0min = 0xFFFFFFFFu;
    }
    __syncthreads();  // (also orders the R stores before any lane's loads of the prefix)
    if (wave == 0) {
        atomicMax(&sh.nbmax, nb);
        atomicMin(&sh.b0min, b0);
    }
    __syncthreads();
    const uint32_t NB = sh.nbmax, B0 = sh.b0min;
    uint4 raw[4];
    auto schedule = [&](uint32_t blk) {
        uint32_t w[16];
        sha256_block_words(w, raw, len, blk);
        if (blk + 1 < nb) sha256_block_load(raw, p, blk + 1);
        keep_loads_ahead();
        sha256_schedule_k(w, [&](int t, uint32_t v) { sh.wk[blk & 1][t][lane] = v; });
    };
    if (wave == 0 && b0 < nb) {
        sha256_block_load(raw, p, b0);
        schedule(b0);
    }
    __syncthreads();
    // each lane's blocks are [b0, nb): wave 0 schedules block blk + 1 while
    // wave 1 compresses block blk
    for (uint32_t blk = B0; blk < NB; ++blk) {
        if (wave == 0) {
            if (blk >= b0 && blk + 1 < nb) schedule(blk + 1);
        } else if (wave == 1 && blk >= b0 && blk < nb) {
            // all 64 round inputs into registers first: the LDS reads are in
            // flight together instead of one per round
            uint32_t x[64];
#pragma unroll
            for (int t = 0; t < 64; ++t) x[t] = sh.wk[blk & 1][t][lane];
            sha256_compress_wk(st, [&](int t) { return x[t]; });
        }
        __syncthreads();
    }
    if (wave == 1 && own) {
        typename F::Node d;
#pragma unroll
        for (int i = 0; i < F::NW; ++i) d.w[i] = st[i];
        F::store(d, dig_out + (size_t)r * F::kN);
    }
}

// H_msg of n requests, 64 per 128-thread workgroup (SHA-256 families).
template <class F>
__global__ __launch_bounds__(128) void k_hmsg_pair(const typename F::Key* __restrict__ keys,
                                                   const DMsg* __restrict__ reqs, uint32_t n, uint32_t pad,
                                                   uint8_t* blob, uint8_t* __restrict__ R_out,
                                                   uint8_t* __restrict__ dig_out, uint32_t prio) {
    __shared__ HmsgPairLds sh;
    wave_priority(prio);
    hmsg_pair_block<F>(keys, reqs, blockIdx.x * 64, n, pad, blob, R_out, dig_out, sh);
}

__global__ __launch_bounds__(WG) void k_gather(const uint8_t* __restrict__ src,
                                               const uint64_t* __restrict__ idx, uint32_t n,
                                               uint32_t nbytes, uint8_t* __restrict__ dst);

}  // namespace xm
