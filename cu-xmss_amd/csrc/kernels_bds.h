// kernels_bds.h -- the BDS traversal state machine on the device.
//
// One workgroup per signature (ceil(LEN/64) waves) advances that key's
// serialized state in place, in HBM: bds_round, bds_treehash_update,
// treehash_update, bds_state_update and deep_state_swap of
// cuxmss/xmss_core_fast.c:165-626,917-1121, working directly on the sk bytes
// (the reference's (de)serialize pair, :38-136, maps the same fields; reading
// and writing them in place is equivalent because nothing else touches them
// during a signature).  The leaves the signature consumes were computed
// before, in the order a host dry run of the same code requested them, so the
// kernel reads them from a cursor.  Control flow is uniform across the
// workgroup (every thread reads the same bytes); every write is made by
// thread 0, or by the chain lanes for a WOTS signature, and followed by a
// barrier.  thash_h runs on thread 0.  The serialized states are staged in
// LDS for the launch (at most 28 KB for the largest parameter set); the
// lower-layer WOTS signatures stay in HBM.
#pragma once
#include "kernels.h"
#include "kernels_lat.h"

namespace xm {

struct BdsP {
    uint32_t th, d, full_height, index_bytes, bds_k, wots_sig_bytes, sig_bytes;
    uint32_t state_bytes;  // one serialized bds_state
};

struct DBds {
    uint64_t sk_off;      // OID-less sk in the sk arena
    uint64_t leaf_off;    // first planned leaf (node index) in the leaf arena
    uint64_t tail_off;    // signature bytes after R and the layer-0 WOTS signature
    uint64_t idx;         // index being signed
    uint32_t key;         // key context
    uint32_t nmerge;      // planned merges of this job
    uint64_t merge_off;   // first planned merge (F::HPre) in the merge arena
    BdsP P;               // parameter shape of this key
    uint64_t res;         // device address of the key's resident sk (index not advanced), or 0
    uint32_t sk_bytes;    // OID-less sk size
    uint32_t maxlev;      // levels of the merge DAG (0: merges hashed as the state machine reaches them)
    uint64_t lit_off;     // first literal (n bytes each) in the literal arena
    uint32_t wots_row;    // first pre-image row of its upper-layer WOTS signatures
    uint32_t nwots;       // planned upper-layer WOTS signatures (beyond: whole chains)
    uint32_t nleaf;       // planned leaves (from leaf_off)
    uint32_t lv_off;      // LDS offset their values are staged at, or 0 (read from HBM)
};

// dst[0, len) = src[0, len), by every thread of the workgroup (no barrier),
// with several loads in flight per thread: a strided loop of one-byte copies
// waits out a whole memory latency per iteration (the compiler issues the
// next load only after the store), which over the kilobytes of an XMSS^MT sk
// came to tens of microseconds per signature.  Word copies when the two
// pointers share their alignment mod 4, the rest byte by byte; the ranges
// must not overlap.
__device__ __forceinline__ void wg_copy(uint8_t* dst, const uint8_t* src, uint32_t len) {
    constexpr uint32_t U = 8;
    const uint32_t t = threadIdx.x, T = blockDim.x;
    uint32_t done = 0;
    if ((((uintptr_t)dst ^ (uintptr_t)src) & 3) == 0) {
        uint32_t head = (uint32_t)(0u - (uint32_t)(uintptr_t)src) & 3u;
        if (head > len) head = len;
        if (t < head) dst[t] = src[t];
        const uint32_t nw = (len - head) >> 2;
        uint32_t* d = reinterpret_cast<uint32_t*>(dst + head);
        const uint32_t* w = reinterpret_cast<const uint32_t*>(src + head);
        for (uint32_t base = 0; base < nw; base += U * T) {
            uint32_t v[U];
#pragma unroll
            for (uint32_t k = 0; k < U; ++k) {
                const uint32_t i = base + k * T + t;
                v[k] = w[i < nw ? i : nw - 1];
            }
#pragma unroll
            for (uint32_t k = 0; k < U; ++k) {
                const uint32_t i = base + k * T + t;
                if (i < nw) d[i] = v[k];
            }
        }
        done = head + 4 * nw;
    }
    const uint32_t rest = len - done;
    const uint8_t* sb = src + done;
    uint8_t* db = dst + done;
    for (uint32_t base = 0; base < rest; base += 4 * U * T) {
        uint8_t v[4 * U];
#pragma unroll
        for (uint32_t k = 0; k < 4 * U; ++k) {
            const uint32_t i = base + k * T + t;
            v[k] = sb[i < rest ? i : rest - 1];
        }
#pragma unroll
        for (uint32_t k = 0; k < 4 * U; ++k) {
            const uint32_t i = base + k * T + t;
            if (i < rest) db[i] = v[k];
        }
    }
}

// Sources of a planned merge's halves (kind in the top 4 bits, index below).
enum : uint32_t { DAG_LIT = 0u << 28, DAG_LEAF = 1u << 28, DAG_MERGE = 2u << 28, DAG_INDEX = (1u << 28) - 1 };

template <class F>
struct BdsDev {
    uint8_t* sk;       // the serialized states, staged in LDS for the launch
    uint8_t* sk_glob;  // the OID-less sk in HBM (lower-layer WOTS signatures live there)
    const BdsP& P;
    const uint8_t* leaves;
    uint32_t cur;
    const typename F::HPre* merges;  // the planned merges' address-only work, in hashing order
    uint32_t mcur, nmerge;
    const uint8_t* mval = nullptr;   // the planned merges' values, when computed up front (LDS)
    const uint8_t* wpre = nullptr;   // pre-image rows of the planned upper-layer WOTS signatures
    size_t wstride = 0;
    uint32_t wcur = 0, nwots = 0;
    const typename F::Key& key;
    uint32_t tid;

    static constexpr uint32_t n = F::kN;

    __device__ BdsDev(uint8_t* lds, uint8_t* skg, const BdsP& P_, const uint8_t* lv, const typename F::HPre* mg,
                      uint32_t nmg, const typename F::Key& k)
        : sk(lds), sk_glob(skg), P(P_), leaves(lv), cur(0), merges(mg), mcur(0), nmerge(nmg), key(k),
          tid(threadIdx.x) {}

    // ---- layout (xmss_core_fast.c:38-78), offsets within the states region ----
    __device__ size_t st(uint32_t i) const { return (size_t)i * P.state_bytes; }
    __device__ size_t states_base() const { return P.index_bytes + 4 * n; }
    __device__ size_t o_stack(uint32_t i) const { return st(i); }
    __device__ size_t o_stackoffset(uint32_t i) const { return st(i) + (size_t)(P.th + 1) * n; }
    __device__ size_t o_stacklevels(uint32_t i) const { return o_stackoffset(i) + 4; }
    __device__ size_t o_auth(uint32_t i) const { return o_stacklevels(i) + P.th + 1; }
    __device__ size_t o_keep(uint32_t i) const { return o_auth(i) + (size_t)P.th * n; }
    __device__ size_t o_th(uint32_t i, uint32_t j) const { return o_keep(i) + (size_t)(P.th >> 1) * n + (size_t)j * (7 + n); }
    __device__ size_t o_retain(uint32_t i) const { return o_th(i, P.th - P.bds_k); }
    __device__ size_t o_next_leaf(uint32_t i) const {
        return o_retain(i) + (size_t)((1u << P.bds_k) - P.bds_k - 1) * n;
    }
    __device__ uint8_t* wots_sigs() const { return sk_glob + states_base() + (size_t)(2 * P.d - 1) * P.state_bytes; }

    __device__ uint32_t rd(size_t off, int bytes) const {
        uint32_t v = 0;
        for (int b = 0; b < bytes; ++b) v = (v << 8) | sk[off + b];
        return v;
    }
    __device__ void wr(size_t off, int bytes, uint32_t v) {
        wr_nosync(off, bytes, v);
        __syncthreads();
    }
    // the same write without the barrier: a run of them ends in one
    __device__ void wr_nosync(size_t off, int bytes, uint32_t v) {
        if (tid == 0)
            for (int b = bytes - 1; b >= 0; --b) {
                sk[off + b] = (uint8_t)v;
                v >>= 8;
            }
    }
    // field accessors
    __device__ uint32_t stackoffset(uint32_t i) const { return rd(o_stackoffset(i), 4); }
    __device__ void set_stackoffset(uint32_t i, uint32_t v) { wr(o_stackoffset(i), 4, v); }
    __device__ uint32_t level(uint32_t i, uint32_t k) const { return sk[o_stacklevels(i) + k]; }
    __device__ void set_level(uint32_t i, uint32_t k, uint32_t v) { wr(o_stacklevels(i) + k, 1, v); }
    __device__ uint32_t th_h(uint32_t i, uint32_t j) const { return sk[o_th(i, j)]; }
    __device__ uint32_t th_next(uint32_t i, uint32_t j) const { return rd(o_th(i, j) + 1, 4); }
    __device__ uint32_t th_usage(uint32_t i, uint32_t j) const { return sk[o_th(i, j) + 5]; }
    __device__ uint32_t th_done(uint32_t i, uint32_t j) const { return sk[o_th(i, j) + 6]; }
    __device__ size_t th_node(uint32_t i, uint32_t j) const { return o_th(i, j) + 7; }
    __device__ uint32_t next_leaf(uint32_t i) const { return rd(o_next_leaf(i), 4); }

    // node copies (n bytes, disjoint ranges), one byte per lane, then a barrier
    __device__ void copy(uint8_t* dst, const uint8_t* src) {
        if (tid < n) dst[tid] = src[tid];
        __syncthreads();
    }
    // swap of two disjoint byte ranges of the states
    __device__ void swap(size_t a, size_t b, size_t len) {
        for (size_t x = tid; x < len; x += blockDim.x) {
            const uint8_t t = sk[a + x];
            sk[a + x] = sk[b + x];
            sk[b + x] = t;
        }
        __syncthreads();
    }
    // dst <- thash_h(in[0..2n)) at ADRS a (thread 0); dst may alias in.  The
    // merges come in the order the host's dry run recorded them, so the
    // address-only part (PRF key and masks, first outer block) was computed
    // in a wide launch before this one (F::pre_h); past the planned count the
    // whole thash_h runs here.
    __device__ void hash_h(uint8_t* dst, const uint8_t* in, const uint32_t (&a)[8]) {
        if (mval && mcur < nmerge) {  // computed before the state machine ran
            copy(dst, mval + (size_t)mcur * n);
            ++mcur;
            return;
        }
        // node buffers are not 4-byte aligned in the serialized state: the
        // lanes gather the 2n input bytes into an aligned buffer first
        __shared__ alignas(16) uint8_t hb[3 * F::kN];
        if (tid < 2 * n) hb[tid] = in[tid];
        __syncthreads();
        if (tid < 64) {  // thread 0's wave, every lane on the same hash (mirror_lanes)
            typename F::Node L, R, o;
            F::load(L, hb);
            F::load(R, hb + n);
            keep_in_vgprs(L);
            if (mcur < nmerge) {
                F::fin_h(o, merges[mcur], L, R);
            } else {
                uint32_t ad[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) ad[i] = a[i];
                F::thash_h(o, key, ad, L, R);
            }
            if (tid == 0) F::store(o, hb + 2 * n);
        }
        ++mcur;
        __syncthreads();
        if (tid < n) dst[tid] = hb[2 * n + tid];
        __syncthreads();
    }
    __device__ static void load_bytes(typename F::Node& x, const uint8_t* p) {
        alignas(8) uint8_t b[F::kN];
        for (uint32_t i = 0; i < F::kN; ++i) b[i] = p[i];
        F::load(x, b);
    }
    __device__ static void store_bytes(const typename F::Node& x, uint8_t* p) {
        alignas(8) uint8_t b[F::kN];
        F::store(x, b);
        for (uint32_t i = 0; i < F::kN; ++i) p[i] = b[i];
    }
    __device__ void leaf(uint8_t* dst) {
        copy(dst, leaves + (size_t)cur * n);
        ++cur;
    }

    // treehash_minheight_on_stack (xmss_core_fast.c:195-207)
    __device__ uint32_t minheight(uint32_t s, uint32_t j) const {
        uint32_t r = P.th;
        const uint32_t so = stackoffset(s), use = th_usage(s, j);
        for (uint32_t i = 0; i < use; ++i) {
            const uint32_t l = level(s, so - i - 1);
            if (l < r) r = l;
        }
        return r;
    }

    // treehash_update (xmss_core_fast.c:406-452) on instance j of state s;
    // nb is a 2n-byte scratch (LDS).  The instance's stack usage and the
    // state's stack offset live in registers while it runs (every thread
    // keeps the same copies), and thread 0 writes them through to the state:
    // a field that thread 0 rewrites is never read back by the other waves in
    // the same step, which a planned merge (no barrier before the writes)
    // would otherwise race with.
    __device__ void treehash_update(uint32_t s, uint32_t j, const uint32_t (&addr)[8], uint8_t* nb) {
        uint32_t nd[8] = {addr[0], addr[1], addr[2], 2, 0, 0, 0, 0};
        const uint32_t next = th_next(s, j);
        uint32_t use = th_usage(s, j), so = stackoffset(s);
        uint32_t nodeheight = 0;
        leaf(nb);
        // (the stack levels below the offset are not written in the loop)
        while (use > 0 && level(s, so - 1) == nodeheight) {
            if (mval && mcur < nmerge) {  // a planned merge: its value is known, its inputs unused
                if (tid < n) nb[tid] = mval[(size_t)mcur * n + tid];
                ++mcur;
            } else {  // (stack top, node) into the 2n bytes at nb, each lane its own byte
                if (tid < n) {
                    nb[n + tid] = nb[tid];
                    nb[tid] = sk[o_stack(s) + (size_t)(so - 1) * n + tid];
                }
                __syncthreads();
                nd[5] = nodeheight;
                nd[6] = next >> (nodeheight + 1);
                hash_h(nb, nb, nd);
            }
            ++nodeheight;
            --use;
            --so;
            wr_nosync(o_th(s, j) + 5, 1, use);
            wr_nosync(o_stackoffset(s), 4, so);
            __syncthreads();
        }
        if (nodeheight == th_h(s, j)) {
            copy(sk + th_node(s, j), nb);
            wr(o_th(s, j) + 6, 1, 1);
        } else {
            copy(sk + o_stack(s) + (size_t)so * n, nb);
            wr_nosync(o_th(s, j) + 5, 1, use + 1);
            wr_nosync(o_stacklevels(s) + so, 1, nodeheight);
            wr_nosync(o_stackoffset(s), 4, so + 1);
            wr_nosync(o_th(s, j) + 1, 4, next + 1);
            __syncthreads();
        }
    }

    // bds_treehash_update (xmss_core_fast.c:458-491), `char` return kept
    __device__ uint32_t bds_treehash_update(uint32_t s, uint32_t updates, const uint32_t (&addr)[8], uint8_t* nb) {
        const uint32_t nth = P.th - P.bds_k;  // at most 20 instances
        uint32_t used = 0;
        for (uint32_t u = 0; u < updates; ++u) {
            // The instance with the lowest node height on the stack (the
            // reference scans them in order, keeping the first strictly
            // lowest below th): lane i of every wave takes instance i, and a
            // min over (height, i) across the wave picks the same one.
            const uint32_t i = tid & 63;
            uint32_t key = 0xFFFFFFFFu;
            if (i < nth) {
                uint32_t low;
                if (th_done(s, i)) low = P.th;
                else if (th_usage(s, i) == 0) low = i;
                else low = minheight(s, i);
                key = (low << 8) | i;
            }
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) key = min(key, (uint32_t)__shfl_xor((int)key, o));
            const uint32_t lev = (key >> 8) < P.th ? (key & 0xFF) : nth;
            if (lev == nth) break;
            treehash_update(s, lev, addr, nb);
            ++used;
        }
        return (uint32_t)(int)(signed char)(updates - used);
    }

    // where a merged right child of height h and index j is recorded (:262-269, :534-541)
    __device__ uint8_t* record_slot(uint32_t s, uint32_t h, uint32_t j) {
        if (j == 1) return sk + o_auth(s) + (size_t)h * n;
        if (h < P.th - P.bds_k && j == 3) return sk + th_node(s, h);
        if (h >= P.th - P.bds_k)
            return sk + o_retain(s) + (size_t)((1u << (P.th - 1 - h)) + h - P.th + ((j - 3) >> 1)) * n;
        return nullptr;
    }

    // bds_state_update (xmss_core_fast.c:497-552)
    __device__ void bds_state_update(uint32_t s, const uint32_t (&addr)[8]) {
        const uint32_t idx = next_leaf(s);
        if (idx == (1u << P.th)) return;
        uint32_t nd[8] = {addr[0], addr[1], addr[2], 2, 0, 0, 0, 0};
        uint32_t so = stackoffset(s);
        leaf(sk + o_stack(s) + (size_t)so * n);
        set_level(s, so, 0);
        set_stackoffset(s, ++so);
        while (so > 1 && level(s, so - 1) == level(s, so - 2)) {
            const uint32_t nodeh = level(s, so - 1);
            uint8_t* slot = record_slot(s, nodeh, idx >> nodeh);
            if (slot) copy(slot, sk + o_stack(s) + (size_t)(so - 1) * n);
            nd[5] = nodeh;
            nd[6] = idx >> (nodeh + 1);
            hash_h(sk + o_stack(s) + (size_t)(so - 2) * n, sk + o_stack(s) + (size_t)(so - 2) * n, nd);
            set_level(s, so - 2, level(s, so - 2) + 1);
            set_stackoffset(s, --so);
        }
        wr(o_next_leaf(s), 4, idx + 1);
    }

    // bds_round (xmss_core_fast.c:559-626); buf is a 2n-byte scratch
    __device__ void bds_round(uint32_t s, uint32_t leaf_idx, const uint32_t (&addr)[8], uint8_t* buf) {
        const uint32_t th = P.th, nth = th - P.bds_k;
        uint32_t tau = th;
        for (uint32_t i = 0; i < th; ++i)
            if (!((leaf_idx >> i) & 1)) {
                tau = i;
                break;
            }
        if (tau > 0) {
            copy(buf, sk + o_auth(s) + (size_t)(tau - 1) * n);
            copy(buf + n, sk + o_keep(s) + (size_t)((tau - 1) >> 1) * n);
        }
        if (!((leaf_idx >> (tau + 1)) & 1) && (tau < th - 1))
            copy(sk + o_keep(s) + (size_t)(tau >> 1) * n, sk + o_auth(s) + (size_t)tau * n);
        if (tau == 0) {
            leaf(sk + o_auth(s));
        } else {
            uint32_t nd[8] = {addr[0], addr[1], addr[2], 2, 0, tau - 1, leaf_idx >> tau, 0};
            hash_h(sk + o_auth(s) + (size_t)tau * n, buf, nd);
            for (uint32_t i = 0; i < tau; ++i) {
                if (i < nth) {
                    copy(sk + o_auth(s) + (size_t)i * n, sk + th_node(s, i));
                } else {
                    const uint32_t offset = (1u << (th - 1 - i)) + i - th;
                    const uint32_t rowidx = ((leaf_idx >> i) - 1) >> 1;
                    copy(sk + o_auth(s) + (size_t)i * n, sk + o_retain(s) + (size_t)(offset + rowidx) * n);
                }
            }
            const uint32_t lim = tau < nth ? tau : nth;
            for (uint32_t i = 0; i < lim; ++i) {
                const uint32_t startidx = leaf_idx + 1 + 3 * (1u << i);
                if (startidx < (1u << th)) {
                    wr(o_th(s, i), 1, i);
                    wr(o_th(s, i) + 1, 4, startidx);
                    wr(o_th(s, i) + 6, 1, 0);
                    wr(o_th(s, i) + 5, 1, 0);
                }
            }
        }
    }

    // deep_state_swap (xmss_core_fast.c:165-193): states occupy equal, disjoint byte ranges
    __device__ void swap_states(uint32_t a, uint32_t b) { swap(st(a), st(b), P.state_bytes); }

    // wots_sign of the n bytes at msg (in sk) into sig (in sk), all chain lanes
    __device__ void wots_sign(uint8_t* sig, const uint8_t* msg, uint32_t layer, uint64_t tree, uint32_t ots,
                              uint8_t* scratch) {
        // copy the message to an aligned scratch first (it may move during the swap)
        if (tid < n) scratch[tid] = msg[tid];
        __syncthreads();
        if (wcur < nwots) {
            // planned: chain starts and every step's address-only work come
            // from the pre-image launch (LAT_SIGN, all steps), so each step
            // is one compression
            const uint8_t* S = wpre + (size_t)wcur * wstride;
            ++wcur;
            // chain c on lane c; the lanes past the last chain in its wave
            // walk a copy (mirror_lanes)
            const uint32_t c = mirror_lanes(tid, F::LEN);
            if (c < (uint32_t)F::LEN) {
                typename F::Node x = reinterpret_cast<const typename F::Node*>(S + LatLayout<F>::starts)[c];
                const typename F::FPre* fp = reinterpret_cast<const typename F::FPre*>(S + LatLayout<F>::fpre) + c * 15;
                const uint32_t steps = chain_len<F>(scratch, c);
                for (uint32_t s = 0; s < steps; ++s) F::fin_f(x, fp[s], x);
                if (c == tid) store_bytes(x, sig + (size_t)c * n);
            }
            __syncthreads();
            return;
        }
        const uint32_t c = mirror_lanes(tid, F::LEN);
        if (c < (uint32_t)F::LEN) {
            uint32_t a[8] = {layer, (uint32_t)(tree >> 32), (uint32_t)tree, 0, ots, c, 0, 0};
            typename F::Node x;
            F::prf_keygen(x, key, a);
            run_chain<F>(x, key, a, 0, chain_len<F>(scratch, c));
            if (c == tid) store_bytes(x, sig + (size_t)c * n);
        }
        __syncthreads();
    }
};

// The hashing of BDS signatures, in three launches on one stream (so that a
// signature costs no cross-stream waits: the host's launch calls, as much as
// the GPU, bound a single signature's latency):
//   k_bds_pre:    the address-only work of the planned leaves (LAT_LEAF
//                 rows), of the layer-0 WOTS signatures (LAT_SIGN rows, all 15
//                 steps of every chain) and of the planned tree merges (one
//                 merge per lane);
//   k_bds_fin:    one workgroup per planned leaf;
//   k_bds_update: the state update of each signature.
// A signature launched alone usually needs only the first two: the last
// workgroup of k_bds_fin runs its state update (BdsFinDag).
// R / H_msg (one message per lane) and the layer-0 WOTS signatures (one
// workgroup each) go one of two ways.  Short messages: H_msg in the first
// rows of k_bds_pre and the signatures beside the leaves in k_bds_fin.  Long
// messages (H_msg a long serial chain of compressions): H_msg beside the
// leaves in k_bds_fin and the signatures beside the state updates in
// k_bds_update, so that H_msg overlaps the leaves.
// The leaves and signatures share one scratch area (rows nleaf .. for the
// signatures).
template <class F>
__global__ __launch_bounds__(WG) void k_bds_pre(const typename F::Key* __restrict__ keys,
                                                 const DMsg* __restrict__ msgs, uint32_t pad,
                                                 uint8_t* __restrict__ blob, uint8_t* __restrict__ R_out,
                                                 uint8_t* __restrict__ digests, uint32_t hrows, uint32_t nmsg,
                                                 const DLeaf* __restrict__ leaves, uint32_t nleaf,
                                                 const DLeaf* __restrict__ sigs, uint32_t nsign,
                                                 uint8_t* __restrict__ scratch, size_t stride,
                                                 const DThash* __restrict__ merges, uint32_t nmerge,
                                                 typename F::HPre* __restrict__ merge_out, HmsgMid* __restrict__ mid,
                                                 uint32_t head) {
    const uint32_t t = blockIdx.x * WG + threadIdx.x;
    if (head && blockIdx.y == gridDim.y - 1) {  // the last row: the first blocks of long messages' H_msg
        if constexpr (F::kPairHmsg) {
            const uint32_t r = mirror_lanes(t, nmsg);
            if (r < nmsg) hmsg_head<F>(keys, msgs, r, pad, blob, R_out, digests, mid, head, r == t);
        }
        return;
    }
    if (blockIdx.y < hrows) {
        hmsg_mirrored<F>(keys, msgs, blockIdx.y * gridDim.x * WG + t, nmsg, pad, blob, R_out, digests);
        return;
    }
    const uint32_t y = blockIdx.y - hrows;
    if (y < nleaf) {
        lat_pre_one<F, LAT_LEAF>(keys, leaves, y, t, nullptr, scratch, stride);
    } else if (y < nleaf + nsign) {
        lat_pre_one<F, LAT_SIGN>(keys, sigs, y - nleaf, t, nullptr, scratch + (size_t)nleaf * stride, stride);
    } else {
        const uint32_t i = (y - nleaf - nsign) * gridDim.x * WG + t;
        if (i >= nmerge) return;
        const DThash q = merges[i];
        uint32_t a[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] = q.adrs[j];
        typename F::HPre p;
        F::pre_h(p, keys[q.key], a);
        merge_out[i] = p;
    }
}

// Whether k_bds_fin's H_msg blocks run two waves per 64 messages
// (hmsg_pair_block): SHA-256 families whose leaf workgroups have two waves.
template <class F>
struct BdsFinPair {
    static constexpr bool value = F::kPairHmsg && ((F::LEN + 63) / 64) * 64 == 128;
};
// LDS of k_bds_fin's merge-DAG block when it also runs the state update
// (BdsFinDag::state): the update's states, merge values and leaves.
constexpr uint32_t kFinStateLds = 32768;
template <class F, bool PAIR = BdsFinPair<F>::value>
struct BdsFinLds {
    union {
        typename F::Node pk[F::LEN];
        HmsgPairLds pair;
        alignas(16) uint8_t raw[kFinStateLds];
    };
};
template <class F>
struct BdsFinLds<F, false> {
    union {
        typename F::Node pk[F::LEN];
        alignas(16) uint8_t raw[kFinStateLds];
    };
};

// The planned merges of one signature's state update hashed level by level,
// one lane per merge, into mv (n bytes per merge): in the update's LDS, or
// beside H_msg in k_bds_fin (BdsFinDag).  q0: 3 words per merge (sources and
// level, Engine::BdsJob::plan_dag); pre: the merges' address-only work; every
// thread of the workgroup calls it.
template <class F>
__device__ __forceinline__ void dag_load(typename F::Node& x, uint32_t src, const uint8_t* lits, const uint8_t* leaves,
                         const uint8_t* mval) {
    const uint32_t i = src & DAG_INDEX;
    const uint8_t* p = (src & ~DAG_INDEX) == DAG_LIT ? lits : (src & ~DAG_INDEX) == DAG_LEAF ? leaves : mval;
    F::load(x, p + (size_t)i * F::kN);
}
template <class F>
__device__ void merge_dag(const uint32_t* q0, uint32_t nmerge, uint32_t maxlev, const uint8_t* lits,
                          const uint8_t* leaves, const typename F::HPre* pre, uint8_t* mv) {
    constexpr uint32_t n = F::kN;
    for (uint32_t lev = 1; lev <= maxlev; ++lev) {
        for (uint32_t m0 = 0; m0 < nmerge; m0 += blockDim.x) {
            const uint32_t m = m0 + threadIdx.x;
            const bool mine = m < nmerge && q0[3 * m + 2] == lev;
            const uint64_t act = __ballot(mine);
            if (act == 0) continue;
            // a lane with no merge of this level hashes a copy of the next
            // one in its wave, cyclically (mirror_lanes)
            const uint32_t lane = threadIdx.x & 63;
            const uint64_t rot = lane ? (act >> lane) | (act << (64 - lane)) : act;
            const uint32_t src = m - lane + ((lane + (uint32_t)__builtin_ctzll(rot)) & 63);
            const uint32_t* q = q0 + 3 * src;
            typename F::Node L, R, o;
            dag_load<F>(L, q[0], lits, leaves, mv);
            dag_load<F>(R, q[1], lits, leaves, mv);
            F::fin_h(o, pre[src], L, R);
            if (mine) F::store(o, mv + (size_t)m * n);
        }
        __threadfence_block();
        __syncthreads();
    }
}

// The two halves of a state update (bds_update_job): STAGE needs only the sk
// (the resident sk into the arena, the states into LDS, the signature tail
// from the old states), RUN the planned leaves (their staging, the merges,
// the state machine, the write-back).
enum : uint32_t { BDS_PHASE_STAGE = 1, BDS_PHASE_RUN = 2, BDS_PHASE_ALL = 3 };
template <class F>
__device__ __forceinline__ void bds_update_job(const typename F::Key& key, const DBds& job, uint8_t* skg,
                               const uint8_t* __restrict__ leaf_arena,
                               const typename F::HPre* __restrict__ merge_arena, const uint32_t* __restrict__ dag,
                               const uint8_t* __restrict__ lit_arena, const uint8_t* __restrict__ wpre, size_t wstride,
                               uint8_t* __restrict__ tails, uint8_t* __restrict__ scratch_arena,
                               const uint8_t* premerged, uint8_t* lds_states, uint32_t phases);

// The merge DAG of a one-signature launch hashed in k_bds_fin, beside H_msg
// and the WOTS signature, instead of in the update launch: its last block
// waits until the blocks of the job's `wait` planned leaves have each
// counted in at sync[0] (a bounded wait: after `spin` polls it gives up and
// the update hashes the DAG as usual), hashes the DAG into vals and sets
// sync[1] = 1.  on == 0: off.
struct BdsFinDag {
    uint32_t* sync;
    uint32_t on, wait, spin, nmerge, maxlev;
    const uint32_t* q0;
    const uint8_t* lits;
    const void* pre;
    uint8_t* vals;
    // state: the block instead runs job 0's whole state update
    // (bds_update_job) in its LDS, the merges included (vals unused), and
    // sets sync[1] = 2, and the update launch skips the job
    uint32_t state;
    // select (a long message's chain values walked beside H_msg): the
    // values block counts in at sync[2], and the H_msg block, once its
    // digest is out and the values are in (the same bounded wait), picks the
    // WOTS signature and sets sync[3] = 1; with state and select both done
    // nothing is left for the update launch
    uint32_t select;
    const DBds* jobs;
    uint8_t* sk_arena;
    uint8_t* tails;
    const uint8_t* wpre;
    size_t wstride;
};

// The waits of k_bds_fin's blocks on blocks dispatched before them (so that
// the wait always ends, bounded anyway): whether `want` blocks had counted in
// at *count within `spin` polls.  Every thread of the workgroup calls it; a
// true return makes the counted blocks' writes visible to all of them.
__device__ __forceinline__ bool fin_wait(const uint32_t* count, uint32_t want, uint32_t spin) {
    __shared__ uint32_t ready;
    if (threadIdx.x == 0) {
        uint32_t polls = 0;
        while (__hip_atomic_load(count, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < want && polls < spin) {
            __builtin_amdgcn_s_sleep(2);
            ++polls;
        }
        ready = __hip_atomic_load(count, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) >= want;
    }
    __syncthreads();
    return ready;
}
// ... and the counting in, once the whole workgroup's writes are done (when on)
__device__ __forceinline__ void fin_count_in(bool on, uint32_t* count) {
    if (!on) return;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        atomicAdd(count, 1u);
    }
}

// mid (long messages, SHA-256 families): k_bds_pre hashed the first blocks of
// each H_msg (hmsg_head) and the H_msg blocks resume from there.
// Blocks: nleaf leaves (or, for rows marked LAT_ENDS_OUT / LAT_ENDS_IN, the
// chains into slot p2 of `ends` / the L-tree from it), then nsig WOTS
// signatures (short messages), then the H_msg blocks (long messages): 64
// messages each by two waves (BdsFinPair), else one message per lane.
template <class F>
__global__ __launch_bounds__(192) void k_bds_fin(const typename F::Key* __restrict__ keys,
                                                  const DMsg* __restrict__ msgs, uint32_t nmsg, uint32_t pad,
                                                  uint8_t* __restrict__ blob, uint8_t* __restrict__ R_out,
                                                  uint8_t* __restrict__ digests, const DLeaf* __restrict__ leaves,
                                                  uint32_t nleaf, const DLeaf* __restrict__ sigs, uint32_t nsig,
                                                  const uint8_t* scratch,  // not restrict: the chain
                                                  size_t stride,           // loops' loads stay ahead
                                                  uint8_t* __restrict__ leaf_out, uint8_t* __restrict__ sig_out,
                                                  uint8_t* __restrict__ ends, uint32_t nvals,
                                                  uint8_t* __restrict__ vals, BdsFinDag dag,
                                                  const HmsgMid* __restrict__ mid) {
    __shared__ BdsFinLds<F> sh;
    if (dag.on && blockIdx.x == gridDim.x - 1) {  // the merge DAG, once the planned leaves are in
        const typename F::HPre* pre = reinterpret_cast<const typename F::HPre*>(dag.pre);
        const DBds job = dag.jobs[0];
        // (state) what needs no leaf goes first, while the leaves are hashed;
        // should the wait give up, the update launch redoes it, to the same
        // bytes
        if (dag.state)
            bds_update_job<F>(keys[job.key], job, dag.sk_arena + job.sk_off, leaf_out, pre, dag.q0, dag.lits,
                              dag.wpre, dag.wstride, dag.tails, nullptr, nullptr, sh.raw, BDS_PHASE_STAGE);
        if (!fin_wait(dag.sync, dag.wait, dag.spin)) return;  // the update launch does it all
        if (dag.state) {  // the merges into LDS and the state update, while H_msg or the WOTS signature runs
            bds_update_job<F>(keys[job.key], job, dag.sk_arena + job.sk_off, leaf_out, pre, dag.q0, dag.lits,
                              dag.wpre, dag.wstride, dag.tails, nullptr, nullptr, sh.raw, BDS_PHASE_RUN);
            __syncthreads();
            if (threadIdx.x == 0) {
                __threadfence();
                dag.sync[1] = 2;
            }
        } else if (dag.maxlev) {  // the merges alone, for the update launch
            merge_dag<F>(dag.q0, dag.nmerge, dag.maxlev, dag.lits, leaf_out, pre, dag.vals);
            if (threadIdx.x == 0) {
                __threadfence();
                dag.sync[1] = 1;
            }
        }
        return;
    }
    if (blockIdx.x < nleaf) {
        const DLeaf q = leaves[blockIdx.x];
        uint8_t* e = ends + (size_t)q.p2 * F::LEN * F::kN;
        if (q.p1 & LAT_ENDS_OUT) lat_leaf_chains<F>(blockIdx.x, scratch, stride, e);
        else if (q.p1 & LAT_ENDS_IN) lat_leaf_ltree<F>(blockIdx.x, sh.pk, scratch, stride, e, leaf_out);
        else lat_fin_one<F, LAT_LEAF>(blockIdx.x, sh.pk, leaves, nullptr, nullptr, nullptr, scratch, stride, leaf_out);
        fin_count_in(dag.on && blockIdx.x < dag.wait, dag.sync);  // a planned leaf of the DAG block's job
    } else if (blockIdx.x < nleaf + nsig) {
        lat_fin_one<F, LAT_SIGN>(blockIdx.x - nleaf, sh.pk, sigs, digests, nullptr, nullptr,
                                 scratch + (size_t)nleaf * stride, stride, sig_out);
    } else if (blockIdx.x < nleaf + nsig + nvals) {
        // long messages, one signature: every value along its WOTS chains,
        // beside H_msg (the update launch then picks the signature)
        const uint32_t j = blockIdx.x - nleaf - nsig;
        lat_sig_values<F>(j, scratch + (size_t)nleaf * stride, stride, vals + (size_t)j * F::LEN * 16 * F::kN);
        fin_count_in(dag.select, dag.sync + 2);
    } else {
        if constexpr (BdsFinPair<F>::value)
            hmsg_pair_block<F>(keys, msgs, (blockIdx.x - nleaf - nsig - nvals) * 64, nmsg, pad, blob, R_out,
                               digests, sh.pair, mid);
        else
            hmsg_mirrored<F>(keys, msgs, (blockIdx.x - nleaf - nsig - nvals) * blockDim.x + threadIdx.x, nmsg,
                             pad, blob, R_out, digests);
        // one message (select): its WOTS signature from the values, which
        // blocks dispatched before this one walk
        if (dag.select) {
            __syncthreads();  // the digest is out
            if (!fin_wait(dag.sync + 2, nvals, dag.spin)) return;  // the update launch picks it
            lat_sig_select<F>(digests, vals, sig_out);
            __syncthreads();
            if (threadIdx.x == 0) {
                __threadfence();
                dag.sync[3] = 1;
            }
        }
    }
}

// The state part of one BDS signature per workgroup (everything after R, H_msg
// and the layer-0 WOTS signature, which run in their own launches): the auth
// paths and lower-layer WOTS signatures go to the signature tail, then the
// traversal updates the state at skg.  Its node scratch is 4n bytes of LDS
// (scratch_arena is no longer read).

template <class F>
__device__ __forceinline__ void bds_update_one(const typename F::Key& key, const DBds& job, uint8_t* skg,
                               const uint8_t* __restrict__ leaf_arena,
                               const typename F::HPre* __restrict__ merge_arena, const uint32_t* __restrict__ dag,
                               const uint8_t* __restrict__ lit_arena, const uint8_t* __restrict__ wpre, size_t wstride,
                               uint8_t* __restrict__ tails, uint8_t* __restrict__ scratch_arena,
                               const uint8_t* premerged, uint8_t* lds_states, uint32_t phases) {
    const BdsP& P = job.P;
    BdsDev<F> B(lds_states, skg, P, leaf_arena + job.leaf_off * F::kN, merge_arena + job.merge_off, job.nmerge,
                key);
    B.wpre = wpre + (size_t)job.wots_row * wstride;
    B.wstride = wstride;
    B.nwots = job.nwots;
    // the state machine's node scratch (4n bytes) in LDS: its copies are
    // one byte per lane behind a barrier, and each through HBM cost about a
    // microsecond (a treehash step is several of them)
    __shared__ alignas(16) uint8_t scr[4 * F::kN];
    uint8_t* scratch = scr;
    (void)scratch_arena;
    uint8_t* tail = tails + job.tail_off;
    const uint32_t n = F::kN, th = P.th, d = P.d;
    const uint64_t idx = job.idx;
    const uint32_t region = (2 * d - 1) * P.state_bytes;
    if (phases & BDS_PHASE_STAGE) {
        // stage the serialized states in LDS (written back when the update is
        // done), then the signature tail: auth of layer 0, then (WOTS
        // signature, auth) of layers 1..d-1
        wg_copy(lds_states, skg + B.states_base(), region);
        __syncthreads();
        size_t o = 0;
        for (uint32_t i = 0; i < d; ++i) {
            if (i > 0) {
                const uint8_t* w = B.wots_sigs() + (size_t)(i - 1) * P.wots_sig_bytes;
                wg_copy(tail + o, w, P.wots_sig_bytes);
                o += P.wots_sig_bytes;
            }
            const uint8_t* a = B.sk + B.o_auth(i);
            wg_copy(tail + o, a, th * n);
            o += (size_t)th * n;
        }
        __syncthreads();
    }
    if (!(phases & BDS_PHASE_RUN)) return;
    // the planned leaves: the state machine takes them one at a time,
    // each a read of HBM that another launch wrote (about 4 us apiece
    // measured) when read in place
    if (job.lv_off) {
        uint8_t* lv = lds_states + job.lv_off;
        wg_copy(lv, B.leaves, job.nleaf * n);
        B.leaves = lv;
    }
    __syncthreads();  // the merges below read the staged leaves of every lane
    // The planned merges whose sources the host's dry run resolved (node
    // values of the sk, planned leaves, earlier merges) are hashed here level
    // by level, one lane per merge, into LDS after the states; the state
    // machine below then copies merge k where it would hash it.  Only merges
    // of one chain depend on each other, so this takes a few merge latencies
    // instead of one per merge.
    if (job.maxlev) {
        uint8_t* mv = lds_states + ((region + 15) & ~15u);
        if (premerged) {  // hashed beside H_msg (BdsFinDag)
            wg_copy(mv, premerged, job.nmerge * n);
        } else {
            merge_dag<F>(dag + 3 * job.merge_off, job.nmerge, job.maxlev, lit_arena + job.lit_off * n, B.leaves,
                         merge_arena + job.merge_off, mv);
        }
        B.mval = mv;
    }
    __syncthreads();
    struct WriteBack {
        uint8_t* l;
        uint8_t* g;
        uint32_t len;
        __device__ ~WriteBack() {
            __syncthreads();
            wg_copy(g, l, len);
        }
    } wb{lds_states, skg + B.states_base(), region};
    if (d == 1) {
        // xmss_core_sign (xmss_core_fast.c:836-839); the host routes the last index elsewhere
        if (idx < (1ull << th) - 1) {
            const uint32_t addr[8] = {0, 0, 0, 0, (uint32_t)idx, 0, 0, 0};
            B.bds_round(0, (uint32_t)idx, addr, scratch);
            B.bds_treehash_update(0, (th - P.bds_k) >> 1, addr, scratch);
        }
        return;
    }
    // xmssmt_core_sign (xmss_core_fast.c:1068-1113)
    uint64_t idx_tree = idx >> th;
    uint32_t idx_leaf = (uint32_t)(idx & ((1u << th) - 1));
    uint32_t addr[8] = {0, (uint32_t)((idx_tree + 1) >> 32), (uint32_t)(idx_tree + 1), 0, 0, 0, 0, 0};
    int needswap_upto = -1;
    uint32_t updates = (th - P.bds_k) >> 1;
    if ((1 + idx_tree) * (1ull << th) + idx_leaf < (1ull << P.full_height)) B.bds_state_update(d, addr);
    for (uint32_t i = 0; i < d; ++i) {
        if (!(((idx + 1) & ((1ull << ((i + 1) * th)) - 1)) == 0)) {
            idx_leaf = (uint32_t)((idx >> (th * i)) & ((1u << th) - 1));
            idx_tree = idx >> (th * (i + 1));
            addr[0] = i;
            addr[1] = (uint32_t)(idx_tree >> 32);
            addr[2] = (uint32_t)idx_tree;
            if ((int)i == needswap_upto + 1) B.bds_round(i, idx_leaf, addr, scratch);
            updates = B.bds_treehash_update(i, updates, addr, scratch);
            addr[1] = (uint32_t)((idx_tree + 1) >> 32);
            addr[2] = (uint32_t)(idx_tree + 1);
            if ((1 + idx_tree) * (1ull << th) + idx_leaf < (1ull << (P.full_height - th * i))) {
                if (i > 0 && updates > 0 && B.next_leaf(d + i) < (1ull << P.full_height)) {
                    B.bds_state_update(d + i, addr);
                    --updates;
                }
            }
        } else if (idx < (1ull << P.full_height) - 1) {
            B.swap_states(d + i, i);
            const uint64_t t = (idx + 1) >> ((i + 2) * th);
            const uint32_t ots = (uint32_t)(((idx >> ((i + 1) * th)) + 1) & ((1u << th) - 1));
            B.wots_sign(B.wots_sigs() + (size_t)i * P.wots_sig_bytes, B.sk + B.o_stack(i), i + 1, t, ots,
                        scratch + 2 * n);
            B.set_stackoffset(d + i, 0);
            B.wr(B.o_next_leaf(d + i), 4, 0);
            --updates;  // WOTS signing counts as one update (:1107)
            needswap_upto = (int)i;
            for (uint32_t j = 0; j < th - P.bds_k; ++j) B.wr(B.o_th(i, j) + 6, 1, 1);
        }
    }
}

// One signature's state update (bds_update_one) with its sk at skg, in the
// sk arena; for a key whose state is resident in HBM (job.res), the arena
// copy is filled from it first, with the index advanced as the signature
// requires (idx + 1, big-endian), and the updated bytes go back to it at the
// end, so the arena holds the new sk for the host's copy either way.  Every
// thread of the workgroup calls it.
template <class F>
__device__ __forceinline__ void bds_update_job(const typename F::Key& key, const DBds& job, uint8_t* skg,
                               const uint8_t* __restrict__ leaf_arena,
                               const typename F::HPre* __restrict__ merge_arena, const uint32_t* __restrict__ dag,
                               const uint8_t* __restrict__ lit_arena, const uint8_t* __restrict__ wpre, size_t wstride,
                               uint8_t* __restrict__ tails, uint8_t* __restrict__ scratch_arena,
                               const uint8_t* premerged, uint8_t* lds_states, uint32_t phases) {
    uint8_t* res = reinterpret_cast<uint8_t*>(job.res);
    if (res && (phases & BDS_PHASE_STAGE)) {
        wg_copy(skg, res, job.sk_bytes);
        __syncthreads();
        if (threadIdx.x < job.P.index_bytes) {
            const uint32_t sh = 8 * (job.P.index_bytes - 1 - threadIdx.x);
            skg[threadIdx.x] = sh < 64 ? (uint8_t)((job.idx + 1) >> sh) : 0;
        }
        __syncthreads();
    }
    bds_update_one<F>(key, job, skg, leaf_arena, merge_arena, dag, lit_arena, wpre, wstride, tails, scratch_arena,
                      premerged, lds_states, phases);
    if (res && (phases & BDS_PHASE_RUN)) {
        __syncthreads();
        wg_copy(res, skg, job.sk_bytes);
    }
}

// One signature's state update per workgroup (bds_update_job).
template <class F>
__global__ __launch_bounds__(192) void k_bds_update(const typename F::Key* __restrict__ keys, const DBds* __restrict__ jobs,
                             uint8_t* __restrict__ sk_arena, const uint8_t* __restrict__ leaf_arena,
                             const typename F::HPre* __restrict__ merge_arena, const uint32_t* __restrict__ dag,
                             const uint8_t* __restrict__ lit_arena, const uint8_t* __restrict__ wpre, size_t wstride,
                             uint8_t* __restrict__ tails, uint8_t* __restrict__ scratch_arena, uint32_t njobs,
                             const DLeaf* __restrict__ sigs, const uint8_t* __restrict__ digests,
                             const uint8_t* __restrict__ sign_scratch, uint8_t* __restrict__ sig_out,
                             const uint8_t* __restrict__ vals, const uint32_t* dag_sync, const uint8_t* dag_vals) {
    if (blockIdx.x >= njobs) {  // long messages: the layer-0 WOTS signature of job blockIdx.x - njobs
        const uint32_t j = blockIdx.x - njobs;
        if (vals)  // its chain values were walked beside H_msg
            lat_sig_select<F>(digests + (size_t)j * F::kN, vals + (size_t)j * F::LEN * 16 * F::kN,
                              sig_out + (size_t)j * F::LEN * F::kN);
        else
            lat_fin_one<F, LAT_SIGN>(j, nullptr, sigs, digests, nullptr, nullptr, sign_scratch, wstride, sig_out);
        return;
    }
    if (blockIdx.x == 0 && dag_sync && dag_sync[1] == 2) return;  // done in k_bds_fin (BdsFinDag::state)
    extern __shared__ __align__(16) uint8_t lds_dyn[];
    const DBds job = jobs[blockIdx.x];
    // job 0's merges, if k_bds_fin hashed them (dag_sync[1] set)
    const uint8_t* premerged = dag_vals && blockIdx.x == 0 && dag_sync[1] == 1 ? dag_vals : nullptr;
    bds_update_job<F>(keys[job.key], job, sk_arena + job.sk_off, leaf_arena, merge_arena, dag, lit_arena, wpre,
                      wstride, tails, scratch_arena, premerged, lds_dyn, BDS_PHASE_ALL);
}

}  // namespace xm
