// launch.h -- kernel launch helper with optional per-kernel HIP-event timing,
// and the per-family launch table (one implementation per fam_*.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "engine.h"
#include "kernels.h"
#include "kernels_bds.h"
#include "kernels_lat.h"

namespace xm {

[[noreturn]] void fatal(const char* what);
[[noreturn]] void hip_failed(hipError_t e, const char* file, int line, const char* what);

#define HIPCK(x)                                                  \
    do {                                                          \
        hipError_t e_ = (x);                                      \
        if (e_ != hipSuccess) hip_failed(e_, __FILE__, __LINE__, #x); \
    } while (0)

// Device memory of the library's workspaces, key pools and resident serving
// state goes through these: they count the bytes held and refuse an
// allocation beyond the limit set by xmss_mi355x_set_device_limit (0: none).
void* dev_malloc(size_t bytes);
void dev_free(void* p, size_t bytes);

struct Launcher {
    hipStream_t s = nullptr;
    hipStream_t s_main = nullptr, s_aux = nullptr;  // s is one of the two
    bool prof = false;
    // A sizing pass: launches are skipped (the caller skips its copies too),
    // so that a launch sequence sizes every workspace before it is captured.
    bool dry = false;
    struct Pending {
        const char* name;
        hipEvent_t e0, e1;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    std::map<std::string, std::pair<uint64_t, double>> stat;

    hipEvent_t ev() {
        if (pool.empty()) {
            hipEvent_t e;
            HIPCK(hipEventCreate(&e));
            return e;
        }
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    // Events bracket each launch on the same stream, so elapsed time between
    // them is that kernel's duration.
    template <typename K, typename... Args>
    void run(const char* name, K kernel, dim3 grid, dim3 block, Args... args) {
        run_lds(name, kernel, grid, block, 0, args...);
    }
    template <typename K, typename... Args>
    void run_lds(const char* name, K kernel, dim3 grid, dim3 block, size_t lds, Args... args) {
        if (grid.x == 0 || grid.y == 0 || dry) return;
        hipEvent_t e0 = nullptr;
        if (prof) {
            e0 = ev();
            HIPCK(hipEventRecord(e0, s));
        }
        hipLaunchKernelGGL(kernel, grid, block, lds, s, args...);
        HIPCK(hipGetLastError());
        if (prof) {
            hipEvent_t e1 = ev();
            HIPCK(hipEventRecord(e1, s));
            pending.push_back({name, e0, e1});
        }
    }
    void harvest() {
        for (auto& p : pending) {
            float ms = 0;
            HIPCK(hipEventElapsedTime(&ms, p.e0, p.e1));
            auto& st = stat[p.name];
            st.first += 1;
            st.second += ms;
            pool.push_back(p.e0);
            pool.push_back(p.e1);
        }
        pending.clear();
    }
};

static inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Kernel launches of one hash family; all pointers are device pointers.
struct FamOps {
    virtual ~FamOps() {}
    virtual size_t key_bytes() const = 0;
    virtual void prepare(Launcher& L, const uint8_t* seeds, int nkeys, void* keys) = 0;
    virtual void chains(Launcher& L, const void* keys, const DRun* runs, int nruns, uint32_t leaf_off,
                        uint32_t leaves, uint8_t* ends) = 0;
    virtual void ltree_level(Launcher& L, const void* keys, const DRun* runs, int nruns, uint32_t leaf_off,
                             uint32_t leaves, uint32_t l, uint32_t height, const uint8_t* src,
                             uint32_t src_stride, uint8_t* dst, uint32_t dst_stride, uint8_t* out,
                             uint64_t out_stride, uint64_t out_base) = 0;
    virtual void level(Launcher& L, const void* keys, const DRun* runs, int nruns, uint32_t th,
                       uint32_t level, uint8_t* nodes) = 0;
    virtual void wots(Launcher& L, const void* keys, const DWots* reqs, uint32_t n, const uint8_t* msgs,
                      const uint8_t* in, uint8_t* out) = 0;
    virtual void leaf_wg(Launcher& L, const void* keys, const DLeaf* reqs, uint32_t n, bool verify,
                         const uint8_t* msgs, const uint8_t* sigs, const uint8_t* auth, uint32_t th,
                         uint8_t* out) = 0;
    // Latency-shaped leaves / WOTS signatures / verify layers (kernels_lat.h):
    // scratch must hold n x lat_bytes(th) with th the largest tree height of
    // the requests (0 except for verification); n <= 65535.  lat = lat_pre +
    // lat_fin on the same requests and scratch.
    virtual size_t lat_bytes(uint32_t th) const = 0;
    virtual void lat(Launcher& L, const void* keys, uint32_t mode, const DLeaf* reqs, uint32_t n,
                     const uint8_t* msgs, const uint8_t* sigs, const uint8_t* auth, uint32_t th,
                     uint8_t* scratch, uint8_t* out) = 0;
    virtual void lat_pre(Launcher& L, const void* keys, uint32_t mode, const DLeaf* reqs, uint32_t n,
                         const uint8_t* msgs, uint32_t th, uint8_t* scratch) = 0;
    // H_msg of nmsg messages (digests) with lat_pre(LAT_VERIFY) of n requests
    // (all of them LAT_ALL_STEPS) in one launch; n + ceil(nmsg / 64) <= 65535.
    virtual void verify_pre(Launcher& L, const void* keys, const DMsg* msgs, uint32_t nmsg, uint32_t pad,
                            uint8_t* blob, uint8_t* digests, const DLeaf* reqs, uint32_t n, uint32_t th,
                            uint8_t* scratch) = 0;
    // Every verify layer of n requests (request j: items base[L] + j for
    // L < its depth, requests sorted deepest first; dmax <= 16), one
    // workgroup per request; DLeaf p2 = auth offset from auth.
    // Layers L0 .. L1-1 of n requests (items layer-major from base[]); with
    // nspec > 0 it also hashes the speculative chain ends of the nspec items
    // of layers >= 1 into spec, and with nspec = 0 and spec set, those items
    // take their chain ends from spec.
    virtual void verify_layers(Launcher& L, const DLeaf* reqs, uint32_t n, const uint32_t* base, uint32_t dmax,
                               uint32_t L0, uint32_t L1, const uint8_t* digests, const uint8_t* sigs,
                               const uint8_t* auth, uint32_t th, const uint8_t* scratch, uint8_t* roots,
                               uint8_t* spec, uint32_t nspec) = 0;
    // LAT_LEAF for many requests: lat_pre, then one lane per chain and one
    // workgroup per L-tree (ends: n x LEN nodes of scratch) -> out[r] = leaf.
    virtual void lat_leaves_dense(Launcher& L, const void* keys, const DLeaf* reqs, uint32_t n, uint8_t* scratch,
                                  uint8_t* ends, uint8_t* out) = 0;
    virtual void lat_fin(Launcher& L, uint32_t mode, const DLeaf* reqs, uint32_t n, const uint8_t* msgs,
                         const uint8_t* sigs, const uint8_t* auth, uint32_t th, const uint8_t* scratch,
                         uint8_t* out) = 0;
    // Lane-dense finish of n verify requests (same layer) whose k_lat_pre
    // scratch is ready: ends and tmp hold n x LEN and n x ceil(LEN/2) nodes.
    // For LAT_VERIFY requests of tree height th: lat_pre_h computes only the
    // H pre-images (address only), pre_need the F pre-images of only the
    // chain steps their signatures need, from their messages.
    virtual void lat_pre_h(Launcher& L, const void* keys, const DLeaf* reqs, uint32_t n, uint32_t th,
                           uint8_t* scratch, uint32_t prio = 0) = 0;
    virtual void pre_need(Launcher& L, const void* keys, const DLeaf* reqs, uint32_t n, const uint8_t* msgs,
                          uint32_t th, uint8_t* scratch, uint32_t prio = 0) = 0;
    // Large verification batches, lane-dense: the chains (ends[r][c]) and
    // then the L-tree and auth walk (roots in out).  tail: 0 = one launch per
    // L-tree height, then the auth walks; 1 = the whole L-tree through LDS in
    // one launch, then the auth walks; 2 = L-tree and auth walk in one
    // launch.  prio: the waves' issue priority (wave_priority).
    virtual void latb_chains(Launcher& L, uint32_t n, const uint8_t* msgs, const uint8_t* sigs,
                             const uint8_t* scratch, size_t stride, uint8_t* ends, uint32_t prio,
                             bool paired = false) = 0;
    virtual void latb_tail(Launcher& L, const DLeaf* reqs, uint32_t n, const uint8_t* auth, const uint8_t* scratch,
                           size_t stride, uint8_t* ends, uint8_t* tmp, uint8_t* out, uint32_t tail,
                           uint32_t prio) = 0;
    // Subtree levels of a few runs split into address-only and finishing
    // work: pre_tree fills 2^th - 1 pre-images per run (tree_pre_slot
    // order), levels_fin builds levels 0 .. th-1 from them.
    virtual void pre_tree(Launcher& L, const void* keys, const DRun* runs, int nruns, uint32_t th, void* pre) = 0;
    virtual void levels_fin(Launcher& L, int nruns, uint32_t th, uint8_t* nodes, const void* pre) = 0;
    virtual void thash_h(Launcher& L, const void* keys, const DThash* reqs, uint32_t n, const uint8_t* in,
                         uint8_t* out) = 0;
    virtual void hmsg(Launcher& L, const void* keys, const DMsg* reqs, uint32_t n, uint32_t pad,
                      uint8_t* blob, uint8_t* R, uint8_t* dig, uint32_t prio = 0) = 0;
    // H_msg of verification requests only (no R to compute): for the SHA-256
    // families two waves per 64 messages (k_hmsg_pair), else k_hmsg.
    virtual void hmsg_verify(Launcher& L, const void* keys, const DMsg* reqs, uint32_t n, uint32_t pad,
                             uint8_t* blob, uint8_t* dig, uint32_t prio = 0) = 0;
    virtual size_t hpre_bytes() const = 0;
    // BDS signatures: R and H_msg of the nsign messages, the planned leaves
    // (LAT_LEAF requests) and the layer-0 WOTS signatures (LAT_SIGN requests
    // with LAT_ALL_STEPS) in one pre-image launch, which also does the
    // planned merges' address-only work, and one launch finishing the leaves
    // beside R / H_msg of the nmsg messages.  The nsign signature requests
    // (the first nmsg sign the digests, the rest are the upper-layer roots
    // the state update signs) only get their pre-images here; the state
    // update launch finishes the first nmsg.  scratch holds (nleaf + nsign)
    // x lat_bytes(0).
    // long: R / H_msg beside the leaves and the nmsg layer-0 signatures
    // left to bds_update (else H_msg first and the signatures beside the
    // leaves, into sig_out).  ends: the chain-end slots of leaf rows marked
    // LAT_ENDS_OUT / LAT_ENDS_IN (LEN x n bytes each).  vals (long messages
    // only, or null): every value along the nmsg layer-0 signatures' chains
    // (LEN x 16 nodes each) is hashed beside H_msg into it, and bds_update
    // then picks the signatures from it instead of walking the chains.
    virtual void bds_lat(Launcher& L, const void* keys, const DMsg* msgs, uint32_t nmsg, uint32_t pad, uint8_t* blob,
                         uint8_t* R, uint8_t* digests, const DLeaf* leaves, uint32_t nleaf, const DLeaf* sigs,
                         uint32_t nsign, uint8_t* scratch, const DThash* merges, uint32_t nmerge, void* merge_out,
                         uint8_t* leaf_out, uint8_t* sig_out, bool long_msgs, uint8_t* ends, uint8_t* vals,
                         const BdsFinDag& dag, void* mid) = 0;
    // The state updates, one workgroup per job; dag / lits: the planned
    // merges' sources (3 words per merge, parallel to merge_arena) and the
    // literal node values (jobs with maxlev 0 hash their merges one by one).
    // ... and, with sigs set (long messages), beside them the n layer-0 WOTS
    // signatures (sigs / digests, pre-images at sign_scratch) into sig_out.
    virtual void bds_update(Launcher& L, const void* keys, size_t lds, const DBds* jobs, uint32_t n,
                            uint8_t* sk_arena, const uint8_t* leaf_arena, const void* merge_arena,
                            const uint32_t* dag, const uint8_t* lits, const uint8_t* wpre, uint8_t* tails,
                            uint8_t* scratch, const DLeaf* sigs, const uint8_t* digests,
                            const uint8_t* sign_scratch, uint8_t* sig_out, const uint8_t* vals,
                            const uint32_t* dag_sync, const uint8_t* dag_vals) = 0;
};

template <class F>
struct FamOpsT final : FamOps {
    using K = typename F::Key;
    size_t key_bytes() const override { return sizeof(K); }
    void prepare(Launcher& L, const uint8_t* seeds, int nkeys, void* keys) override {
        L.run("prepare", k_prepare<F>, dim3(cdiv(nkeys, WG)), dim3(WG), seeds, nkeys, (K*)keys);
    }
    void chains(Launcher& L, const void* keys, const DRun* runs, int nruns, uint32_t leaf_off,
                uint32_t leaves, uint8_t* ends) override {
        L.run("chains", k_chains<F>, dim3(cdiv((uint64_t)leaves * F::LEN, WG), nruns), dim3(WG),
              (const K*)keys, runs, leaf_off, leaves, ends);
    }
    void ltree_level(Launcher& L, const void* keys, const DRun* runs, int nruns, uint32_t leaf_off,
                     uint32_t leaves, uint32_t l, uint32_t height, const uint8_t* src, uint32_t src_stride,
                     uint8_t* dst, uint32_t dst_stride, uint8_t* out, uint64_t out_stride,
                     uint64_t out_base) override {
        L.run("ltree", k_ltree_level<F>, dim3(cdiv((uint64_t)leaves * ((l + 1) / 2), WG), nruns), dim3(WG),
              (const K*)keys, runs, leaf_off, leaves, l, height, src, src_stride, dst, dst_stride, out,
              out_stride, out_base);
    }
    void level(Launcher& L, const void* keys, const DRun* runs, int nruns, uint32_t th, uint32_t lev,
               uint8_t* nodes) override {
        L.run("level", k_level<F>, dim3(cdiv(1ull << (th - lev - 1), WG), nruns), dim3(WG), (const K*)keys,
              runs, th, lev, nodes);
    }
    void wots(Launcher& L, const void* keys, const DWots* reqs, uint32_t n, const uint8_t* msgs,
              const uint8_t* in, uint8_t* out) override {
        L.run("wots", k_wots<F>, dim3(cdiv((uint64_t)n * F::LEN, WG)), dim3(WG), (const K*)keys, reqs, n, msgs,
              in, out);
    }
    void leaf_wg(Launcher& L, const void* keys, const DLeaf* reqs, uint32_t n, bool verify,
                 const uint8_t* msgs, const uint8_t* sigs, const uint8_t* auth, uint32_t th,
                 uint8_t* out) override {
        const dim3 block(((F::LEN + 63) / 64) * 64);
        if (verify)
            L.run("verify_wg", k_leaf_wg<F, true>, dim3(n), block, (const K*)keys, reqs, msgs, sigs, auth, th,
                  out);
        else
            L.run("leaf_wg", k_leaf_wg<F, false>, dim3(n), block, (const K*)keys, reqs, msgs, sigs, auth, th,
                  out);
    }
    size_t lat_bytes(uint32_t th) const override { return LatLayout<F>::bytes(th); }
    template <uint32_t MODE>
    void lat_pre_t(Launcher& L, const void* keys, const DLeaf* reqs, uint32_t n, const uint8_t* msgs,
                   uint32_t th, uint8_t* scratch) {
        static const char* name[3] = {"lat_pre_leaf", "lat_pre_sign", "lat_pre_verify"};
        const uint32_t items = F::LEN * 17 - 1 + th;  // starts, 15 F per chain, L-tree merges, auth levels
        L.run(name[MODE], k_lat_pre<F, MODE>, dim3(cdiv(items, WG), n), dim3(WG), (const K*)keys, reqs, msgs,
              scratch, LatLayout<F>::bytes(th));
    }
    template <uint32_t MODE>
    void lat_fin_t(Launcher& L, const DLeaf* reqs, uint32_t n, const uint8_t* msgs, const uint8_t* sigs,
                   const uint8_t* auth, uint32_t th, const uint8_t* scratch, uint8_t* out) {
        static const char* name[3] = {"lat_leaf", "lat_sign", "lat_verify"};
        L.run(name[MODE], k_lat_fin<F, MODE>, dim3(n), dim3(((F::LEN + 63) / 64) * 64), reqs, msgs, sigs, auth,
              scratch, LatLayout<F>::bytes(th), out);
    }
    void lat_pre(Launcher& L, const void* keys, uint32_t mode, const DLeaf* reqs, uint32_t n, const uint8_t* msgs,
                 uint32_t th, uint8_t* scratch) override {
        if (mode == LAT_LEAF) lat_pre_t<LAT_LEAF>(L, keys, reqs, n, msgs, th, scratch);
        else if (mode == LAT_SIGN) lat_pre_t<LAT_SIGN>(L, keys, reqs, n, msgs, th, scratch);
        else lat_pre_t<LAT_VERIFY>(L, keys, reqs, n, msgs, th, scratch);
    }
    void verify_pre(Launcher& L, const void* keys, const DMsg* msgs, uint32_t nmsg, uint32_t pad, uint8_t* blob,
                    uint8_t* digests, const DLeaf* reqs, uint32_t n, uint32_t th, uint8_t* scratch) override {
        const uint32_t xb = cdiv(F::LEN * 17 - 1 + th, WG);
        const uint32_t hrows = cdiv(nmsg, (uint64_t)xb * WG);
        L.run("verify_pre", k_ver_pre<F>, dim3(xb, hrows + n), dim3(WG), (const K*)keys, msgs, nmsg, pad, blob,
              digests, hrows, reqs, scratch, LatLayout<F>::bytes(th));
    }
    void verify_layers(Launcher& L, const DLeaf* reqs, uint32_t n, const uint32_t* base, uint32_t dmax,
                       uint32_t L0, uint32_t L1, const uint8_t* digests, const uint8_t* sigs, const uint8_t* auth,
                       uint32_t th, const uint8_t* scratch, uint8_t* roots, uint8_t* spec,
                       uint32_t nspec) override {
        LayerBase lb{};
        for (uint32_t i = 0; i <= dmax && i < 17; ++i) lb.b[i] = base[i];
        const uint32_t threads = ((F::LEN + 63) / 64) * 64;
        const uint32_t sblocks = cdiv((uint64_t)nspec * F::LEN * 16, threads);
        L.run("verify_layers", k_lat_verify_layers<F>, dim3(n + sblocks), dim3(threads), reqs, lb, L0, L1, digests,
              sigs, auth, scratch, LatLayout<F>::bytes(th), roots, n, spec, dmax > 1 ? base[1] : 0u, nspec);
    }
    void lat_leaves_dense(Launcher& L, const void* keys, const DLeaf* reqs, uint32_t n, uint8_t* scratch,
                          uint8_t* ends, uint8_t* out) override {
        const size_t stride = LatLayout<F>::bytes(0);
        lat_pre_t<LAT_LEAF>(L, keys, reqs, n, nullptr, 0, scratch);
        L.run("lat_chains_all", k_lat_chains_all<F>, dim3(cdiv((uint64_t)n * F::LEN, WG)), dim3(WG), n,
              (const uint8_t*)scratch, stride, ends);
        L.run("lat_ltree", k_lat_ltree<F>, dim3(n), dim3(LtreeThreads<F>::value), (const uint8_t*)ends,
              (const uint8_t*)scratch, stride, out);
    }
    void lat_pre_h(Launcher& L, const void* keys, const DLeaf* reqs, uint32_t n, uint32_t th,
                   uint8_t* scratch, uint32_t prio = 0) override {
        L.run("lat_pre_h", k_lat_pre_h<F>, dim3(cdiv((uint64_t)n * (F::LEN - 1 + th), WG)), dim3(WG),
              (const K*)keys, reqs, n, th, scratch, LatLayout<F>::bytes(th), prio);
    }
    void pre_need(Launcher& L, const void* keys, const DLeaf* reqs, uint32_t n, const uint8_t* msgs, uint32_t th,
                  uint8_t* scratch, uint32_t prio = 0) override {
        L.run("pre_need", k_pre_need<F>, dim3(n), dim3(256), (const K*)keys, reqs, msgs, scratch,
              LatLayout<F>::bytes(th), prio);
    }
    void lat_fin(Launcher& L, uint32_t mode, const DLeaf* reqs, uint32_t n, const uint8_t* msgs, const uint8_t* sigs,
                 const uint8_t* auth, uint32_t th, const uint8_t* scratch, uint8_t* out) override {
        if (mode == LAT_LEAF) lat_fin_t<LAT_LEAF>(L, reqs, n, msgs, sigs, auth, th, scratch, out);
        else if (mode == LAT_SIGN) lat_fin_t<LAT_SIGN>(L, reqs, n, msgs, sigs, auth, th, scratch, out);
        else lat_fin_t<LAT_VERIFY>(L, reqs, n, msgs, sigs, auth, th, scratch, out);
    }
    void lat(Launcher& L, const void* keys, uint32_t mode, const DLeaf* reqs, uint32_t n, const uint8_t* msgs,
             const uint8_t* sigs, const uint8_t* auth, uint32_t th, uint8_t* scratch, uint8_t* out) override {
        lat_pre(L, keys, mode, reqs, n, msgs, th, scratch);
        lat_fin(L, mode, reqs, n, msgs, sigs, auth, th, scratch, out);
    }
    void pre_tree(Launcher& L, const void* keys, const DRun* runs, int nruns, uint32_t th, void* pre) override {
        L.run("pre_tree", k_pre_tree<F>, dim3(cdiv((1ull << th) - 1, WG), nruns), dim3(WG), (const K*)keys, runs, th,
              (typename F::HPre*)pre);
    }
    void levels_fin(Launcher& L, int nruns, uint32_t th, uint8_t* nodes, const void* pre) override {
        const auto* hp = (const typename F::HPre*)pre;
        const uint32_t top = th > 6 ? th - 6 : 0;  // from here on one workgroup per run
        for (uint32_t lev = 0; lev < top; ++lev)
            L.run("level_fin", k_level_fin<F>, dim3(cdiv(1ull << (th - lev - 1), WG), nruns), dim3(WG), th, lev,
                  nodes, hp);
        L.run("top_fin", k_top_fin<F>, dim3(nruns), dim3(64), th, top, nodes, hp);
    }
    void latb_chains(Launcher& L, uint32_t n, const uint8_t* msgs, const uint8_t* sigs, const uint8_t* scratch,
                     size_t stride, uint8_t* ends, uint32_t prio, bool paired = false) override {
        if (paired)
            L.run("latb_chains2", k_latb_chains2<F>, dim3(cdiv((uint64_t)n * F::LEN, 2 * LATB_CHAINS_WG)),
                  dim3(LATB_CHAINS_WG), n, msgs, sigs, scratch, stride, ends, prio);
        else
            L.run("latb_chains", k_latb_chains<F>, dim3(cdiv((uint64_t)n * F::LEN, LATB_CHAINS_WG)),
                  dim3(LATB_CHAINS_WG), n, msgs, sigs, scratch, stride, ends, prio);
    }
    void latb_tail(Launcher& L, const DLeaf* reqs, uint32_t n, const uint8_t* auth, const uint8_t* scratch,
                   size_t stride, uint8_t* ends, uint8_t* tmp, uint8_t* out, uint32_t tail, uint32_t prio) override {
        if (tail != 0) {
            const dim3 grid(cdiv(n, LtreeBatch<F>::R)), block(LtreeBatch<F>::THREADS);
            if (tail == 2) {
                L.run("latb_tree_auth", k_latb_tree<F, true>, grid, block, reqs, n, (const uint8_t*)ends, auth,
                      scratch, stride, out, prio);
                return;
            }
            L.run("latb_tree", k_latb_tree<F, false>, grid, block, reqs, n, (const uint8_t*)ends, auth, scratch,
                  stride, tmp, prio);
            L.run("latb_auth", k_latb_auth<F>, dim3(cdiv(n, WG)), dim3(WG), reqs, n, (const uint8_t*)tmp, 1u, auth,
                  scratch, stride, out, prio);
            return;
        }
        uint32_t l = F::LEN, off = 0;
        bool from_a = true;
        const uint32_t sa = F::LEN, sb = (F::LEN + 1) / 2;
        while (l > 1) {
            L.run("latb_ltree", k_latb_ltree<F>, dim3(cdiv((uint64_t)n * ((l + 1) / 2), WG)), dim3(WG), n, l, off,
                  (const uint8_t*)(from_a ? ends : tmp), from_a ? sa : sb, from_a ? tmp : ends, from_a ? sb : sa,
                  scratch, stride, prio);
            off += l >> 1;
            l = (l + 1) >> 1;
            from_a = !from_a;
        }
        L.run("latb_auth", k_latb_auth<F>, dim3(cdiv(n, WG)), dim3(WG), reqs, n,
              (const uint8_t*)(from_a ? ends : tmp), from_a ? sa : sb, auth, scratch, stride, out, prio);
    }
    void thash_h(Launcher& L, const void* keys, const DThash* reqs, uint32_t n, const uint8_t* in,
                 uint8_t* out) override {
        L.run("thash_h", k_thash_h<F>, dim3(cdiv(n, WG)), dim3(WG), (const K*)keys, reqs, n, in, out);
    }
    void hmsg(Launcher& L, const void* keys, const DMsg* reqs, uint32_t n, uint32_t pad, uint8_t* blob,
              uint8_t* R, uint8_t* dig, uint32_t prio = 0) override {
        L.run("hmsg", k_hmsg<F>, dim3(cdiv(n, WG)), dim3(WG), (const K*)keys, reqs, n, pad, blob, R, dig, prio);
    }
    void hmsg_verify(Launcher& L, const void* keys, const DMsg* reqs, uint32_t n, uint32_t pad, uint8_t* blob,
                     uint8_t* dig, uint32_t prio = 0) override {
        if constexpr (F::kPairHmsg)
            L.run("hmsg_pair", k_hmsg_pair<F>, dim3(cdiv(n, 64)), dim3(128), (const K*)keys, reqs, n, pad, blob,
                  (uint8_t*)nullptr, dig, prio);
        else
            hmsg(L, keys, reqs, n, pad, blob, nullptr, dig, prio);
    }
    size_t hpre_bytes() const override { return sizeof(typename F::HPre); }
    void bds_lat(Launcher& L, const void* keys, const DMsg* msgs, uint32_t nmsg, uint32_t pad, uint8_t* blob,
                 uint8_t* R, uint8_t* digests, const DLeaf* leaves, uint32_t nleaf, const DLeaf* sigs, uint32_t nsign,
                 uint8_t* scratch, const DThash* merges, uint32_t nmerge, void* merge_out, uint8_t* leaf_out,
                 uint8_t* sig_out, bool long_msgs, uint8_t* ends, uint8_t* vals, const BdsFinDag& dag,
                 void* mid) override {
        const size_t stride = LatLayout<F>::bytes(0);
        const uint32_t xb = cdiv(F::LEN * 17 - 1, WG);  // LAT_LEAF / LAT_SIGN items per request
        const uint32_t mrows = cdiv(nmerge, (uint64_t)xb * WG);
        const uint32_t hrows = long_msgs ? 0 : cdiv(nmsg, (uint64_t)xb * WG);
        const uint32_t threads = ((F::LEN + 63) / 64) * 64;
        // one long message (mid set): its first H_msg blocks in the last row
        // of the pre-image launch, one wave (mirrored lanes)
        HmsgMid* hm = BdsFinPair<F>::value && long_msgs && nmsg == 1 ? (HmsgMid*)mid : nullptr;
        const uint32_t head = hm ? kHmsgHeadBlocks : 0;
        L.run("bds_pre", k_bds_pre<F>, dim3(xb, hrows + nleaf + nsign + mrows + (head ? 1 : 0)), dim3(WG),
              (const K*)keys, msgs, pad, blob, R, digests, hrows, nmsg, leaves, nleaf, sigs, nsign, scratch, stride,
              merges, nmerge, (typename F::HPre*)merge_out, hm, head);
        const uint32_t nsig = long_msgs ? 0 : nmsg, nlate = long_msgs ? nmsg : 0;
        const uint32_t nvals = long_msgs && vals ? nmsg : 0;
        // late H_msg: 64 messages per workgroup by its two waves (k_bds_fin)
        const uint32_t per = BdsFinPair<F>::value ? 64 : threads;
        // the merge-DAG block, if any, is the last one: it waits for the
        // planned leaves' blocks, which are dispatched before it
        L.run("bds_fin", k_bds_fin<F>, dim3(nleaf + nsig + nvals + cdiv(nlate, per) + (dag.on ? 1 : 0)),
              dim3(threads), (const K*)keys, msgs, nlate, pad, blob, R, digests, leaves, nleaf, sigs, nsig,
              (const uint8_t*)scratch, stride, leaf_out, sig_out, ends, nvals, vals, dag, (const HmsgMid*)hm);
    }
    void bds_update(Launcher& L, const void* keys, size_t lds, const DBds* jobs, uint32_t n,
                    uint8_t* sk_arena, const uint8_t* leaf_arena, const void* merge_arena, const uint32_t* dag,
                    const uint8_t* lits, const uint8_t* wpre, uint8_t* tails, uint8_t* scratch, const DLeaf* sigs,
                    const uint8_t* digests, const uint8_t* sign_scratch, uint8_t* sig_out,
                    const uint8_t* vals, const uint32_t* dag_sync, const uint8_t* dag_vals) override {
        L.run_lds("bds_update", k_bds_update<F>, dim3(sigs ? 2 * n : n), dim3(((F::LEN + 63) / 64) * 64), lds, (const K*)keys,
                  jobs, sk_arena, leaf_arena, (const typename F::HPre*)merge_arena, dag, lits, wpre,
                  LatLayout<F>::bytes(0), tails, scratch, n, sigs, digests, sign_scratch, sig_out, vals, dag_sync,
                  dag_vals);
    }
};

FamOps* ops_sha2_24();
FamOps* ops_sha2_32();
FamOps* ops_sha2_64();
FamOps* ops_shake128_32();
FamOps* ops_shake256_24();
FamOps* ops_shake256_32();
FamOps* ops_shake256_64();

}  // namespace xm
