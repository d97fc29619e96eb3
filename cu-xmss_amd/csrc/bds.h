// bds.h -- the BDS signing state machine (core_bds.cpp) and the hashing
// backend it runs on.  The C API drives it with a backend that launches on the
// engine directly; the serving loop (serve.cpp) with one that merges the work
// of many keys into shared launches.
#pragma once
#include <stdint.h>

#include <tuple>
#include <vector>

#include "../../include/params.h"
#include "engine.h"

namespace xm {

struct LeafId {
    uint32_t layer;
    uint64_t tree;
    uint32_t leaf;
    bool operator<(const LeafId& o) const {
        return std::tie(layer, tree, leaf) < std::tie(o.layer, o.tree, o.leaf);
    }
    bool operator==(const LeafId& o) const { return layer == o.layer && tree == o.tree && leaf == o.leaf; }
};

// Every hash one BDS signature needs, for the key bound last.
struct BdsOps {
    virtual ~BdsOps() {}
    // The key the following calls hash for (SK_SEED, SK_PRF, PUB_SEED as the
    // signature reads them from sk, i.e. after a possible wipe).
    virtual void bind_key(const unsigned char* sk_seed, const unsigned char* sk_prf,
                          const unsigned char* pub_seed) = 0;
    // gen_leaf_wots of one leaf (xmss_commons.c:104-113)
    virtual void leaf(unsigned char* out, const LeafId& id) = 0;
    // thash_h with a full ADRS (hash.c:125-155)
    virtual void thash_h(unsigned char* out, const unsigned char* in, const uint32_t adrs[8]) = 0;
    // wots_sign of an n-byte message at (layer, tree, ots) (wots.c:136-154)
    virtual void wots_sign(unsigned char* sig, const unsigned char* msg, uint32_t layer, uint64_t tree,
                           uint32_t ots) = 0;
    // R = PRF(SK_PRF, idx) and H_msg(R, root, idx, M) (xmss_core_fast.c:788-797)
    virtual void hmsg_sign(unsigned char* R, unsigned char* digest, uint64_t idx, const unsigned char* root,
                           const unsigned char* m, uint64_t mlen) = 0;
};

// Backend that computes nothing and records which leaves and which tree
// merges (their ADRS) were asked for, in order, and what each merge hashes.
// Each leaf and merge result it hands back is a tag (a fixed n-5 byte
// pattern, a kind byte, a 4-byte index), so that the inputs of a later
// merge name their sources: a tag is a leaf or an earlier merge, any other
// bytes are node values the sk itself holds (kept as literals).
struct RecordingOps final : BdsOps {
    unsigned n;
    std::vector<LeafId> leaves;
    std::vector<uint32_t> merges;     // 8 ADRS words per thash_h call
    std::vector<uint32_t> dag;        // 3 words per merge: left source, right source, level
    std::vector<unsigned char> lits;  // n bytes per literal source
    std::vector<uint32_t> wots_up;    // 4 words per upper-layer WOTS signature: layer, tree hi, tree lo, ots
    explicit RecordingOps(unsigned n_) : n(n_) {}
    void bind_key(const unsigned char*, const unsigned char*, const unsigned char*) override {}
    void leaf(unsigned char* out, const LeafId& id) override;
    void thash_h(unsigned char* out, const unsigned char*, const uint32_t*) override;
    void wots_sign(unsigned char* sig, const unsigned char*, uint32_t layer, uint64_t tree, uint32_t ots) override;
    void hmsg_sign(unsigned char* R, unsigned char* digest, uint64_t, const unsigned char*, const unsigned char*,
                   uint64_t) override;
};

// One BDS signature (xmss_core_fast.c:710-851 for d = 1, :917-1121 for d > 1)
// with all hashing on `ops`.  sk is the OID-less secret key.
int bds_sign(const xmss_params* p, unsigned char* sk, unsigned char* sm, unsigned long long* smlen,
             const unsigned char* m, unsigned long long mlen, BdsOps& ops);

// The leaves bds_sign would consume for this sk (a dry run on a copy):
// distinct and sorted, or every request in consumption order.
std::vector<LeafId> bds_plan_leaves(const xmss_params* p, const unsigned char* sk, unsigned long long mlen,
                                    bool in_order = false);

// The merges (thash_h ADRS, 8 words each) bds_sign would hash for this sk, in
// order, beside the leaves in order (a dry run on a copy).
// A source is kind | index (Engine::BdsJob::plan_dag); the level of a merge is
// 1 + the deepest level among the merges it depends on (leaves and literals
// are level 0), so that merges of one level can be hashed side by side.
// post, if given, receives the dry run's sk after the signature: its
// traversal structure (indices, stack heights, flags) is the real one, its
// node values are tags.
void bds_plan(const xmss_params* p, const unsigned char* sk, unsigned long long mlen, std::vector<LeafId>& leaves,
              std::vector<uint32_t>& merges, std::vector<uint32_t>* dag = nullptr,
              std::vector<unsigned char>* lits = nullptr, std::vector<uint32_t>* wots_up = nullptr,
              std::vector<unsigned char>* post = nullptr);

// The distinct leaves the signature made with `sk` will consume, from a dry
// run on a copy.  Which leaves a signature asks for depends on the
// traversal's structure and the index only, never on node values, so sk
// may be a dry run's `post` state.  Empty when that signature is the last
// index's or beyond (those take the host path, bds_sign).
std::vector<LeafId> bds_leaves_of(const xmss_params* p, const unsigned char* sk);

// A signature whose traversal runs on the device (Engine::bds_sign_batch).
struct BdsDeviceSign {
    std::vector<uint32_t> layer, leaf, merges, dag, wots_up;
    std::vector<unsigned char> lits;
    std::vector<uint64_t> tree;
    std::vector<unsigned char> seeds;  // SK_SEED, SK_PRF, PUB_SEED, root
    unsigned long long idx = 0;
    std::vector<LeafId> next;  // the leaves of the key's next signature (bds_leaves_of)
    Engine::BdsJob job{};
};
// Host part before the launch: index, plan, index advance, message copy.
// Returns false for the last index / an exhausted key (use bds_sign).
bool bds_device_prepare(const xmss_params* p, unsigned char* sk, unsigned char* sm, const unsigned char* m,
                        unsigned long long mlen, BdsDeviceSign& ds);
void bds_device_finish(const xmss_params* p, unsigned char* sm, unsigned long long* smlen, unsigned long long mlen,
                       const BdsDeviceSign& ds);
Engine::BdsShape bds_shape(const xmss_params* p);

}  // namespace xm
