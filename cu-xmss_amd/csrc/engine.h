// engine.h -- host-side interface of the device engine (internal to the
// library; the public C ABI lives in include/).
//
// The engine owns three HIP streams on the current device (two for kernels,
// one for the copies of launched batches) and grow-only workspaces.  Calls
// are synchronous -- inputs are host byte arrays, outputs are written back
// to host byte arrays before the call returns -- except the batch calls
// below, whose launch / wait split lets a caller overlap host work with the
// GPU.  The cores (core_plain.cpp, core_bds.cpp, ext.cpp, serve.cpp) express
// all hashing through these calls; no hash is ever computed on the host.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <sys/types.h>
#include <unistd.h>

#include <algorithm>
#include <condition_variable>
#include <functional>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace xm {

// Hash families: the (func, n) pairs accepted by core_hash (cuxmss/hash.c:29-65).
enum Fam : int {
    F_SHA2_24 = 0,
    F_SHA2_32,
    F_SHA2_64,
    F_SHAKE128_32,
    F_SHAKE256_24,
    F_SHAKE256_32,
    F_SHAKE256_64,
    F_COUNT
};

// A failed HIP call, an exhausted device allocation or a request the launch
// sequences cannot hold: thrown inside the library and returned as -3 at the
// C ABI (core_common.h, guarded), never an abort of the caller's process.
struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Device bytes held by the library (workspaces, key pools, resident serving
// state) and the limit on them (0: none); set_device_limit returns the old one.
size_t device_bytes();
size_t set_device_limit(size_t bytes);

// One device allocation carved up by a thread-safe first-fit allocator
// (256-byte granules): the serving loop's whole device memory, so that its
// HBM budget bounds real memory.  alloc returns nullptr when no free range
// is large enough.
class DeviceArena {
public:
    explicit DeviceArena(size_t bytes);  // throws DeviceError when the device cannot hold it
    // An arena over memory the caller owns (a block of another arena).
    DeviceArena(void* base, size_t bytes);
    ~DeviceArena();
    void* alloc(size_t bytes);
    void release(void* p);
    bool owns(const void* p) const;
    void* base() const { return base_; }
    size_t size() const { return size_; }
    size_t in_use() const;
    size_t peak() const;

private:
    uint8_t* base_ = nullptr;
    size_t size_ = 0;
    bool owner_ = true;
    mutable std::mutex mu_;
    std::map<size_t, size_t> free_;  // offset -> length
    std::map<size_t, size_t> used_;  // offset -> length
    size_t in_use_ = 0, peak_ = 0;
};

// While an ArenaScope is alive, the device allocations this thread's engine
// calls make come from the arena (exact sizes, no spare growth) instead of
// hipMalloc, and the arena counts them; frees return them.
struct ArenaScope {
    explicit ArenaScope(DeviceArena* a);
    ~ArenaScope();
    DeviceArena* prev;
};

// VALU issue-rate probe (probe.hip): lane-ops/s and kernel ms of one launch
// of instruction mix `kind` (0: SHA-256 chain step; 1-5: single classes).
constexpr int VALU_PROBE_KINDS = 6;
int valu_probe(int kind, uint32_t blocks, uint32_t iters, double* lane_ops_per_s, double* ms);

// func codes of xmss_params.func (params.h:7-9)
int fam_of(unsigned func, unsigned n);  // -1 when core_hash would fail
unsigned fam_n(int fam);
unsigned fam_len(int fam);              // wots_len for w = 16
unsigned fam_pad(int fam);              // padding_len

struct KeySeeds {
    const uint8_t* sk_seed;  // n bytes or nullptr (zeros)
    const uint8_t* sk_prf;
    const uint8_t* pub_seed;
};

// A run of leaves of one (layer, tree) of one key: leaves leaf0 .. leaf0+count-1.
struct LeafRun {
    uint32_t key;
    uint32_t layer;
    uint64_t tree;
    uint32_t leaf0;
};

// A node of a built subtree: tree number in the last build_trees call,
// height (0 = leaves) and index at that height.
struct NodeRef {
    uint32_t tree;
    uint32_t level;
    uint32_t index;
};

enum WotsMode : uint32_t { WOTS_PKGEN = 0, WOTS_SIGN = 1, WOTS_VERIFY = 2 };

struct WotsReq {
    uint32_t key, layer;
    uint64_t tree;
    uint32_t ots;
    uint32_t mode;  // WotsMode
};

struct LtreeReq {
    uint32_t key, layer;
    uint64_t tree;
    uint32_t ltree;
};

struct RootReq {
    uint32_t key, layer;
    uint64_t tree;
    uint32_t leaf_idx;
};

struct ThashReq {
    uint32_t key;
    uint32_t adrs[8];
};

// H_msg request: byte string prefix || M with the prefix laid out as in
// hash_message (cuxmss/hash.c:107-120).  For sign requests R is computed on
// the device from SK_PRF and idx and written into the prefix first.
struct MsgReq {
    uint32_t key;
    uint32_t sign;
    uint64_t idx;
    const uint8_t* R;     // verify: n bytes from the signature; sign: unused
    const uint8_t* root;  // n bytes
    const uint8_t* m;
    uint64_t mlen;
};

// Host worker threads for parallel_for, started on first use and kept for
// the process (a forked child starts its own).  One parallel_for at a time
// uses them; a concurrent caller starts threads of its own.
class HostPool {
public:
    static HostPool* get(size_t workers) {
        static std::mutex mu;
        static HostPool* pool = nullptr;
        std::lock_guard<std::mutex> lk(mu);
        if (!pool || pool->pid_ != getpid()) pool = new HostPool(workers);  // a fork's pool is not reused
        return pool;
    }
    size_t workers() const { return nw_; }
    bool try_acquire() { return busy_.try_lock(); }
    void release() { busy_.unlock(); }
    // Runs f(k) for k in [0, t), t <= workers() + 1: f(0) on the caller, the
    // rest on the workers; returns when all are done.
    void run(size_t t, const std::function<void(size_t)>& f) {
        {
            std::lock_guard<std::mutex> lk(m_);
            task_ = &f;
            ntask_ = t;
            pending_ = t - 1;
            ++gen_;
        }
        go_.notify_all();
        f(0);
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [&] { return pending_ == 0; });
    }

private:
    explicit HostPool(size_t w) : nw_(w), pid_(getpid()) {
        for (size_t k = 1; k <= w; ++k) std::thread([this, k] { loop(k); }).detach();
    }
    void loop(size_t k) {
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(size_t)>* t;
            {
                std::unique_lock<std::mutex> lk(m_);
                go_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
                if (k >= ntask_) continue;
                t = task_;
            }
            (*t)(k);
            std::lock_guard<std::mutex> lk(m_);
            if (--pending_ == 0) done_.notify_one();
        }
    }
    const size_t nw_;
    const pid_t pid_;
    std::mutex busy_, m_;
    std::condition_variable go_, done_;
    const std::function<void(size_t)>* task_ = nullptr;
    size_t ntask_ = 0, pending_ = 0;
    uint64_t gen_ = 0;
};

// Runs f(i) for every i in [0, n) on up to eight host threads (this one
// included) when there are at least `grain` items per thread; the items must
// touch disjoint memory.  Used for the per-request staging copies of large
// batches.
template <class Fn>
void parallel_for(size_t n, size_t grain, Fn&& f) {
    const size_t hw = std::max<size_t>(1, std::thread::hardware_concurrency());
    const size_t t = std::min<size_t>(std::min<size_t>(8, hw), n / std::max<size_t>(grain, 1));
    if (t <= 1) {
        for (size_t i = 0; i < n; ++i) f(i);
        return;
    }
    const size_t chunk = (n + t - 1) / t;
    const std::function<void(size_t)> part = [&f, chunk, n](size_t k) {
        for (size_t i = k * chunk; i < std::min(n, (k + 1) * chunk); ++i) f(i);
    };
    HostPool* pool = HostPool::get(std::min<size_t>(8, hw) - 1);
    if (t <= pool->workers() + 1 && pool->try_acquire()) {
        pool->run(t, part);
        pool->release();
        return;
    }
    std::vector<std::thread> th;
    for (size_t k = 1; k < t; ++k) th.emplace_back(part, k);
    part(0);
    for (std::thread& x : th) x.join();
}

class Engine {
public:
    // The engine of the current HIP device, created on first use.  Returns
    // nullptr (and prints why) when no usable GPU is present.
    static Engine* get();
    // A second engine on the current device with its own stream and buffers,
    // for work that runs concurrently with the shared engine (the serving
    // loop's verifications).  Not shared; the caller serialises its use.
    static Engine* make_private();
    // Hands a private engine back for reuse by a later make_private.
    static void release_private(Engine* e);

    // Device memory that outlives calls (resident key state of the serving
    // loop): allocation, release and a synchronous host-to-device copy on the
    // current device.  device_alloc returns nullptr when the allocation fails
    // or would pass the library's device limit.
    static void* device_alloc(size_t bytes);
    static void device_free(void* p, size_t bytes);
    static void upload(void* dst, const void* src, size_t bytes);

    // After a DeviceError: back to a consistent launcher state (main stream,
    // no pending timings), streams drained.
    void recover();

    void set_keys(int fam, int nkeys, const KeySeeds* seeds);
    // Overwrites the key contexts and seeds of the last set_keys with zeros
    // (queued on the engine's stream).
    void scrub_keys();

    // Device-resident key contexts that outlive a call: after pool_use(fam)
    // the request `key` fields of the calls below index the pool of that
    // family (until the next set_keys).  Slots are found by a keyed digest
    // of the seeds; no seed bytes are kept on the host.
    // pool_add: a pinned slot (lives until pool_release);
    int pool_add(int fam, const KeySeeds& seeds);
    void pool_release(int fam, int slot);
    // pool_slot: the slot of these seeds, else a cached slot -- at most
    // kPoolCached of them per family, the least recently used one of an
    // earlier call (pool_begin) evicted and zeroed first;
    int pool_slot(int fam, const KeySeeds& seeds);
    void pool_begin();
    // pool_forget: drops the cached slot of these seeds (a wiped key).
    void pool_forget(int fam, const KeySeeds& seeds);
    size_t pool_cached(int fam) const;
    // Chain ends hashed one signature ahead, for a key that signs one
    // signature after another: the next bds_sign_batch of one job for the
    // key of these seeds also hashes the chains of the leaves `ids`
    // (LeafRun.key unused, one leaf each) -- those its next signature will
    // consume -- into device slots, beside its own WOTS signature; a later
    // signature of the key then hashes only the L-trees of those leaves.
    // The slots follow one key (found by the keyed digest of its seeds).
    void ends_ahead(int fam, const KeySeeds& seeds, const LeafRun* ids, int count);
    // Forgets every chain end held (a wiped key).
    void ends_ahead_drop();
    void pool_use(int fam);
    size_t pool_key_bytes(int fam) const;
    static size_t pool_key_bytes_of(int fam);
    // Synchronous device-to-device copy / zero fill (outside any stream order).
    static void copy_device(void* dst, const void* src, size_t bytes);
    static void clear_device(void* p, size_t bytes);
    // Frees every device buffer the engine holds -- workspaces, spare batches,
    // key pools -- and zeroes the pools first (a private engine handed to a
    // server, whose device memory must be its arena alone).
    void trim();
    // Key contexts of `fam` kept by the caller (the serving loop's key
    // tables): the request `key` fields index `table` until the next
    // set_keys or pool_use.  prepare_keys writes the contexts of `nkeys`
    // keys (3n seed bytes each) into `table` at slot `first` (synchronous).
    void use_keys(int fam, const void* table);
    void prepare_keys(int fam, const KeySeeds* seeds, int nkeys, void* table, int first);
    // Frees every workspace buffer of the engine (not its key pool or
    // batches): under an ArenaScope, nothing is held between calls.
    void release_workspace();

    // Full subtrees of height th, one per run (leaf0 must be 0).  All node
    // layers stay on the device for fetch_nodes until the next call.
    void build_trees(int fam, unsigned th, int nruns, const LeafRun* runs);
    void fetch_nodes(int fam, unsigned th, int nrefs, const NodeRef* refs, uint8_t* out);

    // gen_leaf_wots for `count` consecutive leaves of each run (xmss_commons.c:104-113).
    void leaves(int fam, int nruns, const LeafRun* runs, unsigned count, uint8_t* out);

    void wots(int fam, int n, const WotsReq* reqs, const uint8_t* msgs, const uint8_t* in,
              uint8_t* out);
    // One layer of signature verification per request (xmss_commons.c:362-375):
    // WOTS pk from the signature, L-tree, then the auth-path walk -> root.
    void verify_layer(int fam, unsigned th, int n, const RootReq* reqs, const uint8_t* msgs,
                      const uint8_t* sigs, const uint8_t* auth, uint8_t* roots);
    void ltree(int fam, int n, const LtreeReq* reqs, const uint8_t* pks, uint8_t* out);
    void compute_root(int fam, unsigned th, int n, const RootReq* reqs, const uint8_t* leaves,
                      const uint8_t* auth, uint8_t* out);
    void thash_h(int fam, int n, const ThashReq* reqs, const uint8_t* in, uint8_t* out);
    void hmsg(int fam, int n, const MsgReq* reqs, uint8_t* R_out, uint8_t* digest_out);

    // ---- device-resident batches (plain-core sign, verify) ----
    // A batch is staged to the device by *_prepare, computed by run() with no
    // host traffic, and read back by collect().  All requests of a batch share
    // one parameter set.
    struct SignItem {
        const uint8_t* sk_seed;
        const uint8_t* sk_prf;
        const uint8_t* pub_seed;
        const uint8_t* root;  // pub root (H_msg input)
        uint64_t idx;         // signature index
        const uint8_t* m;
        uint64_t mlen;
    };
    struct VerifyItem {
        const uint8_t* pk;  // root || PUB_SEED
        uint64_t idx;
        const uint8_t* R;
        const uint8_t* wots;  // d consecutive (WOTS sig || auth path) layers, as in sm
        const uint8_t* m;
        uint64_t mlen;
    };
    struct Batch;
    // one_key: all items sign with the same key (its seeds from items[0]);
    // items whose signatures use the same (layer, tree) share one subtree build.
    // wait = false returns before the staging copy has completed: the next
    // prepare on this engine must then come after a run + collect (whose
    // copies are ordered after it on the same stream).
    Batch* sign_prepare(int fam, unsigned th, unsigned d, int count, const SignItem* items, bool one_key = false,
                        bool wait = true);
    // A signing batch of the job API: its run sizes every workspace in a pass
    // without launches, then captures its launches into one HIP graph and
    // launches that (replayed if the batch runs again).
    void set_graphed(Batch* b, bool on);
    Batch* verify_prepare(int fam, unsigned th, unsigned d, int count, const VerifyItem* items);
    // New inputs for a verification batch of the same shape (count items of
    // its parameter set), staged into its buffers; its captured launch graph
    // stays valid unless the messages no longer fit (then it is dropped).
    void verify_refill(Batch* b, const VerifyItem* items);
    // A verification batch whose launches all go on the engine's main stream
    // (default: its large batches split into two groups on two streams).
    // Batches of several engines then run side by side on as many hardware
    // queues instead of contending for them (mixed jobs).
    void set_one_stream(Batch* b, bool on);
    // The stream of this engine a verification batch's launches go to (0:
    // main, 1: aux, 2: the copy stream).  The three streams of an engine are
    // created one after another, so HIP gives them three different hardware
    // queues: the sets of a mixed job, each kept to one stream (set_one_stream)
    // on one engine, run side by side whatever streams the process created
    // before.
    void set_stream(Batch* b, int sel);
    // wait_done = false returns once the work is queued; wait() then blocks
    // until this engine's streams are idle.  A verification batch's launches
    // are captured into a HIP graph on its second run and replayed from then
    // on (not while per-kernel timing is on).
    void run(Batch* b, bool wait_done = true);
    void wait();
    // launch() queues a batch's work and returns; wait_batch() blocks until
    // that work is done (no lock is needed for it).  A launched batch's
    // refill, collect and free wait for it by themselves; its input copy runs
    // on a copy stream and its result copy on the aux stream, so a
    // verification batch can be staged and another collected while it runs.
    void launch(Batch* b);
    void wait_batch(Batch* b);
    // sign: per item R (n) || d x (WOTS sig || auth path) written to out[i]
    //       (sig_bytes - index_bytes bytes each).
    // verify: ok[i] = 1 when the recomputed root equals the pk root.
    void collect_sign(Batch* b, uint8_t* const* out);
    void collect_verify(Batch* b, int* ok);
    void free_batch(Batch* b);

    // ---- BDS signatures with the traversal on the device ----
    // For each job: R and H_msg, the planned leaves (in the order the
    // signature consumes them), the layer-0 WOTS signature and the state
    // update of the serialized sk, in three launches on one stream per chunk
    // of jobs (two for a job launched alone, sig_values); sk (or its resident
    // copy dsk) is updated in place.  key is the job's slot in the device key
    // pool.
    struct BdsJob {
        const uint8_t* sk_seed;
        const uint8_t* sk_prf;
        const uint8_t* pub_seed;
        const uint8_t* pub_root;
        uint32_t key;          // pool slot
        uint8_t* sk;           // OID-less sk, index already advanced
        uint8_t* dsk;          // resident copy of the sk in HBM (index not yet advanced), or null:
                               // the update runs on it in place and sk receives the result
        uint64_t idx;          // index being signed
        const uint8_t* m;
        uint64_t mlen;
        const uint32_t* plan_layer;  // planned leaves
        const uint64_t* plan_tree;
        const uint32_t* plan_leaf;
        uint32_t nplan;
        const uint32_t* plan_merge;  // planned tree merges: 8 ADRS words each, in hashing order
        uint32_t nmerge;
        // What each planned merge hashes, or null (merges hashed one after
        // another as the state machine reaches them): 3 words per merge --
        // the sources of its left and right halves (DAG_LIT | index into
        // plan_lit, DAG_LEAF | index of a planned leaf, DAG_MERGE | index of
        // an earlier merge) and its level (1 + the deepest merge it needs).
        const uint32_t* plan_dag;
        const uint8_t* plan_lit;     // n bytes per literal (node values the sk holds)
        uint32_t nlit;
        // XMSS^MT subtree roots the update signs (layer >= 1), 4 words each:
        // layer, tree hi, tree lo, ots -- their address-only work is done in
        // the pre-image launch; null: whole chains in the update
        const uint32_t* plan_wots;
        uint32_t nwots;
        // a signature launched alone (the one-call path, a serving group of
        // one): with a long message its layer-0 WOTS chains are walked to
        // every value beside H_msg, and the merge DAG and the state update
        // run in the finishing launch (Engine::bds_bytes counts the memory)
        bool sig_values;
        uint8_t* R_out;        // n
        uint8_t* wots0_out;    // wots_sig_bytes
        uint8_t* tail_out;     // sig_bytes - index_bytes - n - wots_sig_bytes
        struct Shape {
            unsigned th, d, full_height, index_bytes, bds_k, wots_sig_bytes, sig_bytes;
            unsigned long long sk_bytes;
        } shape;
    };
    using BdsShape = BdsJob::Shape;
    enum : uint32_t { DAG_LIT = 0u << 28, DAG_LEAF = 1u << 28, DAG_MERGE = 2u << 28, DAG_INDEX = (1u << 28) - 1 };

    // Verification of signatures of keys in the device pool of `fam`
    // (pool_slot with PUB_SEED set), any mix of parameter shapes, in one
    // launch sequence: one staging copy in, H_msg beside the address-only work
    // of every layer, then one verify-layer launch per XMSS^MT layer, one copy
    // out.  ok[i] = 1 when signature i is valid.
    struct VerifyJob {
        uint32_t key;            // pool slot
        unsigned th, d, index_bytes;
        const uint8_t* pk_root;  // n bytes
        const uint8_t* sig;      // idx || R || d x (WOTS signature || auth path)
        const uint8_t* m;
        uint64_t mlen;
    };
    void verify_many(int fam, int count, const VerifyJob* jobs, int* ok);
    // One verification with no device allocation: every buffer lives in a
    // fixed device array of static_capacity() bytes per GPU (pk_root must
    // point at root || PUB_SEED).  Returns 0 (ok set), or -4 when the request
    // needs more than the array holds (static_bytes).
    int verify_static(int fam, const VerifyJob& job, int* ok);
    static size_t static_capacity();
    static size_t static_bytes(int fam, unsigned th, unsigned d, uint64_t mlen);
    // Jobs may mix parameter shapes of one hash family; their keys are pool
    // slots of that family (BdsJob::key).
    void bds_sign_batch(int fam, int count, BdsJob* jobs);
    // Device bytes one BDS signature / one verification takes in
    // bds_sign_batch / verify_many (its share of a launch group, rounding
    // included), by the formulas those calls allocate with.
    static size_t bds_bytes(int fam, const BdsJob& job);
    static size_t verify_bytes(int fam, unsigned th, unsigned d, uint64_t mlen);

    // Per-kernel HIP-event timing (off by default).
    void set_profiling(bool on);
    bool profiling() const { return prof_on_; }
    struct KStat {
        std::string name;
        uint64_t count;
        double total_ms;
    };
    std::vector<KStat> stats() const;
    void reset_stats();

    void* stream() const { return stream_; }
    int device() const { return dev_; }

    // implementation detail, public for the launch helpers
    struct Impl;
    Impl* impl_;

private:
    Engine();
    void bds_sign_chunk(int fam, int count, BdsJob* jobs);
    void* stream_ = nullptr;
    int dev_ = 0;
    bool prof_on_ = false;
};

// Autotuner of the verification-batch schedule (engine.hip): forces a plan
// for jobs run from now on (plan >= 0; -1: tuned per shape, the default);
// returns the previous setting.  verify_plans lists the candidate plans;
// tuned_plans describes the settled shapes, one line each.
int set_verify_plan(int plan);
size_t verify_plans(uint32_t* out, size_t max);
std::string tuned_plans();

}  // namespace xm
