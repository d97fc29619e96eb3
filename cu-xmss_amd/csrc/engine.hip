// engine.hip -- the device engine: stream, grow-only workspace, host <-> device
// staging, and the sequencing of kernel launches for each operation.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <random>
#include <array>
#include <string>
#include <tuple>
#include <unordered_map>

#include "engine.h"
#include "launch.h"

namespace xm {

[[noreturn]] void fatal(const char* what) {
    fprintf(stderr, "xmss_mi355x: %s\n", what);
    fflush(stderr);
    throw DeviceError(what);
}

[[noreturn]] void hip_failed(hipError_t e, const char* file, int line, const char* what) {
    char msg[512];
    snprintf(msg, sizeof msg, "HIP error %s at %s:%d (%s)", hipGetErrorString(e), file, line, what);
    fprintf(stderr, "xmss_mi355x: %s\n", msg);
    fflush(stderr);
    throw DeviceError(msg);
}

// Bytes of device memory the library holds, and the limit on them.
static std::atomic<size_t> g_dev_bytes{0};
static std::atomic<size_t> g_dev_limit{0};

// The arena this thread's device allocations come from (ArenaScope), if any.
static thread_local DeviceArena* t_arena = nullptr;

ArenaScope::ArenaScope(DeviceArena* a) : prev(t_arena) { t_arena = a; }
ArenaScope::~ArenaScope() { t_arena = prev; }

DeviceArena::DeviceArena(size_t bytes) {
    size_ = (std::max<size_t>(bytes, 256) + 255) & ~(size_t)255;
    base_ = (uint8_t*)dev_malloc(size_);  // counted once, as a whole
    free_[0] = size_;
}
DeviceArena::DeviceArena(void* base, size_t bytes) : base_((uint8_t*)base), size_(bytes & ~(size_t)255), owner_(false) {
    free_[0] = size_;
}
DeviceArena::~DeviceArena() {
    if (!owner_) return;
    try {
        dev_free(base_, size_);
    } catch (const DeviceError&) {
    }
}
void* DeviceArena::alloc(size_t bytes) {
    bytes = (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
    std::lock_guard<std::mutex> lk(mu_);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < bytes) continue;
        const size_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > bytes) free_[off + bytes] = len - bytes;
        used_[off] = bytes;
        in_use_ += bytes;
        peak_ = std::max(peak_, in_use_);
        return base_ + off;
    }
    return nullptr;
}
void DeviceArena::release(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(mu_);
    const size_t off = (size_t)((uint8_t*)p - base_);
    auto u = used_.find(off);
    if (u == used_.end()) return;
    size_t start = off, len = u->second;
    in_use_ -= len;
    used_.erase(u);
    auto nx = free_.lower_bound(start);
    if (nx != free_.end() && nx->first == start + len) {  // merge with the next free range
        len += nx->second;
        nx = free_.erase(nx);
    }
    if (nx != free_.begin()) {  // and with the previous one
        auto pv = std::prev(nx);
        if (pv->first + pv->second == start) {
            start = pv->first;
            len += pv->second;
            free_.erase(pv);
        }
    }
    free_[start] = len;
}
bool DeviceArena::owns(const void* p) const { return p >= base_ && p < base_ + size_; }
size_t DeviceArena::in_use() const {
    std::lock_guard<std::mutex> lk(mu_);
    return in_use_;
}
size_t DeviceArena::peak() const {
    std::lock_guard<std::mutex> lk(mu_);
    return peak_;
}

void* dev_malloc(size_t bytes) {
    if (DeviceArena* a = t_arena) {
        void* p = a->alloc(bytes);
        if (!p) {
            char msg[160];
            snprintf(msg, sizeof msg, "%zu device bytes do not fit in the serving budget's free space", bytes);
            fatal(msg);
        }
        return p;
    }
    const size_t lim = g_dev_limit.load();
    const size_t now = g_dev_bytes.fetch_add(bytes) + bytes;
    if (lim && now > lim) {
        g_dev_bytes.fetch_sub(bytes);
        char msg[160];
        snprintf(msg, sizeof msg, "device allocation of %zu bytes beyond the library's limit of %zu", bytes, lim);
        fatal(msg);
    }
    void* p = nullptr;
    const hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        g_dev_bytes.fetch_sub(bytes);
        (void)hipGetLastError();  // an out-of-memory error is not sticky: clear it
        char msg[160];
        snprintf(msg, sizeof msg, "hipMalloc of %zu bytes failed (%s)", bytes, hipGetErrorString(e));
        fatal(msg);
    }
    return p;
}

void dev_free(void* p, size_t bytes) {
    if (!p) return;
    if (t_arena && t_arena->owns(p)) {
        t_arena->release(p);
        return;
    }
    g_dev_bytes.fetch_sub(bytes);
    HIPCK(hipFree(p));
}

size_t device_bytes() { return g_dev_bytes.load(); }
size_t set_device_limit(size_t bytes) { return g_dev_limit.exchange(bytes); }

int fam_of(unsigned func, unsigned n) {
    // the (func, n) pairs core_hash accepts (cuxmss/hash.c:36-63)
    if (func == 0 && n == 24) return F_SHA2_24;
    if (func == 0 && n == 32) return F_SHA2_32;
    if (func == 0 && n == 64) return F_SHA2_64;
    if (func == 1 && n == 32) return F_SHAKE128_32;
    if (func == 2 && n == 24) return F_SHAKE256_24;
    if (func == 2 && n == 32) return F_SHAKE256_32;
    if (func == 2 && n == 64) return F_SHAKE256_64;
    return -1;
}
unsigned fam_n(int fam) {
    static const unsigned n[F_COUNT] = {24, 32, 64, 32, 24, 32, 64};
    return n[fam];
}
unsigned fam_len(int fam) { return 2 * fam_n(fam) + 3; }
unsigned fam_pad(int fam) { return fam_n(fam) == 24 ? 4 : fam_n(fam); }

static FamOps* ops_of(int fam) {
    switch (fam) {
        case F_SHA2_24: return ops_sha2_24();
        case F_SHA2_32: return ops_sha2_32();
        case F_SHA2_64: return ops_sha2_64();
        case F_SHAKE128_32: return ops_shake128_32();
        case F_SHAKE256_24: return ops_shake256_24();
        case F_SHAKE256_32: return ops_shake256_32();
        case F_SHAKE256_64: return ops_shake256_64();
    }
    fatal("unsupported hash family");
}

// Leaves, WOTS signatures and verify layers go through the two-launch
// kernels of kernels_lat.h whenever their scratch fits: they are faster at
// every batch size measured, not only for single requests (4096
// XMSS-SHA2_10_256 verifications: 2.4 M/s, against 1.3 M/s with the
// one-workgroup-per-request kernel, whose lanes each walk a whole chain).
constexpr size_t kLatScratch = 4ull << 30;
// Up to this many subtrees per build, their levels use the split
// (pre-image + finish) kernels.
constexpr int kSplitRuns = 64;
// From this many requests of one layer on, a verification batch finishes
// lane-dense (per chain, per L-tree pair, per auth walk) instead of with one
// workgroup per request.
constexpr int kLatDense = 256;
// The schedule of a large verification batch (verify_dense), a bit set:
// VPLAN_PRIO            the long-message group's kernels at wave priority 2
// VPLAN_PRE_H_FIRST     the H pre-images on a third stream beside H_msg
//                       (else on the main stream after its F pre-images)
// VPLAN_PRIO_HMSG       every H_msg at wave priority 2
// tail (bits 2-3)       0: one launch per L-tree height, then the auth walks;
//                       1: the L-tree through LDS in one launch, then the auth
//                       walks; 2: L-tree and auth walk in one launch
// div (bits 4-7)        the long group is the longest 1/div of the batch (0: 4)
// VPLAN_SHORT_FIRST     the short messages' group launched first: the stream
//                       that starts first starts tens of us earlier, and the
//                       short group holds three quarters of the work
// VPLAN_PAIRED          two chains of complementary length per lane (k_latb_chains2)
constexpr uint32_t VPLAN_PRIO = 1, VPLAN_PRE_H_FIRST = 2, VPLAN_PRIO_HMSG = 256, VPLAN_SHORT_FIRST = 512,
                   VPLAN_PAIRED = 1024;
static inline uint32_t vplan_tail(uint32_t p) { return (p >> 2) & 3; }
static inline uint32_t vplan_div(uint32_t p) { return ((p >> 4) & 15) ? ((p >> 4) & 15) : 4; }
constexpr uint32_t kVplanDefault = (1u << 2) | 512u;  // tail 1, two groups, the short group first
uint32_t verify_plan_at_prepare();  // the plan a new verification batch starts with
// Up to this many requests per layer, a signing batch's WOTS signatures come
// from the latency kernels (address-only work of the needed steps wide, then
// one compression per step) instead of one lane per whole chain (k_wots).
constexpr int kLatWots = 256;
// Up to this many XMSS^MT requests per verify_many call, the upper layers'
// chains are hashed speculatively for every start digit (16x their work)
// beside layer 0, taking the chains off each upper layer's serial path.
constexpr int kSpecVerify = 64;
// Up to this many leaves per signing batch, its subtrees' leaves come from
// the latency kernels (the throughput kernels would leave most SIMDs idle).
constexpr uint64_t kLatLeaves = 16384;
// Dynamic LDS of k_bds_update: the serialized states plus, when they fit in
// this much, the values of the planned merges.
constexpr size_t kBdsLds = 64 * 1024;
// BDS signatures: messages up to this long are hashed before the leaves.
constexpr uint64_t kBdsShortMsg = 512;
// Polls (about 0.16 us apart: about 0.65 ms in all, several times the
// longest leaf) the merge-DAG block of a one-call BDS signature's
// finishing launch makes for the planned leaves before it leaves the DAG to
// the update launch (BdsFinDag); XMSS_MI355X_DAG_SPIN overrides it, and 0
// keeps the DAG in the update.
static uint32_t dag_spin() {
    static const uint32_t v = [] {
        const char* e = getenv("XMSS_MI355X_DAG_SPIN");
        return e ? (uint32_t)strtoul(e, nullptr, 10) : (uint32_t)(1u << 12);
    }();
    return v;
}
static_assert(Engine::DAG_LEAF == DAG_LEAF && Engine::DAG_MERGE == DAG_MERGE && Engine::DAG_INDEX == DAG_INDEX,
              "merge source codes of the host plan and the kernel");
static bool use_lat(FamOps* o, size_t count, uint32_t th) {
    return count > 0 && count <= 65535 && count * o->lat_bytes(th) <= kLatScratch;
}

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    template <class T = uint8_t>
    T* need(size_t bytes) {
        bytes = std::max<size_t>(bytes, 64);
        if (bytes > cap && t_arena) {  // arena: exact sizes, freed after each call
            release();
            cap = (bytes + 255) & ~(size_t)255;
            p = dev_malloc(cap);
            return (T*)p;
        }
        if (bytes > cap) {
            release();
            size_t c = std::max(bytes, cap + cap / 2);
            c = (c + 4095) & ~(size_t)4095;
            try {
                p = dev_malloc(c);
            } catch (const DeviceError&) {  // the 1.5x growth did not fit: try the exact size
                c = (bytes + 4095) & ~(size_t)4095;
                p = dev_malloc(c);
            }
            cap = c;
        }
        return (T*)p;
    }
    void release() {
        void* q = p;
        const size_t c = cap;
        p = nullptr;
        cap = 0;
        dev_free(q, c);
    }
};

// Grow-only pinned host buffer (hipHostMalloc), for single-copy staging.
struct PinnedBuf {
    uint8_t* p = nullptr;
    size_t cap = 0;
    uint8_t* need(size_t bytes) {
        bytes = std::max<size_t>(bytes, 256);
        if (bytes > cap) {
            if (p) {
                HIPCK(hipDeviceSynchronize());  // no copy may still read or write the old buffer
                HIPCK(hipHostFree(p));
            }
            cap = (std::max(bytes, cap + cap / 2) + 4095) & ~(size_t)4095;
            HIPCK(hipHostMalloc((void**)&p, cap, hipHostMallocDefault));
        }
        return p;
    }
};

struct BdsBufs;
struct VerifyBufs;
void destroy_batch(Engine::Batch* b);

// Keyed 128-bit digest of a key's seed bytes: the pool finds a key's slot by
// it, so no seed bytes are kept on the host.  Two SipHash-2-4 values
// (Aumasson and Bernstein, 2012) under independent random per-process keys.
struct Digest {
    uint64_t a = 0, b = 0;
    bool operator==(const Digest& o) const { return a == o.a && b == o.b; }
};
struct DigestHash {
    size_t operator()(const Digest& d) const { return (size_t)(d.a ^ (d.b * 0x9E3779B97F4A7C15ull)); }
};
[[maybe_unused]] static inline uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
static uint64_t siphash24(uint64_t k0, uint64_t k1, const uint8_t* m, size_t len) {
    uint64_t v0 = k0 ^ 0x736f6d6570736575ull, v1 = k1 ^ 0x646f72616e646f6dull;
    uint64_t v2 = k0 ^ 0x6c7967656e657261ull, v3 = k1 ^ 0x7465646279746573ull;
    auto round = [&] {
        v0 += v1; v1 = rotl64(v1, 13); v1 ^= v0; v0 = rotl64(v0, 32);
        v2 += v3; v3 = rotl64(v3, 16); v3 ^= v2;
        v0 += v3; v3 = rotl64(v3, 21); v3 ^= v0;
        v2 += v1; v1 = rotl64(v1, 17); v1 ^= v2; v2 = rotl64(v2, 32);
    };
    size_t i = 0;
    for (; i + 8 <= len; i += 8) {
        uint64_t w = 0;
        for (int b = 7; b >= 0; --b) w = (w << 8) | m[i + b];
        v3 ^= w;
        round();
        round();
        v0 ^= w;
    }
    uint64_t last = (uint64_t)(len & 0xff) << 56;
    for (size_t b = 0; i + b < len; ++b) last |= (uint64_t)m[i + b] << (8 * b);
    v3 ^= last;
    round();
    round();
    v0 ^= last;
    v2 ^= 0xff;
    for (int r = 0; r < 4; ++r) round();
    return v0 ^ v1 ^ v2 ^ v3;
}
static Digest seed_digest(unsigned n, const KeySeeds& s) {
    static const std::array<uint64_t, 4> key = [] {
        std::random_device rd;
        std::array<uint64_t, 4> k;
        for (uint64_t& x : k) x = ((uint64_t)rd() << 32) ^ rd();
        return k;
    }();
    uint8_t b[3 * 64] = {0};
    if (s.sk_seed) memcpy(b, s.sk_seed, n);
    if (s.sk_prf) memcpy(b + n, s.sk_prf, n);
    if (s.pub_seed) memcpy(b + 2 * n, s.pub_seed, n);
    Digest d{siphash24(key[0], key[1], b, 3 * n), siphash24(key[2], key[3], b, 3 * n)};
    volatile uint8_t* vb = b;
    for (unsigned i = 0; i < 3 * n; ++i) vb[i] = 0;
    return d;
}

// Device-resident key contexts of one hash family.  Pinned slots (pool_add:
// the serving loop's keys) live until released; cached slots (pool_slot: keys
// met by one-call signatures and verifications) are evicted least recently
// used beyond kPoolCached, never within the call that uses them (pool_begin
// starts a call).  A freed slot's context is overwritten with zeros.
constexpr size_t kPoolCached = 1024;
struct KeyPool {
    void* p = nullptr;
    size_t cap = 0;  // slots allocated
    struct Slot {
        Digest d;
        uint64_t used = 0;  // epoch of the last use
        bool live = false, pinned = false;
    };
    std::vector<Slot> slot;
    std::vector<int> free_slots;
    std::unordered_map<Digest, int, DigestHash> by_digest;
    size_t cached = 0;
};

// The allocation-free verification's working memory (verify_static): one
// file-scope device array per GPU, holding every buffer of one call -- staged
// inputs, key context, pre-images, digests, roots.  Its address comes from
// hipGetSymbolAddress once per device (Engine::get); host code uses only that
// address, never the array's name.
constexpr size_t kStaticScratch = 16u << 20;
namespace noalloc {
static __device__ __align__(256) uint8_t scratch[kStaticScratch];
}

// Chain ends hashed one signature ahead (Engine::ends_ahead): the chains of
// the leaves a key's next signature will consume are hashed by this
// signature's finishing launch, beside its WOTS signature, into slots of
// `buf`, and the next signature hashes only their L-trees.  `have`: the
// leaves of key d (family fam) whose chain ends a finished launch left in a
// slot; `want`: the request for the coming launch, for key want_d.
struct EndsAhead {
    using Id = std::tuple<uint32_t, uint64_t, uint32_t>;  // layer, tree, leaf
    DevBuf buf;
    void* at = nullptr;  // buf.p when `have` was filled
    int fam = -1, want_fam = -1;
    Digest d, want_d;
    std::map<Id, uint32_t> have;
    std::vector<Id> want;
    void clear() {
        have.clear();
        want.clear();
        fam = want_fam = -1;
    }
};
constexpr uint32_t kEndsSlots = 64;

struct Engine::Impl {
    Launcher L;
    EndsAhead ahead;
    void* static_scratch = nullptr;  // this device's noalloc::scratch
    KeyPool pool[F_COUNT];
    const void* kp_override = nullptr;  // pool in use, or null for `keys`
    BdsBufs* bds = nullptr;             // buffers of bds_sign_batch
    VerifyBufs* ver = nullptr;          // buffers of verify_many
    const void* kp() const { return kp_override ? kp_override : keys.p; }
    Engine::Batch* spare[2] = {nullptr, nullptr};  // reusable verify / sign batches
    DevBuf keys, seeds, desc, ends, tmp, nodes, a, b, c, d, idx, lat, tpre;
    // input copies of verification batches: a batch's inputs go in here while
    // other batches' kernels run on s_main / s_aux
    hipStream_t s_copy = nullptr;
    uint64_t epoch = 1;  // pool_begin: slots used since then are not evicted
    bool external = false;  // kp_override is a caller's key table (use_keys)
    int key_fam = -1;
    int nkeys = 0;
    void sync() {
        HIPCK(hipStreamSynchronize(L.s));
        L.harvest();
    }
    void h2d(void* dst, const void* src, size_t n) {
        if (n) HIPCK(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, L.s));
    }
    void d2h(void* dst, const void* src, size_t n) {
        if (n) HIPCK(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, L.s));
    }
    FamOps* ops(int fam) {
        if (fam != key_fam) fatal("engine call for a hash family other than the loaded keys");
        return ops_of(fam);
    }
};

Engine::Engine() : impl_(new Impl) {
    HIPCK(hipGetDevice(&dev_));
    HIPCK(hipStreamCreateWithFlags(&impl_->L.s_main, hipStreamNonBlocking));
    HIPCK(hipStreamCreateWithFlags(&impl_->L.s_aux, hipStreamNonBlocking));
    HIPCK(hipStreamCreateWithFlags(&impl_->s_copy, hipStreamNonBlocking));
    impl_->L.s = impl_->L.s_main;
    stream_ = impl_->L.s;
}

Engine* Engine::get() {
    static std::mutex mu;
    static std::map<int, Engine*> engines;
    std::lock_guard<std::mutex> lk(mu);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        fprintf(stderr, "xmss_mi355x: no HIP device available (this library has no CPU path)\n");
        fflush(stderr);
        return nullptr;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    auto it = engines.find(dev);
    if (it != engines.end()) return it->second;
    Engine* e = nullptr;
    try {
        e = new Engine();
        // set-up, once per device with the device current: the address of
        // this device's copy of the allocation-free path's scratch array
        void* ptr = nullptr;
        HIPCK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(noalloc::scratch)));
        e->impl_->static_scratch = ptr;
    } catch (const DeviceError&) {
        return nullptr;
    }
    engines[dev] = e;
    return e;
}

// Private engines are kept for reuse rather than destroyed: their streams
// and device buffers live as long as the process, as the shared engine's do.
static std::mutex& private_mutex() {
    static std::mutex m;
    return m;
}
static std::vector<Engine*>& private_free() {
    static std::vector<Engine*> v;
    return v;
}

Engine* Engine::make_private() {
    {
        std::lock_guard<std::mutex> lk(private_mutex());
        if (!private_free().empty()) {
            Engine* e = private_free().back();
            private_free().pop_back();
            return e;
        }
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return nullptr;
    try {
        return new Engine();
    } catch (const DeviceError&) {
        return nullptr;
    }
}

void Engine::release_private(Engine* e) {
    if (!e) return;
    std::lock_guard<std::mutex> lk(private_mutex());
    private_free().push_back(e);
}

void* Engine::device_alloc(size_t bytes) {
    try {
        return dev_malloc(bytes);
    } catch (const DeviceError&) {
        return nullptr;
    }
}

void Engine::device_free(void* p, size_t bytes) { dev_free(p, bytes); }

void Engine::upload(void* dst, const void* src, size_t bytes) {
    if (bytes) HIPCK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}

__global__ __launch_bounds__(WG) void k_fill32(uint32_t* __restrict__ p, uint64_t words, uint32_t v) {
    const uint64_t g = (uint64_t)blockIdx.x * WG + threadIdx.x;
    if (g < words) p[g] = v;
}

// Zeros `bytes` (a multiple of 4) at p on the launcher's stream: a kernel
// rather than hipMemsetAsync, so that a caller may capture it in a graph.
static void zero_async(Launcher& L, void* p, size_t bytes) {
    const uint64_t words = bytes / 4;
    for (uint64_t w0 = 0; w0 < words; w0 += (uint64_t)WG * 65535) {
        const uint64_t w = std::min<uint64_t>(words - w0, (uint64_t)WG * 65535);
        L.run("fill", k_fill32, dim3(cdiv(w, WG)), dim3(WG), (uint32_t*)p + w0, w, 0u);
    }
}

void Engine::recover() {
    // after an exception: the launcher back on the main stream, pending
    // timings dropped, the streams drained (errors ignored: the next call
    // reports a device that stays broken)
    Impl* I = impl_;
    I->L.s = I->L.s_main;
    for (auto& p : I->L.pending) {
        I->L.pool.push_back(p.e0);
        I->L.pool.push_back(p.e1);
    }
    I->L.pending.clear();
    (void)hipStreamSynchronize(I->L.s_main);
    (void)hipStreamSynchronize(I->L.s_aux);
    (void)hipStreamSynchronize(I->s_copy);
    (void)hipGetLastError();
    // chain ends hashed ahead are not trusted after a failure
    I->ahead.clear();
}

void Engine::scrub_keys() {
    Impl* I = impl_;
    if (I->keys.p) zero_async(I->L, I->keys.p, I->keys.cap);
    if (I->seeds.p) zero_async(I->L, I->seeds.p, I->seeds.cap);
}

void Engine::set_profiling(bool on) {
    prof_on_ = on;
    impl_->L.prof = on;
}
std::vector<Engine::KStat> Engine::stats() const {
    std::vector<KStat> v;
    for (auto& kv : impl_->L.stat) v.push_back({kv.first, kv.second.first, kv.second.second});
    return v;
}
void Engine::reset_stats() { impl_->L.stat.clear(); }

void Engine::set_keys(int fam, int nkeys, const KeySeeds* seeds) {
    Impl* I = impl_;
    const unsigned n = fam_n(fam);
    std::vector<uint8_t> h((size_t)nkeys * 3 * n, 0);
    for (int i = 0; i < nkeys; ++i) {
        uint8_t* d = &h[(size_t)i * 3 * n];
        if (seeds[i].sk_seed) memcpy(d, seeds[i].sk_seed, n);
        if (seeds[i].sk_prf) memcpy(d + n, seeds[i].sk_prf, n);
        if (seeds[i].pub_seed) memcpy(d + 2 * n, seeds[i].pub_seed, n);
    }
    uint8_t* ds = I->seeds.need(h.size());
    I->h2d(ds, h.data(), h.size());
    FamOps* o = ops_of(fam);
    void* dk = I->keys.need(o->key_bytes() * (size_t)nkeys);
    o->prepare(I->L, ds, nkeys, dk);
    I->key_fam = fam;
    I->nkeys = nkeys;
    I->kp_override = nullptr;
    I->external = false;
    HIPCK(hipStreamSynchronize(I->L.s));  // the copy has read h: clear it
    volatile uint8_t* vh = h.data();
    for (size_t i = 0; i < h.size(); ++i) vh[i] = 0;
}

void Engine::ends_ahead(int fam, const KeySeeds& s, const LeafRun* ids, int count) {
    EndsAhead& E = impl_->ahead;
    E.want.clear();
    E.want_fam = fam;
    E.want_d = seed_digest(fam_n(fam), s);
    for (int i = 0; i < count; ++i) E.want.emplace_back(ids[i].layer, ids[i].tree, ids[i].leaf0);
}

void Engine::ends_ahead_drop() { impl_->ahead.clear(); }

// A free slot of the pool (grown when none is left), prepared for `s`.
static int pool_take(Engine::Impl* I, int fam, const KeySeeds& s, bool pinned, const Digest& d) {
    FamOps* o = ops_of(fam);
    KeyPool& P = I->pool[fam];
    const size_t kb = o->key_bytes();
    int slot;
    if (!P.free_slots.empty()) {
        slot = P.free_slots.back();
        P.free_slots.pop_back();
    } else {
        if (P.slot.size() == P.cap) {  // grow, keeping the prepared contexts
            const size_t cap = std::max<size_t>(64, P.cap * 2);
            void* np = dev_malloc(cap * kb);
            if (P.p) {
                HIPCK(hipMemcpyAsync(np, P.p, P.slot.size() * kb, hipMemcpyDeviceToDevice, I->L.s));
                HIPCK(hipStreamSynchronize(I->L.s));
                dev_free(P.p, P.cap * kb);
            }
            P.p = np;
            P.cap = cap;
            if (I->key_fam == fam && I->kp_override) I->kp_override = P.p;
        }
        slot = (int)P.slot.size();
        P.slot.emplace_back();
    }
    const unsigned n = fam_n(fam);
    uint8_t h[3 * 64] = {0};
    if (s.sk_seed) memcpy(&h[0], s.sk_seed, n);
    if (s.sk_prf) memcpy(&h[n], s.sk_prf, n);
    if (s.pub_seed) memcpy(&h[2 * n], s.pub_seed, n);
    uint8_t* ds = I->seeds.need(3 * n);
    I->h2d(ds, h, 3 * n);
    o->prepare(I->L, ds, 1, (uint8_t*)P.p + (size_t)slot * kb);
    zero_async(I->L, ds, ((3 * n + 3) / 4) * 4);
    I->sync();
    volatile uint8_t* vh = h;
    for (unsigned i = 0; i < 3 * n; ++i) vh[i] = 0;
    KeyPool::Slot& S = P.slot[slot];
    S.d = d;
    S.used = I->epoch;
    S.live = true;
    S.pinned = pinned;
    if (!pinned) ++P.cached;
    return slot;
}

// Frees a slot: its context is zeroed and its digest forgotten.
static void pool_drop(Engine::Impl* I, int fam, int slot) {
    KeyPool& P = I->pool[fam];
    KeyPool::Slot& S = P.slot[slot];
    if (!S.live) return;
    auto it = P.by_digest.find(S.d);
    if (it != P.by_digest.end() && it->second == slot) P.by_digest.erase(it);
    if (!S.pinned) --P.cached;
    const size_t kb = ops_of(fam)->key_bytes();
    zero_async(I->L, (uint8_t*)P.p + (size_t)slot * kb, kb);
    S = KeyPool::Slot{};
    P.free_slots.push_back(slot);
}

int Engine::pool_add(int fam, const KeySeeds& s) {
    Impl* I = impl_;
    const Digest d = seed_digest(fam_n(fam), s);
    const int slot = pool_take(I, fam, s, true, d);
    I->pool[fam].by_digest[d] = slot;
    return slot;
}

void Engine::pool_release(int fam, int slot) {
    Impl* I = impl_;
    if (slot < 0 || (size_t)slot >= I->pool[fam].slot.size()) return;
    pool_drop(I, fam, slot);
    I->sync();
}

void Engine::pool_forget(int fam, const KeySeeds& s) {
    Impl* I = impl_;
    KeyPool& P = I->pool[fam];
    auto it = P.by_digest.find(seed_digest(fam_n(fam), s));
    if (it == P.by_digest.end() || P.slot[it->second].pinned) return;
    pool_drop(I, fam, it->second);
    I->sync();
}

void Engine::pool_begin() { ++impl_->epoch; }

size_t Engine::pool_cached(int fam) const { return impl_->pool[fam].cached; }

void Engine::pool_use(int fam) {
    impl_->key_fam = fam;
    impl_->kp_override = impl_->pool[fam].p;
    impl_->external = false;
}

// The keys of a launch sequence of `fam`: the caller's table when one was
// given for this family (use_keys), else the engine's pool.
static void keys_for(Engine* e, Engine::Impl* I, int fam) {
    if (!(I->external && I->key_fam == fam)) e->pool_use(fam);
}

size_t Engine::pool_key_bytes(int fam) const { return ops_of(fam)->key_bytes(); }
size_t Engine::pool_key_bytes_of(int fam) { return ops_of(fam)->key_bytes(); }

void Engine::copy_device(void* dst, const void* src, size_t bytes) {
    if (bytes) HIPCK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
}

void Engine::clear_device(void* p, size_t bytes) {
    if (p && bytes) {
        HIPCK(hipMemset(p, 0, bytes));
        HIPCK(hipDeviceSynchronize());
    }
}

void Engine::use_keys(int fam, const void* table) {
    impl_->key_fam = fam;
    impl_->kp_override = table;
    impl_->external = true;
}

void Engine::prepare_keys(int fam, const KeySeeds* seeds, int nkeys, void* table, int first) {
    Impl* I = impl_;
    FamOps* o = ops_of(fam);
    const unsigned n = fam_n(fam);
    std::vector<uint8_t> h((size_t)nkeys * 3 * n, 0);
    for (int i = 0; i < nkeys; ++i) {
        uint8_t* d = &h[(size_t)i * 3 * n];
        if (seeds[i].sk_seed) memcpy(d, seeds[i].sk_seed, n);
        if (seeds[i].sk_prf) memcpy(d + n, seeds[i].sk_prf, n);
        if (seeds[i].pub_seed) memcpy(d + 2 * n, seeds[i].pub_seed, n);
    }
    uint8_t* ds = I->seeds.need(h.size());
    I->h2d(ds, h.data(), h.size());
    o->prepare(I->L, ds, nkeys, (uint8_t*)table + (size_t)first * o->key_bytes());
    zero_async(I->L, ds, (h.size() + 3) & ~(size_t)3);
    I->sync();
    volatile uint8_t* vh = h.data();
    for (size_t i = 0; i < h.size(); ++i) vh[i] = 0;
}

static std::vector<DRun> pack_runs(int n, const LeafRun* runs) {
    std::vector<DRun> v(n);
    for (int i = 0; i < n; ++i) {
        v[i] = {runs[i].key, runs[i].layer, (uint32_t)(runs[i].tree >> 32), (uint32_t)runs[i].tree,
                runs[i].leaf0, 0, 0, 0};
    }
    return v;
}

// Byte budget of the chain-end staging buffer per launch group.
static const size_t kEndsBudget = (size_t)1 << 30;

void Engine::build_trees(int fam, unsigned th, int nruns, const LeafRun* runs) {
    Impl* I = impl_;
    FamOps* o = I->ops(fam);
    const unsigned n = fam_n(fam), len = fam_len(fam);
    const uint64_t stride = (2ull << th) - 1;
    const uint64_t leaves = 1ull << th;
    const uint64_t per_leaf = (uint64_t)len * n;
    std::vector<DRun> d = pack_runs(nruns, runs);
    DRun* drun = I->desc.need<DRun>(sizeof(DRun) * d.size());
    I->h2d(drun, d.data(), sizeof(DRun) * d.size());
    uint8_t* nodes = I->nodes.need(stride * n * (size_t)nruns);
    const void* dk = I->kp();
    // few trees: their levels are latency-bound, so the address-only part of
    // every internal node runs on the aux stream beside the leaves
    const bool split = nruns <= kSplitRuns;
    void* tpre = split ? I->tpre.need((size_t)nruns * ((1ull << th) - 1) * o->hpre_bytes()) : nullptr;
    hipEvent_t staged = nullptr, pre = nullptr;
    if (split) {
        staged = I->L.ev();
        pre = I->L.ev();
        HIPCK(hipEventRecord(staged, I->L.s_main));
        HIPCK(hipStreamWaitEvent(I->L.s_aux, staged, 0));
        I->L.s = I->L.s_aux;
        o->pre_tree(I->L, dk, drun, nruns, th, tpre);
        HIPCK(hipEventRecord(pre, I->L.s_aux));
        I->L.s = I->L.s_main;
    }
    for (int r0 = 0; r0 < nruns;) {
        int rg = std::min<int>(nruns - r0, 65535);
        uint64_t lpp = leaves;
        while ((uint64_t)rg * lpp * per_leaf > kEndsBudget && lpp > 64) lpp >>= 1;
        while ((uint64_t)rg * lpp * per_leaf > kEndsBudget && rg > 1) rg = (rg + 1) / 2;
        uint8_t* ends = I->ends.need((size_t)rg * lpp * per_leaf);
        uint8_t* tmp = I->tmp.need((size_t)rg * lpp * ((len + 1) / 2) * n);
        uint8_t* tnodes = nodes + (size_t)r0 * stride * n;
        for (uint64_t l0 = 0; l0 < leaves; l0 += lpp) {
            const uint32_t cnt = (uint32_t)std::min<uint64_t>(lpp, leaves - l0);
            o->chains(I->L, dk, drun + r0, rg, (uint32_t)l0, cnt, ends);
            // L-tree height by height, ping-ponging ends <-> tmp
            uint32_t l = len, height = 0;
            bool from_a = true;
            const uint32_t sa = len, sb = (len + 1) / 2;
            while (l > 1) {
                o->ltree_level(I->L, dk, drun + r0, rg, (uint32_t)l0, cnt, l, height, from_a ? ends : tmp,
                               from_a ? sa : sb, from_a ? tmp : ends, from_a ? sb : sa, tnodes, stride, l0);
                from_a = !from_a;
                l = (l + 1) / 2;
                ++height;
            }
        }
        if (split) {
            HIPCK(hipStreamWaitEvent(I->L.s_main, pre, 0));
            o->levels_fin(I->L, rg, th, tnodes, (const uint8_t*)tpre + (size_t)r0 * ((1ull << th) - 1) * o->hpre_bytes());
        } else {
            for (unsigned lev = 0; lev < th; ++lev) o->level(I->L, dk, drun + r0, rg, th, lev, tnodes);
        }
        r0 += rg;
    }
    I->sync();
    if (split) {
        I->L.pool.push_back(staged);
        I->L.pool.push_back(pre);
    }
}

void Engine::fetch_nodes(int fam, unsigned th, int nrefs, const NodeRef* refs, uint8_t* out) {
    Impl* I = impl_;
    const unsigned n = fam_n(fam);
    const uint64_t stride = (2ull << th) - 1;
    std::vector<uint64_t> idx(nrefs);
    for (int i = 0; i < nrefs; ++i) {
        idx[i] = refs[i].tree * stride + ((2ull << th) - (2ull << (th - refs[i].level))) + refs[i].index;
    }
    uint64_t* didx = I->idx.need<uint64_t>(sizeof(uint64_t) * (size_t)nrefs);
    uint8_t* dout = I->b.need((size_t)n * nrefs);
    I->h2d(didx, idx.data(), sizeof(uint64_t) * nrefs);
    I->L.run("gather", k_gather, dim3(cdiv((uint64_t)nrefs * (n / 4), WG)), dim3(WG),
             (const uint8_t*)I->nodes.p, (const uint64_t*)didx, (uint32_t)nrefs, n, dout);
    I->d2h(out, dout, (size_t)n * nrefs);
    I->sync();
}

void Engine::leaves(int fam, int nruns, const LeafRun* runs, unsigned count, uint8_t* out) {
    Impl* I = impl_;
    FamOps* o = I->ops(fam);
    const unsigned n = fam_n(fam);
    const size_t nl = (size_t)nruns * count;
    std::vector<DLeaf> d(nl);
    for (int i = 0; i < nruns; ++i) {
        for (unsigned j = 0; j < count; ++j) {
            d[(size_t)i * count + j] = {runs[i].key, runs[i].layer, (uint32_t)(runs[i].tree >> 32),
                                        (uint32_t)runs[i].tree, runs[i].leaf0 + j, 0, 0, 0};
        }
    }
    DLeaf* dd = I->desc.need<DLeaf>(sizeof(DLeaf) * nl);
    uint8_t* dout = I->b.need(nl * n);
    I->h2d(dd, d.data(), sizeof(DLeaf) * nl);
    if (use_lat(o, nl, 0))
        o->lat(I->L, I->kp(), LAT_LEAF, dd, (uint32_t)nl, nullptr, nullptr, nullptr, 0,
               I->lat.need(nl * o->lat_bytes(0)), dout);
    else
        o->leaf_wg(I->L, I->kp(), dd, (uint32_t)nl, false, nullptr, nullptr, nullptr, 0, dout);
    I->d2h(out, dout, nl * n);
    I->sync();
}

void Engine::verify_layer(int fam, unsigned th, int nreq, const RootReq* reqs, const uint8_t* msgs,
                          const uint8_t* sigs, const uint8_t* auth, uint8_t* roots) {
    Impl* I = impl_;
    FamOps* o = I->ops(fam);
    const unsigned n = fam_n(fam), len = fam_len(fam);
    std::vector<DLeaf> d(nreq);
    for (int i = 0; i < nreq; ++i) {
        d[i] = {reqs[i].key, reqs[i].layer, (uint32_t)(reqs[i].tree >> 32), (uint32_t)reqs[i].tree,
                reqs[i].leaf_idx, th, 0, (uint32_t)(i * th)};
    }
    DLeaf* dd = I->desc.need<DLeaf>(sizeof(DLeaf) * nreq);
    uint8_t* dm = I->a.need((size_t)nreq * n);
    uint8_t* ds = I->c.need((size_t)nreq * len * n);
    uint8_t* da = I->d.need((size_t)nreq * th * n);
    uint8_t* dout = I->b.need((size_t)nreq * n);
    I->h2d(dd, d.data(), sizeof(DLeaf) * nreq);
    I->h2d(dm, msgs, (size_t)nreq * n);
    I->h2d(ds, sigs, (size_t)nreq * len * n);
    I->h2d(da, auth, (size_t)nreq * th * n);
    if (use_lat(o, nreq, th))
        o->lat(I->L, I->kp(), LAT_VERIFY, dd, (uint32_t)nreq, dm, ds, da, th,
               I->lat.need((size_t)nreq * o->lat_bytes(th)), dout);
    else
        o->leaf_wg(I->L, I->kp(), dd, (uint32_t)nreq, true, dm, ds, da, th, dout);
    I->d2h(roots, dout, (size_t)nreq * n);
    I->sync();
}

void Engine::wots(int fam, int nreq, const WotsReq* reqs, const uint8_t* msgs, const uint8_t* in,
                  uint8_t* out) {
    Impl* I = impl_;
    FamOps* o = I->ops(fam);
    const unsigned n = fam_n(fam), len = fam_len(fam);
    std::vector<DWots> d(nreq);
    for (int i = 0; i < nreq; ++i) {
        d[i] = {reqs[i].key, reqs[i].layer, (uint32_t)(reqs[i].tree >> 32), (uint32_t)reqs[i].tree,
                reqs[i].ots, reqs[i].mode, 0, 0};
    }
    DWots* dd = I->desc.need<DWots>(sizeof(DWots) * d.size());
    uint8_t* dm = I->a.need((size_t)nreq * n);
    uint8_t* din = I->c.need((size_t)nreq * len * n);
    uint8_t* dout = I->b.need((size_t)nreq * len * n);
    bool all_sign = use_lat(o, nreq, 0);
    for (int i = 0; i < nreq && all_sign; ++i) all_sign = reqs[i].mode == MODE_SIGN;
    if (all_sign) {
        std::vector<DLeaf> dl(nreq);
        for (int i = 0; i < nreq; ++i) dl[i] = {d[i].key, d[i].layer, d[i].tree_hi, d[i].tree_lo, d[i].ots, 0, 0, 0};
        DLeaf* ddl = I->desc.need<DLeaf>(sizeof(DLeaf) * nreq);
        I->h2d(ddl, dl.data(), sizeof(DLeaf) * nreq);
        I->h2d(dm, msgs, (size_t)nreq * n);
        o->lat(I->L, I->kp(), LAT_SIGN, ddl, (uint32_t)nreq, dm, nullptr, nullptr, 0,
               I->lat.need((size_t)nreq * o->lat_bytes(0)), dout);
        I->d2h(out, dout, (size_t)nreq * len * n);
        I->sync();
        return;
    }
    I->h2d(dd, d.data(), sizeof(DWots) * d.size());
    I->h2d(dm, msgs, (size_t)nreq * n);
    if (in) I->h2d(din, in, (size_t)nreq * len * n);
    o->wots(I->L, I->kp(), dd, (uint32_t)nreq, dm, din, dout);
    I->d2h(out, dout, (size_t)nreq * len * n);
    I->sync();
}

void Engine::thash_h(int fam, int nreq, const ThashReq* reqs, const uint8_t* in, uint8_t* out) {
    Impl* I = impl_;
    FamOps* o = I->ops(fam);
    const unsigned n = fam_n(fam);
    std::vector<DThash> d(nreq);
    for (int i = 0; i < nreq; ++i) {
        memset(&d[i], 0, sizeof(DThash));
        d[i].key = reqs[i].key;
        memcpy(d[i].adrs, reqs[i].adrs, 32);
    }
    DThash* dd = I->desc.need<DThash>(sizeof(DThash) * d.size());
    uint8_t* di = I->a.need((size_t)nreq * 2 * n);
    uint8_t* dout = I->b.need((size_t)nreq * n);
    I->h2d(dd, d.data(), sizeof(DThash) * d.size());
    I->h2d(di, in, (size_t)nreq * 2 * n);
    o->thash_h(I->L, I->kp(), dd, (uint32_t)nreq, di, dout);
    I->d2h(out, dout, (size_t)nreq * n);
    I->sync();
}

void Engine::hmsg(int fam, int nreq, const MsgReq* reqs, uint8_t* R_out, uint8_t* dig_out) {
    Impl* I = impl_;
    FamOps* o = I->ops(fam);
    const unsigned n = fam_n(fam), pad = fam_pad(fam);
    const size_t pb = pad + 3 * n;
    std::vector<DMsg> d(nreq);
    size_t total = 0;
    for (int i = 0; i < nreq; ++i) {
        d[i] = {reqs[i].key, reqs[i].sign, reqs[i].idx, (uint64_t)total, (uint32_t)(pb + reqs[i].mlen), 0};
        total += (pb + reqs[i].mlen + 15) & ~(size_t)15;
    }
    total += kMsgSlack;
    // toByte(2, pad) || R || root || toByte(idx, n) || M  (hash.c:107-120)
    std::vector<uint8_t> blob(total, 0);
    for (int i = 0; i < nreq; ++i) {
        uint8_t* p = &blob[d[i].off];
        p[pad - 1] = 2;
        if (!reqs[i].sign && reqs[i].R) memcpy(p + pad, reqs[i].R, n);
        memcpy(p + pad + n, reqs[i].root, n);
        uint64_t v = reqs[i].idx;
        for (int b = 0; b < 8; ++b) {
            p[pad + 3 * n - 1 - b] = (uint8_t)(v & 0xff);
            v >>= 8;
        }
        if (reqs[i].mlen) memcpy(p + pb, reqs[i].m, reqs[i].mlen);
    }
    DMsg* dd = I->desc.need<DMsg>(sizeof(DMsg) * d.size());
    uint8_t* db = I->c.need(total);
    uint8_t* dR = I->a.need((size_t)nreq * n);
    uint8_t* dD = I->b.need((size_t)nreq * n);
    I->h2d(dd, d.data(), sizeof(DMsg) * d.size());
    I->h2d(db, blob.data(), total);
    o->hmsg(I->L, I->kp(), dd, (uint32_t)nreq, pad, db, dR, dD);
    if (R_out) I->d2h(R_out, dR, (size_t)nreq * n);
    I->d2h(dig_out, dD, (size_t)nreq * n);
    I->sync();
}

__global__ __launch_bounds__(WG) void k_gather(const uint8_t* __restrict__ src,
                                               const uint64_t* __restrict__ idx, uint32_t n,
                                               uint32_t nbytes, uint8_t* __restrict__ dst) {
    const uint32_t words = nbytes / 4;
    const uint32_t g = blockIdx.x * WG + threadIdx.x;
    if (g >= n * words) return;
    const uint32_t item = g / words, w = g % words;
    reinterpret_cast<uint32_t*>(dst)[(size_t)item * words + w] =
        reinterpret_cast<const uint32_t*>(src)[idx[item] * words + w];
}

}  // namespace xm

namespace xm {

// Copies node i of a flat node array into out: used to lay out auth paths
// and roots of built subtrees (see Batch::gather_idx).
struct Engine::Batch {
    bool sign = false;
    int fam = 0;
    unsigned th = 0, d = 0, n = 0, len = 0;
    int count = 0;
    int nruns = 0;  // subtrees to build (count x d, or fewer when items share trees)
    // inputs, staged in one copy into `io`; results in `res`, copied back in one
    DevBuf io, res, keys, dig, nodes, ends, tmp, roots;
    size_t nodes_stride = 0;
    const uint8_t* seeds = nullptr;
    const DMsg* md = nullptr;
    uint8_t* blob = nullptr;
    const DRun* runs = nullptr;
    const DWots* wd = nullptr;
    const uint64_t* gidx = nullptr;
    const DLeaf* vd = nullptr;
    const uint8_t* sigs = nullptr;     // verify: WOTS signatures, layer-major
    const uint8_t* auth_in = nullptr;  // verify: auth paths, layer-major
    const uint8_t* pkroot = nullptr;
    uint8_t* R = nullptr;     // sign: R | auth records | WOTS signatures, contiguous in res
    uint8_t* auth = nullptr;
    uint8_t* wots = nullptr;
    size_t res_bytes = 0;
    int* ok = nullptr;        // verify
    std::vector<int> perm;    // verify: staged position -> the caller's index (longest message first)
    DevBuf lat;               // pre-image scratch of the latency kernels
    DevBuf lreq;              // sign: request descriptors of the latency kernels
    DevBuf wlat, wreq;        // sign: the same for the WOTS signatures (layer-major)
    hipGraphExec_t graph = nullptr;  // the captured launches of run() (verify; graphed sign)
    int nrun = 0;                    // verify: runs since staging
    PinnedBuf pin;                   // host staging of this batch's inputs and results
    size_t seed_off = 0, seed_bytes = 0;  // sign: the seeds' region in `pin` and at `seeds`
    bool one_stream = false;         // verify: all launches on the main stream (set_one_stream)
    bool graphed = false;            // sign: run() captures its launches into `graph`
    int sel = 0;                     // verify: the engine stream its launches go to (set_stream)
    uint32_t plan = 0;               // verify: the schedule of verify_dense (VPLAN_* bits)
    bool tuned = false;              // verify: plan settled (forced, cached or tuned)
    hipEvent_t done = nullptr;       // recorded after the work of launch()
    bool inflight = false;           // launched and not yet waited for
    bool evented = false;            // launched through launch(): results leave on the copy stream
    hipEvent_t staged = nullptr;     // verify: recorded after the staging copy on the copy stream
    bool copying = false;            // verify: that copy may still be reading `pin`
    void drop_graph() {
        if (graph) HIPCK(hipGraphExecDestroy(graph));
        graph = nullptr;
        nrun = 0;
    }
};

// Descriptors of the latency kernels from those of the throughput path:
// every leaf of nruns subtrees of height th (run-major); WOTS signatures.
__global__ __launch_bounds__(WG) void k_runs_to_leaves(const DRun* __restrict__ runs, uint32_t nruns, uint32_t th,
                                                       DLeaf* __restrict__ out) {
    const uint64_t g = (uint64_t)blockIdx.x * WG + threadIdx.x;
    if (g >= ((uint64_t)nruns << th)) return;
    const DRun q = runs[g >> th];
    out[g] = {q.key, q.layer, q.tree_hi, q.tree_lo, q.leaf0 + (uint32_t)(g & ((1ull << th) - 1)), 0, 0, 0};
}
__global__ __launch_bounds__(WG) void k_wots_to_leaves(const DWots* __restrict__ w, uint32_t n,
                                                       DLeaf* __restrict__ out) {
    const uint32_t g = blockIdx.x * WG + threadIdx.x;
    if (g >= n) return;
    const DWots q = w[g];
    out[g] = {q.key, q.layer, q.tree_hi, q.tree_lo, q.ots, 0, 0, 0};
}

__global__ __launch_bounds__(WG) void k_compare(const uint8_t* __restrict__ a, const uint8_t* __restrict__ b,
                                                uint32_t n, uint32_t nbytes, int* __restrict__ ok) {
    const uint32_t i = blockIdx.x * WG + threadIdx.x;
    if (i >= n) return;
    int eq = 1;
    for (uint32_t j = 0; j < nbytes; ++j) eq &= (a[(size_t)i * nbytes + j] == b[(size_t)i * nbytes + j]);
    ok[i] = eq;
}

// Offsets of the parts of one staging area (256-byte aligned).
struct Layout {
    size_t size = 0;
    size_t add(size_t bytes) {
        const size_t o = size;
        size += (bytes + 255) & ~(size_t)255;
        return o;
    }
};

// H_msg inputs: descriptors and the prefixed messages (toByte(2) || R ||
// root || toByte(idx) || M, R filled in on the device for signatures).
struct MsgStage {
    std::vector<DMsg> d;
    size_t total = 0;
};
static MsgStage plan_msgs(int fam, int count, bool sign, const std::vector<MsgReq>& reqs) {
    const unsigned n = fam_n(fam), pad = fam_pad(fam);
    const size_t pb = pad + 3 * n;
    MsgStage m;
    m.d.resize(count);
    for (int i = 0; i < count; ++i) {
        m.d[i] = {reqs[i].key, sign ? 1u : 0u, reqs[i].idx, (uint64_t)m.total, (uint32_t)(pb + reqs[i].mlen), 0};
        m.total += (pb + reqs[i].mlen + 15) & ~(size_t)15;
    }
    m.total += kMsgSlack;
    return m;
}
static void fill_msgs(uint8_t* h, int fam, bool sign, const std::vector<MsgReq>& reqs, const MsgStage& m) {
    const unsigned n = fam_n(fam), pad = fam_pad(fam);
    const size_t pb = pad + 3 * n;
    const size_t cnt = m.d.size();
    // each message's region (up to the next one, or to the end with the
    // slack) is written whole: prefix, message, zeros
    parallel_for(cnt, 512, [&](size_t i) {
        uint8_t* p = h + m.d[i].off;
        const size_t end = i + 1 < cnt ? m.d[i + 1].off : m.total;
        memset(p, 0, pad);
        p[pad - 1] = 2;
        if (sign) memset(p + pad, 0, n);
        else memcpy(p + pad, reqs[i].R, n);
        memcpy(p + pad + n, reqs[i].root, n);
        uint64_t v = reqs[i].idx;
        memset(p + pad + 2 * n, 0, n);
        for (int k = 0; k < 8; ++k) {
            p[pad + 3 * n - 1 - k] = (uint8_t)(v & 0xff);
            v >>= 8;
        }
        if (reqs[i].mlen) memcpy(p + pb, reqs[i].m, reqs[i].mlen);
        memset(p + pb + reqs[i].mlen, 0, end - m.d[i].off - pb - reqs[i].mlen);
    });
    if (!cnt) memset(h, 0, m.total);
}
static void fill_seeds(uint8_t* h, unsigned n, const std::vector<KeySeeds>& ks) {
    memset(h, 0, ks.size() * 3 * n);
    for (size_t i = 0; i < ks.size(); ++i) {
        uint8_t* d = h + i * 3 * n;
        if (ks[i].sk_seed) memcpy(d, ks[i].sk_seed, n);
        if (ks[i].sk_prf) memcpy(d + n, ks[i].sk_prf, n);
        if (ks[i].pub_seed) memcpy(d + 2 * n, ks[i].pub_seed, n);
    }
}

// Blocks until the work of a launched batch is done (no-op otherwise).
static void wait_batch(Engine::Batch* b) {
    if (b->inflight) {
        HIPCK(hipEventSynchronize(b->done));
        b->inflight = false;
    }
    if (b->copying) {  // staged and never run
        HIPCK(hipEventSynchronize(b->staged));
        b->copying = false;
    }
}

static Engine::Batch* take_batch(Engine::Impl* I, bool sign) {
    Engine::Batch* b = I->spare[sign];
    I->spare[sign] = nullptr;
    if (b) b->drop_graph();
    return b ? b : new Engine::Batch;
}

// Frees a batch whose preparation failed (the exception goes on).
struct BatchGuard {
    Engine* e;
    Engine::Batch* b;
    bool ok = false;
    ~BatchGuard() {
        if (ok || !b) return;
        try {
            e->free_batch(b);
        } catch (const DeviceError&) {
        }
    }
};

Engine::Batch* Engine::sign_prepare(int fam, unsigned th, unsigned d, int count, const SignItem* items,
                                    bool one_key, bool wait) {
    Impl* I = impl_;
    Batch* b = take_batch(I, true);
    b->graphed = false;
    BatchGuard guard{this, b};
    b->sign = true;
    b->fam = fam;
    b->th = th;
    b->d = d;
    b->n = fam_n(fam);
    b->len = fam_len(fam);
    b->count = count;
    const unsigned n = b->n;
    FamOps* o = ops_of(fam);
    // one_key: every item signs with the same key (consecutive indices), so
    // one key context serves all and items on the same (layer, tree) share
    // that subtree -- it is built once
    std::vector<KeySeeds> ks(one_key ? 1 : count);
    std::vector<MsgReq> mr(count);
    for (int i = 0; i < count; ++i) {
        const uint32_t key = one_key ? 0u : (uint32_t)i;
        if (!one_key || i == 0) ks[i] = {items[i].sk_seed, items[i].sk_prf, items[i].pub_seed};
        mr[i] = {key, 1, items[i].idx, nullptr, items[i].root, items[i].m, items[i].mlen};
    }
    // one subtree per distinct (key, layer, tree): layer i, tree idx >> (th * (i + 1))
    std::vector<DRun> runs;
    std::map<std::pair<uint32_t, uint64_t>, uint32_t> shared;  // one_key: (layer, tree) -> run
    std::vector<DWots> wd((size_t)count * d);
    std::vector<uint64_t> gi((size_t)count * d * (th + 1));
    b->nodes_stride = (2ull << th) - 1;
    for (int i = 0; i < count; ++i) {
        uint64_t idx = items[i].idx;
        const uint32_t key = one_key ? 0u : (uint32_t)i;
        for (unsigned L = 0; L < d; ++L) {
            const uint32_t leaf = (uint32_t)(idx & ((1ull << th) - 1));
            idx >>= th;
            const size_t t = (size_t)L * count + i;  // layer-major item slots
            uint64_t run;
            auto it = one_key ? shared.find({L, idx}) : shared.end();
            if (it != shared.end()) {
                run = it->second;
            } else {
                run = runs.size();
                runs.push_back({key, L, (uint32_t)(idx >> 32), (uint32_t)idx, 0, 0, 0, 0});
                if (one_key) shared.emplace(std::make_pair(L, idx), (uint32_t)run);
            }
            wd[t] = {key, L, (uint32_t)(idx >> 32), (uint32_t)idx, leaf, MODE_SIGN, 0, 0};
            // auth path (th nodes) then the root, gathered into auth[t]
            for (unsigned j = 0; j < th; ++j) {
                gi[t * (th + 1) + j] =
                    run * b->nodes_stride + ((2ull << th) - (2ull << (th - j))) + ((leaf >> j) ^ 1);
            }
            gi[t * (th + 1) + th] = run * b->nodes_stride + b->nodes_stride - 1;
        }
    }
    b->nruns = (int)runs.size();
    const MsgStage ms = plan_msgs(fam, count, true, mr);
    Layout in;
    const size_t o_seeds = in.add(ks.size() * 3 * n), o_md = in.add(sizeof(DMsg) * count), o_blob = in.add(ms.total),
                 o_runs = in.add(sizeof(DRun) * runs.size()), o_wd = in.add(sizeof(DWots) * wd.size()),
                 o_gi = in.add(sizeof(uint64_t) * gi.size());
    uint8_t* h = b->pin.need(in.size);
    fill_seeds(h + o_seeds, n, ks);
    b->seed_off = o_seeds;
    b->seed_bytes = ks.size() * 3 * n;
    memcpy(h + o_md, ms.d.data(), sizeof(DMsg) * count);
    fill_msgs(h + o_blob, fam, true, mr, ms);
    memcpy(h + o_runs, runs.data(), sizeof(DRun) * runs.size());
    memcpy(h + o_wd, wd.data(), sizeof(DWots) * wd.size());
    memcpy(h + o_gi, gi.data(), sizeof(uint64_t) * gi.size());
    uint8_t* dv = b->io.need(in.size);
    I->h2d(dv, h, in.size);
    b->seeds = dv + o_seeds;
    b->md = reinterpret_cast<const DMsg*>(dv + o_md);
    b->blob = dv + o_blob;
    b->runs = reinterpret_cast<const DRun*>(dv + o_runs);
    b->wd = reinterpret_cast<const DWots*>(dv + o_wd);
    b->gidx = reinterpret_cast<const uint64_t*>(dv + o_gi);
    o->prepare(I->L, b->seeds, (int)ks.size(), b->keys.need(o->key_bytes() * ks.size()));
    Layout out;
    const size_t o_R = out.add((size_t)count * n), o_auth = out.add(gi.size() * n),
                 o_w = out.add((size_t)count * d * b->len * n);
    uint8_t* dr = b->res.need(out.size);
    b->R = dr + o_R;
    b->auth = dr + o_auth;
    b->wots = dr + o_w;
    b->res_bytes = out.size;
    b->dig.need((size_t)count * n);
    b->nodes.need(b->nodes_stride * n * (size_t)b->nruns);
    b->roots.need((size_t)count * d * n);  // layer-major roots
    if (wait) I->sync();
    guard.ok = true;
    return b;
}

// Stages a verification batch's inputs: key contexts, messages, signatures.
// The message blob comes last, so every other device offset depends only on
// the batch's shape (count, d, th, n) and a refill of the same shape keeps
// the pointers a captured graph uses, unless the io buffer has to grow.
static void stage_verify(Engine::Impl* I, FamOps* o, Engine::Batch* b, const Engine::VerifyItem* items) {
    wait_batch(b);  // a launched batch still reads its inputs
    const int count = b->count;
    const unsigned n = b->n, len = b->len, th = b->th, d = b->d;
    // staged longest message first, so that run() can start the rest while
    // the longest messages are still being hashed: a stable counting sort by
    // 64-byte blocks (H_msg's cost), messages of 4 MB and more in one class
    std::vector<int>& ord = b->perm;
    ord.resize(count);
    {
        constexpr uint64_t kClasses = 1u << 16;
        std::vector<uint32_t> at(kClasses + 1, 0);
        auto cls = [&](int i) { return (size_t)(kClasses - 1 - std::min<uint64_t>(items[i].mlen / 64, kClasses - 1)); };
        for (int i = 0; i < count; ++i) ++at[cls(i) + 1];
        for (size_t c = 0; c < kClasses; ++c) at[c + 1] += at[c];
        for (int i = 0; i < count; ++i) ord[at[cls(i)]++] = i;
    }
    std::vector<KeySeeds> ks(count);
    std::vector<MsgReq> mr(count);
    for (int i = 0; i < count; ++i) {
        const Engine::VerifyItem& it = items[ord[i]];
        ks[i] = {nullptr, nullptr, it.pk + n};
        mr[i] = {(uint32_t)i, 0, it.idx, it.R, it.pk, it.m, it.mlen};
    }
    const MsgStage ms = plan_msgs(b->fam, count, false, mr);
    Layout in;
    const size_t o_seeds = in.add(ks.size() * 3 * n), o_md = in.add(sizeof(DMsg) * count),
                 o_vd = in.add(sizeof(DLeaf) * count * d), o_sig = in.add((size_t)count * d * len * n),
                 o_auth = in.add((size_t)count * d * th * n), o_pk = in.add((size_t)count * n),
                 o_blob = in.add(ms.total);
    uint8_t* h = b->pin.need(in.size);
    fill_seeds(h + o_seeds, n, ks);
    memcpy(h + o_md, ms.d.data(), sizeof(DMsg) * count);
    fill_msgs(h + o_blob, b->fam, false, mr, ms);
    DLeaf* vd = reinterpret_cast<DLeaf*>(h + o_vd);
    parallel_for(count, 512, [&](size_t i) {
        const Engine::VerifyItem& it = items[ord[i]];
        uint64_t idx = it.idx;
        const uint8_t* p = it.wots;
        for (unsigned L = 0; L < d; ++L) {
            const uint32_t leaf = (uint32_t)(idx & ((1ull << th) - 1));
            idx >>= th;
            const size_t t = (size_t)L * count + i;
            vd[t] = {(uint32_t)i, L, (uint32_t)(idx >> 32), (uint32_t)idx, leaf, th, LAT_ALL_STEPS, (uint32_t)(i * th)};
            memcpy(h + o_sig + t * len * n, p, (size_t)len * n);
            p += (size_t)len * n;
            memcpy(h + o_auth + t * th * n, p, (size_t)th * n);
            p += (size_t)th * n;
        }
        memcpy(h + o_pk + i * n, it.pk, n);
    });
    uint8_t* const before = (uint8_t*)b->io.p;
    uint8_t* dv = b->io.need(in.size);
    if (before && dv != before) b->drop_graph();  // the buffer moved: the captured pointers are stale
    // The copy and the key contexts go on the copy stream, so that staging
    // this batch overlaps the kernels of batches launched before it.
    HIPCK(hipMemcpyAsync(dv, h, in.size, hipMemcpyHostToDevice, I->s_copy));
    b->seeds = dv + o_seeds;
    b->md = reinterpret_cast<const DMsg*>(dv + o_md);
    b->blob = dv + o_blob;
    b->vd = reinterpret_cast<const DLeaf*>(dv + o_vd);
    b->sigs = dv + o_sig;
    b->auth_in = dv + o_auth;
    b->pkroot = dv + o_pk;
    void* keys = b->keys.need(o->key_bytes() * count);
    I->L.s = I->s_copy;
    o->prepare(I->L, b->seeds, count, keys);
    I->L.s = I->L.s_main;
    // Returns without waiting for the copy: the batch's run waits for it on
    // the device, and the next refill, collect or free of this batch on the
    // host (they reuse `pin`).  With per-kernel timing on, the copy stream is
    // drained here so that its timings can be read at the next wait.
    if (!b->staged) HIPCK(hipEventCreateWithFlags(&b->staged, hipEventDisableTiming));
    HIPCK(hipEventRecord(b->staged, I->s_copy));
    b->copying = true;
    if (I->L.prof) wait_batch(b);
}

Engine::Batch* Engine::verify_prepare(int fam, unsigned th, unsigned d, int count, const VerifyItem* items) {
    Impl* I = impl_;
    Batch* b = take_batch(I, false);
    BatchGuard guard{this, b};
    b->sign = false;
    b->one_stream = false;
    b->plan = verify_plan_at_prepare();
    b->tuned = false;
    b->fam = fam;
    b->th = th;
    b->d = d;
    b->n = fam_n(fam);
    b->len = fam_len(fam);
    b->count = count;
    FamOps* o = ops_of(fam);
    b->ok = b->res.need<int>(sizeof(int) * (size_t)count);
    b->res_bytes = sizeof(int) * (size_t)count;
    b->dig.need((size_t)count * b->n);
    b->roots.need((size_t)count * d * b->n);
    stage_verify(I, o, b, items);
    guard.ok = true;
    return b;
}

void Engine::verify_refill(Batch* b, const VerifyItem* items) { stage_verify(impl_, ops_of(b->fam), b, items); }

void Engine::set_stream(Batch* b, int sel) {
    sel = sel < 0 || sel > 2 ? 0 : sel;
    if (b->sel != sel) b->drop_graph();  // the graph is launched where it was captured
    b->sel = sel;
}

// While alive, the engine's main stream is the one a verification batch was
// given (set_stream); the launches, the capture and the completion event of
// the batch follow it.
struct StreamSel {
    Engine::Impl* I;
    hipStream_t main, cur;
    StreamSel(Engine::Impl* impl, const Engine::Batch* b) : I(impl), main(impl->L.s_main), cur(impl->L.s) {
        if (b->sign || b->sel == 0) return;
        I->L.s_main = I->L.s = b->sel == 1 ? I->L.s_aux : I->s_copy;
    }
    ~StreamSel() {
        I->L.s_main = main;
        I->L.s = cur;
    }
};

void Engine::set_one_stream(Batch* b, bool on) {
    if (b->one_stream != on) {
        b->drop_graph();  // the captured launches follow the stream layout
        b->tuned = false;  // and the best plan may differ
    }
    b->one_stream = on;
}

// ---- autotuner of the verification schedule ---------------------------
// The best plan of verify_dense depends on the parameter set (chain count,
// tree height, layers), the batch size and what shares the GPU (one-stream
// batches of mixed jobs).  A job of a shape not seen before tries every
// candidate plan on its second run: each is captured into a HIP graph and
// replayed three times on the job's own staged inputs (every replay is a
// whole, correct verification), the last two timed with HIP events on the
// stream the graph is launched on.  The fastest is kept -- the default unless
// another beats it by more than 2% -- for the job and for every later job of
// that shape in the process, and appended to the file XMSS_MI355X_TUNE_FILE
// names, from which later processes start.  XMSS_MI355X_TUNE=0 turns tuning
// off (the default plan).
namespace {
const uint32_t kVerifyCands[] = {
    kVplanDefault,                      // L-tree through LDS, then the auth walks; short group first
    kVplanDefault | VPLAN_PRIO_HMSG,    // H_msg at wave priority 2
    1u << 2,                            // the long group first
    VPLAN_SHORT_FIRST,                  // one launch per L-tree height
    kVplanDefault | VPLAN_PAIRED,       // two chains of complementary length per lane
    kVplanDefault | VPLAN_PRE_H_FIRST,  // H pre-images beside H_msg
};
constexpr size_t kNumVerifyCands = sizeof(kVerifyCands) / sizeof(kVerifyCands[0]);

struct VerifyTuner {
    std::mutex mu;
    std::map<uint64_t, uint32_t> best;  // shape key -> plan
    int forced = -1;                    // >= 0: every job uses this plan
    bool off = false;
    bool loaded = false;
    std::string file;

    void load_locked() {
        if (loaded) return;
        loaded = true;
        const char* t = getenv("XMSS_MI355X_TUNE");
        off = t && t[0] == '0';
        const char* f = getenv("XMSS_MI355X_TUNE_FILE");
        if (!f || !*f) return;
        file = f;
        FILE* fp = fopen(f, "r");
        if (!fp) return;
        unsigned long long key;
        unsigned plan;
        char kind[16];
        while (fscanf(fp, "%15s %llx %u", kind, &key, &plan) == 3)
            if (!strcmp(kind, "verify")) best[key] = plan;
        fclose(fp);
    }
    void record_locked(uint64_t key, uint32_t plan) {
        best[key] = plan;
        if (file.empty()) return;
        if (FILE* fp = fopen(file.c_str(), "a")) {
            fprintf(fp, "verify %llx %u\n", (unsigned long long)key, plan);
            fclose(fp);
        }
    }
};
VerifyTuner& vtuner() {
    static VerifyTuner* t = new VerifyTuner;  // lives as long as the process (engines do)
    return *t;
}
uint64_t verify_key(const Engine::Batch* b) {
    uint32_t lg = 0;
    while ((2u << lg) <= (uint32_t)b->count) ++lg;  // floor(log2(count))
    return (uint64_t)b->fam | (uint64_t)b->th << 8 | (uint64_t)b->d << 16 | (uint64_t)lg << 24 |
           (uint64_t)(b->one_stream ? 1 : 0) << 32;
}
}  // namespace

uint32_t verify_plan_at_prepare() {
    VerifyTuner& t = vtuner();
    std::lock_guard<std::mutex> lk(t.mu);
    t.load_locked();
    return t.forced >= 0 ? (uint32_t)t.forced : kVplanDefault;
}

int set_verify_plan(int plan) {
    VerifyTuner& t = vtuner();
    std::lock_guard<std::mutex> lk(t.mu);
    t.load_locked();
    const int prev = t.forced;
    t.forced = plan < 0 ? -1 : plan;
    return prev;
}

size_t verify_plans(uint32_t* out, size_t max) {
    for (size_t i = 0; i < kNumVerifyCands && i < max; ++i) out[i] = kVerifyCands[i];
    return kNumVerifyCands;
}

std::string tuned_plans() {
    VerifyTuner& t = vtuner();
    std::lock_guard<std::mutex> lk(t.mu);
    t.load_locked();
    std::string s;
    char line[96];
    for (auto& kv : t.best) {
        snprintf(line, sizeof line, "verify fam=%u th=%u d=%u log2count=%u one_stream=%u plan=%u\n",
                 (unsigned)(kv.first & 255), (unsigned)(kv.first >> 8 & 255), (unsigned)(kv.first >> 16 & 255),
                 (unsigned)(kv.first >> 24 & 255), (unsigned)(kv.first >> 32 & 1), kv.second);
        s += line;
    }
    return s;
}

// Large verification batches (staged longest message first).  The requests
// split into two groups by message length, each on its own stream: the
// longest 1/div (aux) and the rest (main).  Each group runs H_msg -> per
// layer: F pre-images of only the chain steps its signatures need -> chains
// (sorted by length within each workgroup) -> L-tree -> auth walk.  H_msg is
// serial per message and fills few lanes, so the short messages' pipeline
// starts while the long ones are still being hashed.  The H pre-images of
// every layer and request (L-tree merges and auth levels: address only) run
// once; only the L-trees wait for them.  The plan (VerifyPlan bits) chooses
// where they run, the tail kernels, the split and the wave priorities.
static void verify_dense(Engine::Impl* I, FamOps* o, Engine::Batch* b, uint8_t* vscratch) {
    const unsigned n = b->n, len = b->len, th = b->th, d = b->d;
    const uint32_t count = (uint32_t)b->count;
    const void* dk = b->keys.p;
    const size_t lb = o->lat_bytes(th);
    const uint32_t plan = b->plan;
    const uint32_t div = vplan_div(plan), tail = vplan_tail(plan);
    const uint32_t prio_long = (plan & VPLAN_PRIO) ? 1u : 0u, prio_hmsg = (plan & VPLAN_PRIO_HMSG) ? 1u : 0u;
    const bool hfirst = (plan & VPLAN_PRE_H_FIRST) && !b->one_stream;
    uint8_t* dig = (uint8_t*)b->dig.p;
    uint8_t* roots = (uint8_t*)b->roots.p;
    uint8_t* ends = b->ends.need((size_t)count * len * n);
    uint8_t* tmp = b->tmp.need((size_t)count * ((len + 1) / 2) * n);
    constexpr int G = 2;
    const uint32_t lo[G] = {0, count / div}, hi[G] = {count / div, count};
    const uint32_t prio[G] = {prio_long, 0};
    // one stream: the batch shares the GPU with other engines' batches (a
    // mixed job), which take the other hardware queues
    hipStream_t st[G] = {b->one_stream ? I->L.s_main : I->L.s_aux, I->L.s_main};
    hipStream_t sh = hfirst ? I->s_copy : I->L.s_main;  // the H pre-images
    hipEvent_t staged = I->L.ev(), hpre = I->L.ev(), done[G - 1] = {I->L.ev()};
    HIPCK(hipEventRecord(staged, I->L.s_main));
    for (int g = 0; g < G - 1; ++g) HIPCK(hipStreamWaitEvent(st[g], staged, 0));
    if (hfirst) {  // beside H_msg, on a third stream
        HIPCK(hipStreamWaitEvent(sh, staged, 0));
        I->L.s = sh;
        o->lat_pre_h(I->L, dk, b->vd, count * d, th, vscratch);
        HIPCK(hipEventRecord(hpre, sh));
    }
    // Launches are issued stage by stage across the groups, so that no stream
    // waits for the host to issue the others' launches.
    for (int k = 0; k < G; ++k) {
        const int g = (plan & VPLAN_SHORT_FIRST) ? G - 1 - k : k;
        if (hi[g] == lo[g]) continue;
        I->L.s = st[g];
        o->hmsg_verify(I->L, dk, b->md + lo[g], hi[g] - lo[g], fam_pad(b->fam), b->blob, dig + (size_t)lo[g] * n,
                       prio_hmsg | prio[g]);
    }
    for (unsigned L = 0; L < d; ++L) {
        for (int stage = 0; stage < 3; ++stage) {
            if (L == 0 && stage == 1 && !hfirst) {
                I->L.s = I->L.s_main;
                o->lat_pre_h(I->L, dk, b->vd, count * d, th, vscratch);
                HIPCK(hipEventRecord(hpre, I->L.s_main));
            }
            if (L == 0 && stage == 2)  // the L-trees need the H pre-images
                for (int g = 0; g < G; ++g)
                    if (st[g] != sh) HIPCK(hipStreamWaitEvent(st[g], hpre, 0));
            for (int g = G - 1; g >= 0; --g) {
                const uint32_t a = lo[g], m = hi[g] - lo[g];
                if (!m) continue;
                I->L.s = st[g];
                const size_t t = (size_t)L * count + a;  // first request of the group in this layer
                const uint8_t* msg = L == 0 ? dig + (size_t)a * n : roots + ((size_t)(L - 1) * count + a) * n;
                const DLeaf* vd = b->vd + t;
                uint8_t* scr = vscratch + t * lb;
                uint8_t* e = ends + (size_t)a * len * n;
                if (stage == 0)
                    o->pre_need(I->L, dk, vd, m, msg, th, scr, prio[g]);
                else if (stage == 1)
                    o->latb_chains(I->L, m, msg, b->sigs + t * len * n, scr, lb, e, prio[g],
                                   (plan & VPLAN_PAIRED) != 0);
                else
                    o->latb_tail(I->L, vd, m, b->auth_in + (size_t)L * count * th * n, scr, lb, e,
                                 tmp + (size_t)a * ((len + 1) / 2) * n, roots + t * n, tail, prio[g]);
            }
        }
    }
    for (int g = 0; g < G - 1; ++g) {
        HIPCK(hipEventRecord(done[g], st[g]));
        HIPCK(hipStreamWaitEvent(I->L.s_main, done[g], 0));
    }
    I->L.s = I->L.s_main;
    I->L.run("compare", k_compare, dim3(cdiv(count, WG)), dim3(WG), (const uint8_t*)roots + (size_t)(d - 1) * count * n,
             b->pkroot, count, n, b->ok);
    I->L.pool.push_back(staged);
    I->L.pool.push_back(hpre);
    for (hipEvent_t e : done) I->L.pool.push_back(e);
}

void Engine::wait() { impl_->sync(); }

void Engine::launch(Batch* b) {
    wait_batch(b);
    StreamSel sel(impl_, b);
    run(b, false);
    if (!b->done) HIPCK(hipEventCreateWithFlags(&b->done, hipEventDisableTiming));
    // both paths end on s_main with the aux stream joined into it
    HIPCK(hipEventRecord(b->done, impl_->L.s_main));
    b->inflight = b->evented = true;
}

void Engine::wait_batch(Batch* b) { xm::wait_batch(b); }

// The launches of one verification batch: verify_dense for large batches,
// else H_msg beside the address-only work of every layer (all 15 steps of
// each chain, aux stream), then one finishing launch per layer.
static void verify_launches(Engine::Impl* I, FamOps* o, Engine::Batch* b) {
    const unsigned n = b->n, len = b->len, th = b->th, d = b->d;
    const int count = b->count;
    const void* dk = b->keys.p;
    const size_t nv = (size_t)count * d;
    const bool vlat = use_lat(o, nv, th);
    uint8_t* vscratch = vlat ? b->lat.need(nv * o->lat_bytes(th)) : nullptr;
    if (vlat && count >= kLatDense) {
        verify_dense(I, o, b, vscratch);
        return;
    }
    hipEvent_t staged = nullptr, pre = nullptr;
    if (vlat) {
        staged = I->L.ev();
        pre = I->L.ev();
        HIPCK(hipEventRecord(staged, I->L.s_main));
        HIPCK(hipStreamWaitEvent(I->L.s_aux, staged, 0));
        I->L.s = I->L.s_aux;
        o->lat_pre(I->L, dk, LAT_VERIFY, b->vd, (uint32_t)nv, nullptr, th, vscratch);
        HIPCK(hipEventRecord(pre, I->L.s_aux));
        I->L.s = I->L.s_main;
    }
    o->hmsg(I->L, dk, b->md, (uint32_t)count, fam_pad(b->fam), b->blob, b->R, (uint8_t*)b->dig.p);
    if (vlat) HIPCK(hipStreamWaitEvent(I->L.s_main, pre, 0));
    for (unsigned L = 0; L < d; ++L) {
        const uint8_t* msg_l = L == 0 ? (const uint8_t*)b->dig.p : (const uint8_t*)b->roots.p + (size_t)(L - 1) * count * n;
        const DLeaf* vd = b->vd + (size_t)L * count;
        const uint8_t* sig_l = b->sigs + (size_t)L * count * len * n;
        const uint8_t* auth_l = b->auth_in + (size_t)L * count * th * n;
        uint8_t* root_l = (uint8_t*)b->roots.p + (size_t)L * count * n;
        if (vlat)
            o->lat_fin(I->L, LAT_VERIFY, vd, (uint32_t)count, msg_l, sig_l, auth_l, th,
                       vscratch + (size_t)L * count * o->lat_bytes(th), root_l);
        else
            o->leaf_wg(I->L, dk, vd, (uint32_t)count, true, msg_l, sig_l, auth_l, th, root_l);
    }
    I->L.run("compare", k_compare, dim3(cdiv(count, WG)), dim3(WG),
             (const uint8_t*)b->roots.p + (size_t)(d - 1) * count * n, b->pkroot, (uint32_t)count, n, b->ok);
    if (vlat) {
        I->L.pool.push_back(staged);
        I->L.pool.push_back(pre);
    }
}

static hipGraphExec_t capture_verify(Engine::Impl* I, FamOps* o, Engine::Batch* b) {
    hipGraph_t g;
    hipGraphExec_t x = nullptr;
    HIPCK(hipStreamBeginCapture(I->L.s_main, hipStreamCaptureModeRelaxed));
    verify_launches(I, o, b);
    HIPCK(hipStreamEndCapture(I->L.s_main, &g));
    const hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCK(e);
    return x;
}

// Settles b->plan before the batch's graph is captured: the forced plan, the
// one already found for this shape, or -- for a batch on the lane-dense path
// -- the fastest candidate, timed here (b->graph is then that candidate's
// graph, replayed once more so that its results are the run's).
static void settle_plan(Engine::Impl* I, FamOps* o, Engine::Batch* b) {
    b->tuned = true;
    VerifyTuner& t = vtuner();
    const uint64_t key = verify_key(b);
    {
        std::lock_guard<std::mutex> lk(t.mu);
        t.load_locked();
        if (t.forced >= 0) {
            b->plan = (uint32_t)t.forced;
            return;
        }
        auto it = t.best.find(key);
        if (it != t.best.end()) {
            b->plan = it->second;
            return;
        }
        if (t.off) {
            b->plan = kVplanDefault;
            return;
        }
    }
    const size_t nv = (size_t)b->count * b->d;
    if (!(use_lat(o, nv, b->th) && b->count >= kLatDense)) {  // the plan does not apply
        b->plan = kVplanDefault;
        return;
    }
    hipEvent_t e0 = I->L.ev(), e1 = I->L.ev();
    float best_ms = 0, default_ms = 0;
    uint32_t best_plan = kVplanDefault;
    hipGraphExec_t best_graph = nullptr;
    try {
        for (size_t c = 0; c < kNumVerifyCands; ++c) {
            b->plan = kVerifyCands[c];
            hipGraphExec_t x = capture_verify(I, o, b);
            float ms = 0;
            for (int rep = 0; rep < 3; ++rep) {
                HIPCK(hipEventRecord(e0, I->L.s_main));
                HIPCK(hipGraphLaunch(x, I->L.s_main));
                HIPCK(hipEventRecord(e1, I->L.s_main));
                HIPCK(hipEventSynchronize(e1));
                float r = 0;
                HIPCK(hipEventElapsedTime(&r, e0, e1));
                if (rep == 1 || (rep == 2 && r < ms)) ms = r;
            }
            if (kVerifyCands[c] == kVplanDefault) default_ms = ms;
            if (!best_graph || ms < best_ms) {
                if (best_graph) (void)hipGraphExecDestroy(best_graph);
                best_graph = x, best_ms = ms, best_plan = kVerifyCands[c];
            } else {
                (void)hipGraphExecDestroy(x);
            }
        }
    } catch (...) {
        if (best_graph) (void)hipGraphExecDestroy(best_graph);
        I->L.pool.push_back(e0);
        I->L.pool.push_back(e1);
        b->plan = kVplanDefault;
        throw;
    }
    I->L.pool.push_back(e0);
    I->L.pool.push_back(e1);
    if (best_plan != kVplanDefault && best_ms > 0.98f * default_ms) {
        // not clearly faster: keep the default, captured by the caller (every
        // candidate's last replay was waited for: the graph is idle)
        (void)hipGraphExecDestroy(best_graph);
        best_graph = nullptr;
        best_plan = kVplanDefault;
    }
    b->plan = best_plan;
    if (best_graph) {
        // the last replay was the last candidate's: replay the kept graph so
        // that this run's results come from it like every later run's
        b->graph = best_graph;
        HIPCK(hipGraphLaunch(b->graph, I->L.s_main));
    }
    std::lock_guard<std::mutex> lk(t.mu);
    t.record_locked(key, best_plan);
}

// A verification batch: launched directly on its first run (which sizes
// every buffer), captured into a HIP graph on its second and replayed from
// then on -- one launch call instead of tens.  The capture holds no
// allocation, memset or host copy; with per-kernel timing on, the launches
// stay direct.
static void run_verify(Engine::Impl* I, Engine::Batch* b) {
    FamOps* o = ops_of(b->fam);
    if (b->copying) {  // the inputs are staged on the copy stream (outside any capture)
        HIPCK(hipStreamWaitEvent(I->L.s_main, b->staged, 0));
        b->copying = false;
    }
    if (b->graph && !I->L.prof) {
        HIPCK(hipGraphLaunch(b->graph, I->L.s_main));
        return;
    }
    if (I->L.prof || b->nrun++ < 1) {
        verify_launches(I, o, b);
        return;
    }
    if (!b->tuned) settle_plan(I, o, b);
    if (b->graph) {  // tuning left the best candidate's graph (and its results)
        return;
    }
    b->graph = capture_verify(I, o, b);
    HIPCK(hipGraphLaunch(b->graph, I->L.s_main));
}

void Engine::set_graphed(Batch* b, bool on) {
    if (b->graphed != on) b->drop_graph();
    b->graphed = on;
}

static void sign_launches(Engine::Impl* I, Engine::Batch* b);

void Engine::run(Batch* b, bool wait_done) {
    Impl* I = impl_;
    if (!b->sign) {
        StreamSel sel(I, b);
        run_verify(I, b);
        if (wait_done) I->sync();
        return;
    }
    if (b->graphed && !I->L.prof) {
        if (!b->graph) {
            I->L.dry = true;  // sizes every workspace: the capture allocates nothing
            try {
                sign_launches(I, b);
            } catch (...) {
                I->L.dry = false;
                throw;
            }
            I->L.dry = false;
            hipGraph_t g = nullptr;
            HIPCK(hipStreamBeginCapture(I->L.s_main, hipStreamCaptureModeRelaxed));
            try {
                sign_launches(I, b);
            } catch (...) {
                if (hipStreamEndCapture(I->L.s_main, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
                throw;
            }
            HIPCK(hipStreamEndCapture(I->L.s_main, &g));
            const hipError_t e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIPCK(e);
        }
        HIPCK(hipGraphLaunch(b->graph, I->L.s_main));
    } else {
        sign_launches(I, b);
    }
    if (wait_done) I->sync();
}

// The launches of a signing batch (on the main stream, forking to the aux
// stream and joining it again); with L.dry only the workspaces are sized.
static void sign_launches(Engine::Impl* I, Engine::Batch* b) {
    FamOps* o = ops_of(b->fam);
    const unsigned n = b->n, len = b->len, th = b->th, d = b->d;
    const int count = b->count;
    const void* dk = b->keys.p;
    // WOTS signatures of layer L (on the current stream): few requests with
    // the latency kernels (address-only work wide, then one compression per
    // chain step; scratch per layer), else one lane per chain (k_wots)
    const bool lat_wots = count <= kLatWots && use_lat(o, count, 0);
    DLeaf* dw = nullptr;
    uint8_t* wlat = lat_wots ? b->wlat.need((size_t)d * count * o->lat_bytes(0)) : nullptr;
    auto sign_layer = [&](unsigned L, const uint8_t* msg_l) {
        if (lat_wots)
            o->lat(I->L, dk, LAT_SIGN, dw + (size_t)L * count, (uint32_t)count, msg_l, nullptr, nullptr, 0,
                   wlat + (size_t)L * count * o->lat_bytes(0), b->wots + (size_t)L * count * len * n);
        else
            o->wots(I->L, dk, b->wd + (size_t)L * count, (uint32_t)count, msg_l, nullptr,
                    b->wots + (size_t)L * count * len * n);
    };
    hipEvent_t wdone = nullptr;
    {
        // all subtrees in one pipeline (few: levels split as in build_trees);
        // R / H_msg (needed only by the WOTS signatures) and the levels'
        // address-only work on the aux stream beside the leaves
        const int nruns = b->nruns;
        const bool split = nruns <= kSplitRuns;
        void* tpre = split ? I->tpre.need((size_t)nruns * ((1ull << th) - 1) * o->hpre_bytes()) : nullptr;
        hipEvent_t tstaged = I->L.ev(), tdone = I->L.ev();
        wdone = I->L.ev();
        HIPCK(hipEventRecord(tstaged, I->L.s_main));
        HIPCK(hipStreamWaitEvent(I->L.s_aux, tstaged, 0));
        I->L.s = I->L.s_aux;
        o->hmsg(I->L, dk, b->md, (uint32_t)count, fam_pad(b->fam), b->blob, b->R, (uint8_t*)b->dig.p);
        if (split) o->pre_tree(I->L, dk, b->runs, nruns, th, tpre);
        if (lat_wots) {
            dw = b->wreq.need<DLeaf>((size_t)d * count * sizeof(DLeaf));
            I->L.run("wots_to_leaves", k_wots_to_leaves, dim3(cdiv((uint64_t)d * count, WG)), dim3(WG), b->wd,
                     (uint32_t)(d * count), dw);
        }
        HIPCK(hipEventRecord(tdone, I->L.s_aux));
        // the layer-0 WOTS signatures sign the digests, not a subtree root:
        // on the aux stream beside the leaves
        sign_layer(0, (const uint8_t*)b->dig.p);
        HIPCK(hipEventRecord(wdone, I->L.s_aux));
        I->L.s = I->L.s_main;
        const uint64_t leaves = 1ull << th;
        const uint64_t per_leaf = (uint64_t)len * n;
        const uint64_t stride = b->nodes_stride;
        uint8_t* nodes = (uint8_t*)b->nodes.p;
        const DRun* drun = b->runs;
        // few leaves (a single signature's trees): the latency kernels -- one
        // wide launch of address-only work, then the chains at one
        // compression per step and the L-trees through LDS -- instead of
        // k_chains at about one wave per SIMD and seven L-tree launches
        const uint64_t nlv = (uint64_t)nruns * leaves;
        const bool lat_leaves = split && nlv <= kLatLeaves && use_lat(o, nlv, 0);
        if (lat_leaves) {
            DLeaf* dl = b->lreq.need<DLeaf>(nlv * sizeof(DLeaf));
            I->L.run("runs_to_leaves", k_runs_to_leaves, dim3(cdiv(nlv, WG)), dim3(WG), drun, (uint32_t)nruns, th, dl);
            uint8_t* lv = b->tmp.need(nlv * n);
            o->lat_leaves_dense(I->L, dk, dl, (uint32_t)nlv, b->lat.need(nlv * o->lat_bytes(0)),
                                b->ends.need(nlv * len * n), lv);
            if (!I->L.dry)
                HIPCK(hipMemcpy2DAsync(nodes, stride * n, lv, leaves * n, leaves * n, nruns,
                                       hipMemcpyDeviceToDevice, I->L.s));
            HIPCK(hipStreamWaitEvent(I->L.s_main, tdone, 0));
            o->levels_fin(I->L, nruns, th, nodes, tpre);
        }
        for (int r0 = lat_leaves ? nruns : 0; r0 < nruns;) {
            int rg = std::min<int>(nruns - r0, 65535);
            uint64_t lpp = leaves;
            while ((uint64_t)rg * lpp * per_leaf > kEndsBudget && lpp > 64) lpp >>= 1;
            while ((uint64_t)rg * lpp * per_leaf > kEndsBudget && rg > 1) rg = (rg + 1) / 2;
            uint8_t* ends = b->ends.need((size_t)rg * lpp * per_leaf);
            uint8_t* tmp = b->tmp.need((size_t)rg * lpp * ((len + 1) / 2) * n);
            uint8_t* tnodes = nodes + (size_t)r0 * stride * n;
            for (uint64_t l0 = 0; l0 < leaves; l0 += lpp) {
                const uint32_t cnt = (uint32_t)std::min<uint64_t>(lpp, leaves - l0);
                o->chains(I->L, dk, drun + r0, rg, (uint32_t)l0, cnt, ends);
                uint32_t l = len, height = 0;
                bool from_a = true;
                const uint32_t sa = len, sb = (len + 1) / 2;
                while (l > 1) {
                    o->ltree_level(I->L, dk, drun + r0, rg, (uint32_t)l0, cnt, l, height, from_a ? ends : tmp,
                                   from_a ? sa : sb, from_a ? tmp : ends, from_a ? sb : sa, tnodes, stride, l0);
                    from_a = !from_a;
                    l = (l + 1) / 2;
                    ++height;
                }
            }
            if (split) {
                HIPCK(hipStreamWaitEvent(I->L.s_main, tdone, 0));
                o->levels_fin(I->L, rg, th, tnodes,
                              (const uint8_t*)tpre + (size_t)r0 * ((1ull << th) - 1) * o->hpre_bytes());
            } else {
                for (unsigned lev = 0; lev < th; ++lev) o->level(I->L, dk, drun + r0, rg, th, lev, tnodes);
            }
            r0 += rg;
        }
        HIPCK(hipStreamWaitEvent(I->L.s_main, tdone, 0));  // the digests
        I->L.pool.push_back(tstaged);
        I->L.pool.push_back(tdone);
        // auth paths and roots: auth[t] = th auth nodes || root, t = layer * count + item
        const uint32_t ng = (uint32_t)((size_t)count * d * (th + 1));
        I->L.run("gather", k_gather, dim3(cdiv((uint64_t)ng * (n / 4), WG)), dim3(WG), (const uint8_t*)nodes,
                 b->gidx, ng, n, b->auth);
        // WOTS signatures of the upper layers: layer L signs the root of layer
        // L-1.  Roots sit at the end of each auth record.
        uint8_t* msgs = (uint8_t*)b->roots.p;
        const size_t rec = (size_t)(th + 1) * n;
        for (unsigned L = 1; L < d; ++L) {
            // compact roots of layer L-1 into msgs[L-1]
            uint8_t* dst = msgs + (size_t)(L - 1) * count * n;
            if (!I->L.dry)
                HIPCK(hipMemcpy2DAsync(dst, n, b->auth + ((size_t)(L - 1) * count) * rec + th * n, rec, n, count,
                                       hipMemcpyDeviceToDevice, I->L.s));
            sign_layer(L, dst);
        }
    }
    HIPCK(hipStreamWaitEvent(I->L.s_main, wdone, 0));  // the layer-0 signatures
    I->L.pool.push_back(wdone);
}

// Results of a batch run with run() leave on its own stream, ordered after
// its kernels (one wait); those of a launched batch, once it is done, on the
// aux stream: batches launched after it replay their graphs from the main
// stream, and the next batch's input copy on the copy stream does not delay
// them.  (No fourth stream: HIP maps a process's streams onto a few hardware
// queues -- four by default -- in creation order, and with three streams per
// engine the main streams of engines made one after another, such as the
// serving lanes, land on different queues; a fourth stream in every engine
// put them all on one and serialized their kernels.)
static void fetch_results(Engine::Impl* I, Engine::Batch* b, void* h, const void* src, size_t bytes) {
    if (b->evented) {
        wait_batch(b);
        HIPCK(hipMemcpyAsync(h, src, bytes, hipMemcpyDeviceToHost, I->L.s_aux));
        HIPCK(hipStreamSynchronize(I->L.s_aux));
    } else {
        I->d2h(h, src, bytes);
        I->sync();
    }
}

void Engine::collect_sign(Batch* b, uint8_t* const* out) {
    Impl* I = impl_;
    const unsigned n = b->n, len = b->len, th = b->th, d = b->d;
    const int count = b->count;
    uint8_t* h = b->pin.need(b->res_bytes);
    fetch_results(I, b, h, b->res.p, b->res_bytes);
    const uint8_t* R = h + (b->R - (uint8_t*)b->res.p);
    const uint8_t* auth = h + (b->auth - (uint8_t*)b->res.p);
    const uint8_t* w = h + (b->wots - (uint8_t*)b->res.p);
    parallel_for(count, 512, [&](size_t i) {
        uint8_t* p = out[i];
        memcpy(p, R + i * n, n);
        p += n;
        for (unsigned L = 0; L < d; ++L) {
            const size_t t = (size_t)L * count + i;
            memcpy(p, w + t * len * n, (size_t)len * n);
            p += (size_t)len * n;
            memcpy(p, auth + t * (th + 1) * n, (size_t)th * n);
            p += (size_t)th * n;
        }
    });
}

void Engine::collect_verify(Batch* b, int* ok) {
    Impl* I = impl_;
    int* h = reinterpret_cast<int*>(b->pin.need(sizeof(int) * (size_t)b->count));
    fetch_results(I, b, h, b->ok, sizeof(int) * (size_t)b->count);
    for (int i = 0; i < b->count; ++i) ok[b->perm[i]] = h[i];
}

void Engine::free_batch(Batch* b) {
    wait_batch(b);
    b->evented = false;
    if (b->sign) {
        // a signing batch's seeds (staging, device copy) and key contexts are
        // overwritten with zeros before its buffers are kept or freed
        Impl* I = impl_;
        if (b->seed_bytes) {
            volatile uint8_t* v = b->pin.p + b->seed_off;
            for (size_t i = 0; i < b->seed_bytes; ++i) v[i] = 0;
            zero_async(I->L, (void*)b->seeds, (b->seed_bytes + 3) & ~(size_t)3);
        }
        if (b->keys.p) zero_async(I->L, b->keys.p, b->keys.cap);
        b->seed_bytes = 0;
        I->sync();
    }
    if (!impl_->spare[b->sign]) {  // keep one of each kind: its buffers only grow
        impl_->spare[b->sign] = b;
        return;
    }
    destroy_batch(b);
}

void destroy_batch(Engine::Batch* b) {
    b->drop_graph();
    if (b->done) HIPCK(hipEventDestroy(b->done));
    if (b->staged) HIPCK(hipEventDestroy(b->staged));
    if (b->pin.p) HIPCK(hipHostFree(b->pin.p));
    DevBuf* bufs[] = {&b->io,    &b->res, &b->keys, &b->dig,  &b->nodes, &b->ends,
                      &b->tmp,   &b->roots, &b->lat, &b->lreq, &b->wlat,  &b->wreq};
    for (DevBuf* x : bufs) x->release();
    delete b;
}

}  // namespace xm

namespace xm {


struct BdsBufs {
    DevBuf io, leaf_arena, dig, scratch, lat_leaf, merge_arena, vals, dagv, mid;
    PinnedBuf staging;
};


void Engine::bds_sign_batch(int fam, int count, BdsJob* jobs) {
    // chunks that fit one three-launch sequence: the pre-image launch's rows
    // (planned leaves, signatures, message and merge rows) within the grid
    // limit and the latency kernels' scratch within kLatScratch
    const size_t lb = ops_of(fam)->lat_bytes(0);
    int start = 0;
    size_t rows = 0, extra = 0;
    for (int i = 0; i <= count; ++i) {
        const size_t r = i < count ? jobs[i].nplan + 1 + (jobs[i].plan_wots ? jobs[i].nwots : 0) : 0,
                     x = i < count ? jobs[i].nmerge + 1 : 0;
        const size_t rows2 = rows + r, extra2 = extra + x;
        if (i == count || (i > start && (rows2 + extra2 / WG + 2 > 65535 || rows2 * lb > kLatScratch ||
                                         i - start >= 32767))) {  // the update launch has 2 blocks per job
            if (i > start) bds_sign_chunk(fam, i - start, jobs + start);
            start = i;
            rows = r;
            extra = x;
        } else {
            rows = rows2;
            extra = extra2;
        }
    }
}

void Engine::bds_sign_chunk(int fam, int count, BdsJob* jobs) {
    Impl* I = impl_;
    if (!I->bds) I->bds = new BdsBufs;
    BdsBufs& B = *I->bds;
    FamOps* o = ops_of(fam);
    const unsigned n = fam_n(fam), len = fam_len(fam), pad = fam_pad(fam);
    keys_for(this, I, fam);
    const void* dk = I->kp();
    // per-job layout: planned leaves, sk bytes, signature tail
    std::vector<DLeaf> ld;
    std::vector<DThash> mg;
    std::vector<uint32_t> dag;  // 3 words per merge, parallel to mg
    std::vector<uint8_t> lits;
    std::vector<DLeaf> up;      // upper-layer WOTS signatures of the state updates (pre-images only)
    std::vector<DBds> jd(count);
    std::vector<size_t> tail_bytes(count);
    size_t sk_total = 0, tail_total = 0, lds = 0;
    // chain ends hashed ahead (ends_ahead): a planned leaf whose chain ends
    // a slot holds only gets its L-tree here (LAT_ENDS_IN)
    EndsAhead& E = I->ahead;
    const size_t eb = (size_t)len * n;
    uint8_t* ends = nullptr;
    if (!E.want.empty() || !E.have.empty()) {
        try {  // the slots are an optimization: without memory for them, every leaf is hashed whole
            ends = E.buf.need(kEndsSlots * eb);
        } catch (const DeviceError&) {
            E.clear();
        }
        if (ends != E.at) E.have.clear();  // a new buffer holds no chain ends
        E.at = ends;
    }
    std::vector<EndsAhead::Id> used;
    std::vector<bool> busy(kEndsSlots, false);
    for (int i = 0; i < count; ++i) {
        const BdsShape& s = jobs[i].shape;
        jd[i].leaf_off = ld.size();
        const bool ahead = ends && E.fam == fam && !E.have.empty() &&
                           seed_digest(n, {jobs[i].sk_seed, jobs[i].sk_prf, jobs[i].pub_seed}) == E.d;
        for (uint32_t k = 0; k < jobs[i].nplan; ++k) {
            const uint64_t t = jobs[i].plan_tree[k];
            DLeaf q{jobs[i].key, jobs[i].plan_layer[k], (uint32_t)(t >> 32), (uint32_t)t, jobs[i].plan_leaf[k], 0, 0, 0};
            if (ahead) {
                const EndsAhead::Id id{q.layer, t, q.leaf};
                auto it = E.have.find(id);
                if (it != E.have.end()) {
                    q.p1 = LAT_ENDS_IN;
                    q.p2 = it->second;
                    busy[it->second] = true;
                    used.push_back(id);
                }
            }
            ld.push_back(q);
        }
        jd[i].merge_off = mg.size();
        for (uint32_t k = 0; k < jobs[i].nmerge; ++k) {
            DThash t{};
            t.key = jobs[i].key;
            for (int w = 0; w < 8; ++w) t.adrs[w] = jobs[i].plan_merge[8 * k + w];
            mg.push_back(t);
        }
        tail_bytes[i] = (size_t)s.sig_bytes - s.index_bytes - n - s.wots_sig_bytes;
        jd[i].sk_off = sk_total;
        sk_total += (s.sk_bytes + 7) & ~7ull;
        jd[i].tail_off = tail_total;
        tail_total += tail_bytes[i];
        jd[i].idx = jobs[i].idx;
        jd[i].key = jobs[i].key;
        jd[i].nmerge = jobs[i].nmerge;
        jd[i].res = reinterpret_cast<uint64_t>(jobs[i].dsk);
        jd[i].sk_bytes = (uint32_t)s.sk_bytes;
        BdsP& P = jd[i].P;
        P = {s.th, s.d, s.full_height, s.index_bytes, s.bds_k, s.wots_sig_bytes, s.sig_bytes, 0};
        P.state_bytes = (s.th + 1) * n + 4 + s.th + 1 + s.th * n + (s.th >> 1) * n + (s.th - s.bds_k) * (7 + n) +
                        ((1u << s.bds_k) - s.bds_k - 1) * n + 4;
        const size_t region = (size_t)(2 * s.d - 1) * P.state_bytes;
        // the merge DAG: its values go to LDS after the states, if they fit
        const size_t with_dag = ((region + 15) & ~(size_t)15) + (size_t)jobs[i].nmerge * n;
        jd[i].maxlev = 0;
        jd[i].lit_off = lits.size() / n;
        if (jobs[i].plan_dag && jobs[i].nmerge && with_dag <= kBdsLds) {
            for (uint32_t k = 0; k < jobs[i].nmerge; ++k)
                jd[i].maxlev = std::max(jd[i].maxlev, jobs[i].plan_dag[3 * k + 2]);
            dag.insert(dag.end(), jobs[i].plan_dag, jobs[i].plan_dag + 3 * jobs[i].nmerge);
            lits.insert(lits.end(), jobs[i].plan_lit, jobs[i].plan_lit + (size_t)jobs[i].nlit * n);
        } else {
            dag.resize(dag.size() + 3 * jobs[i].nmerge, 0);
        }
        const size_t used = jd[i].maxlev ? with_dag : region;
        // the planned leaves' values staged in LDS after the states and merges
        const size_t lv_off = (used + 15) & ~(size_t)15, with_lv = lv_off + (size_t)jobs[i].nplan * n;
        jd[i].nleaf = jobs[i].nplan;
        jd[i].lv_off = jobs[i].nplan && with_lv <= kBdsLds ? (uint32_t)lv_off : 0;
        lds = std::max(lds, jd[i].lv_off ? with_lv : used);
        jd[i].wots_row = (uint32_t)up.size();
        jd[i].nwots = jobs[i].plan_wots ? jobs[i].nwots : 0;
        for (uint32_t k = 0; k < jd[i].nwots; ++k) {
            const uint32_t* w = jobs[i].plan_wots + 4 * k;
            up.push_back({jobs[i].key, w[0], w[1], w[2], w[3], 0, LAT_ALL_STEPS, 0});
        }
    }
    // the chains of the leaves the key's next signature will consume
    // (ends_ahead, one-signature launches), as rows after the planned ones,
    // into free slots (LAT_ENDS_OUT)
    const bool same_key = E.fam == fam && E.want_d == E.d;
    std::vector<std::pair<EndsAhead::Id, uint32_t>> fresh;
    const bool matched = ends && count == 1 && E.want_fam == fam && !E.want.empty() &&
                         seed_digest(n, {jobs[0].sk_seed, jobs[0].sk_prf, jobs[0].pub_seed}) == E.want_d;
    if (matched) {
        std::vector<bool> taken = busy;  // slots still read: this launch's L-trees, and what stays held
        if (same_key)
            for (const auto& kv : E.have) taken[kv.second] = true;
        uint32_t next_slot = 0;
        for (const EndsAhead::Id& id : E.want) {
            if (same_key && E.have.count(id)) continue;  // held already
            bool planned = false;                         // hashed whole here anyway
            for (uint32_t k = 0; k < jobs[0].nplan && !planned; ++k)
                planned = std::get<0>(id) == jobs[0].plan_layer[k] && std::get<1>(id) == jobs[0].plan_tree[k] &&
                          std::get<2>(id) == jobs[0].plan_leaf[k];
            if (planned) continue;
            while (next_slot < kEndsSlots && taken[next_slot]) ++next_slot;
            if (next_slot == kEndsSlots) break;
            taken[next_slot] = true;
            const uint64_t t = std::get<1>(id);
            ld.push_back({jobs[0].key, std::get<0>(id), (uint32_t)(t >> 32), (uint32_t)t, std::get<2>(id), 0,
                          LAT_ENDS_OUT, next_slot});
            fresh.emplace_back(id, next_slot);
        }
    }
    // H_msg inputs
    const size_t pb = pad + 3 * n;
    std::vector<DMsg> md(count);
    size_t total = 0;
    for (int i = 0; i < count; ++i) {
        md[i] = {jobs[i].key, 1u, jobs[i].idx, (uint64_t)total, (uint32_t)(pb + jobs[i].mlen), 0};
        total += (pb + jobs[i].mlen + 15) & ~(size_t)15;
    }
    total += kMsgSlack;
    // One pinned staging area and one device arena, so that the inputs go
    // over in one copy and the results come back in one:
    //   [descriptors, H_msg blob | sk arena (updated in place) | R, WOTS, tails]
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_ld = 0, o_md = o_ld + al(sizeof(DLeaf) * ld.size()), o_blob = o_md + al(sizeof(DMsg) * count),
                 o_wd = o_blob + al(total), o_jd = o_wd + al(sizeof(DLeaf) * (count + up.size())),
                 o_mg = o_jd + al(sizeof(DBds) * count), o_dag = o_mg + al(sizeof(DThash) * mg.size()),
                 o_lit = o_dag + al(sizeof(uint32_t) * dag.size()), o_sync = o_lit + al(lits.size()),
                 o_sk = o_sync + al(4 * sizeof(uint32_t)),
                 o_R = o_sk + al(sk_total), o_W = o_R + al((size_t)count * n),
                 o_T = o_W + al((size_t)count * len * n), io_bytes = o_T + al(tail_total);
    uint8_t* h = B.staging.need(io_bytes);
    uint8_t* dv = B.io.need(io_bytes);
    memcpy(h + o_ld, ld.data(), sizeof(DLeaf) * ld.size());
    memcpy(h + o_md, md.data(), sizeof(DMsg) * count);
    memset(h + o_blob, 0, total);
    memcpy(h + o_jd, jd.data(), sizeof(DBds) * count);
    memcpy(h + o_mg, mg.data(), sizeof(DThash) * mg.size());
    memcpy(h + o_dag, dag.data(), sizeof(uint32_t) * dag.size());
    memcpy(h + o_lit, lits.data(), lits.size());
    // BdsFinDag: leaves counted in, DAG (and state) done, values in, signature picked
    memset(h + o_sync, 0, 4 * sizeof(uint32_t));
    const size_t nl = ld.size(), nw = count + up.size();
    if (nl + nw + (mg.size() + count) / WG + 2 > 65535 || !use_lat(o, nl + nw, 0))
        fatal("bds_sign_chunk: batch too large for one launch sequence");
    DLeaf* wd = reinterpret_cast<DLeaf*>(h + o_wd);
    for (int i = 0; i < count; ++i) {
        uint8_t* p = h + o_blob + md[i].off;
        p[pad - 1] = 2;
        memcpy(p + pad + n, jobs[i].pub_root, n);
        uint64_t v = jobs[i].idx;
        for (int k = 0; k < 8; ++k) {
            p[pad + 3 * n - 1 - k] = (uint8_t)(v & 0xff);
            v >>= 8;
        }
        if (jobs[i].mlen) memcpy(p + pb, jobs[i].m, jobs[i].mlen);
        const unsigned th = jobs[i].shape.th;
        const uint64_t tree = jobs[i].idx >> th;
        wd[i] = {jobs[i].key, 0, (uint32_t)(tree >> 32), (uint32_t)tree,
                 (uint32_t)(jobs[i].idx & ((1ull << th) - 1)), 0, LAT_ALL_STEPS, 0};
        if (!jobs[i].dsk) memcpy(h + o_sk + jd[i].sk_off, jobs[i].sk, jobs[i].shape.sk_bytes);
    }
    if (!up.empty()) memcpy(wd + count, up.data(), sizeof(DLeaf) * up.size());
    const DLeaf* dld = reinterpret_cast<const DLeaf*>(dv + o_ld);
    const DMsg* dmd = reinterpret_cast<const DMsg*>(dv + o_md);
    uint8_t* db = dv + o_blob;
    const DLeaf* dw = reinterpret_cast<const DLeaf*>(dv + o_wd);
    const DBds* djobs = reinterpret_cast<const DBds*>(dv + o_jd);
    const DThash* dmg = reinterpret_cast<const DThash*>(dv + o_mg);
    const uint32_t* ddag = reinterpret_cast<const uint32_t*>(dv + o_dag);
    const uint8_t* dlit = dv + o_lit;
    uint8_t* ska = dv + o_sk;
    uint8_t* dR = dv + o_R;
    uint8_t* dW = dv + o_W;
    uint8_t* dT = dv + o_T;
    uint8_t* leaf_arena = B.leaf_arena.need(ld.size() * n);
    void* merge_arena = B.merge_arena.need(std::max<size_t>(1, mg.size()) * o->hpre_bytes());
    I->h2d(dv, h, o_R);
    // three launches on the main stream: R / H_msg with the address-only work
    // of the planned leaves, the layer-0 WOTS signatures (all steps: the
    // digests are computed in the same launch) and the merges, then the
    // leaves' and signatures' finishing work, then the state update (which a
    // lone signature's finishing launch usually runs itself: BdsFinDag)
    // scratch rows: the planned leaves, the layer-0 WOTS signatures, then the
    // upper-layer WOTS signatures of the state updates
    uint8_t* lat = B.lat_leaf.need((nl + nw) * o->lat_bytes(0));
    uint8_t* dig = B.dig.need((size_t)count * n);
    // A message longer than this makes H_msg (a serial chain of compressions)
    // longer than the planned leaves it can overlap: then H_msg runs beside
    // the leaves and the WOTS signatures beside the state updates (measured:
    // 1 KB messages 216 -> 178 us, 4 KB 338 -> 303 us per XMSS-SHA2_16_256
    // signature; 32 B messages stay at 147 us with H_msg first).
    uint64_t mmax = 0;
    for (int i = 0; i < count; ++i) mmax = std::max<uint64_t>(mmax, jobs[i].mlen);
    const bool long_msgs = mmax > kBdsShortMsg;
    // one signature of a long message launched alone (sig_values): its WOTS
    // chains are walked to every value beside H_msg, and the signature is
    // picked once the digest is there (the chains leave the serial path
    // behind H_msg; a group of several keeps the plain walk)
    uint8_t* vals = long_msgs && count == 1 && jobs[0].sig_values ? B.vals.need((size_t)len * 16 * n) : nullptr;
    // ... and its first H_msg blocks are hashed in the pre-image launch
    // (HmsgMid: the SHA-256 families)
    void* mid = vals ? B.mid.need(sizeof(HmsgMid)) : nullptr;
    // one-call signature with a merge DAG: the DAG is hashed in the
    // finishing launch as soon as the planned leaves are in, beside H_msg
    // and the WOTS signature, instead of at the start of the update
    // (and, when its LDS fits the finishing launch's, the state update after
    // it, so that the update launch only picks a long message's signature)
    BdsFinDag fd{};
    const bool fin_state = lds <= kFinStateLds;
    if (count == 1 && jobs[0].sig_values && (jd[0].maxlev || fin_state) && dag_spin()) {
        fd.sync = reinterpret_cast<uint32_t*>(dv + o_sync);
        fd.on = 1;
        fd.wait = jobs[0].nplan;
        fd.spin = dag_spin();
        fd.nmerge = jd[0].nmerge;
        fd.maxlev = jd[0].maxlev;
        fd.q0 = ddag;  // job 0's merges and literals start at offset 0
        fd.lits = dlit;
        fd.pre = merge_arena;
        fd.vals = jd[0].maxlev && !fin_state ? B.dagv.need((size_t)jd[0].nmerge * n) : nullptr;
        fd.state = fin_state;
        fd.jobs = djobs;
        fd.sk_arena = ska;
        fd.tails = dT;
        fd.wpre = lat + (nl + count) * o->lat_bytes(0);
        fd.wstride = o->lat_bytes(0);
        fd.select = fin_state && vals;
    }
    o->bds_lat(I->L, dk, dmd, (uint32_t)count, pad, db, dR, dig, dld, (uint32_t)nl, dw, (uint32_t)nw, lat, dmg,
               (uint32_t)mg.size(), merge_arena, leaf_arena, dW, long_msgs, ends, vals, fd, mid);
    auto update = [&] {
        o->bds_update(I->L, dk, lds, djobs, (uint32_t)count, ska, leaf_arena, merge_arena, ddag, dlit,
                      lat + (nl + count) * o->lat_bytes(0), dT, B.scratch.need((size_t)count * 4 * n),
                      long_msgs ? dw : nullptr, dig, lat + nl * o->lat_bytes(0), dW, vals, fd.sync, fd.vals);
    };
    // (fd.sync, null when the finishing launch has no merge-DAG block, tells
    // the update whether job 0's merges or whole state update were done there)
    if (fd.on && fd.state) {
        // the finishing launch is to do all of it: the update launch only
        // if one of its bounded waits gave up (the results then come back
        // a second time)
        I->d2h(h + o_sync, dv + o_sync, io_bytes - o_sync);
        I->sync();
        const uint32_t* st = reinterpret_cast<const uint32_t*>(h + o_sync);
        if (st[1] != 2 || (fd.select && st[3] != 1)) {
            update();
            I->d2h(h + o_sk, ska, io_bytes - o_sk);
            I->sync();
        }
    } else {
        update();
        I->d2h(h + o_sk, ska, io_bytes - o_sk);
        I->sync();
    }
    for (int i = 0; i < count; ++i) {
        memcpy(jobs[i].sk, h + o_sk + jd[i].sk_off, jobs[i].shape.sk_bytes);
        memcpy(jobs[i].R_out, h + o_R + (size_t)i * n, n);
        memcpy(jobs[i].wots0_out, h + o_W + (size_t)i * len * n, (size_t)len * n);
        memcpy(jobs[i].tail_out, h + o_T + jd[i].tail_off, tail_bytes[i]);
    }
    // the slots now: the chain ends this launch used are spent, the fresh
    // ones are held for the requested key (its earlier ones too, if it is
    // the key the slots held)
    for (const EndsAhead::Id& id : used) E.have.erase(id);
    if (matched) {
        if (!same_key) E.have.clear();
        E.fam = fam;
        E.d = E.want_d;
        for (const auto& f : fresh) E.have[f.first] = f.second;
    }
    E.want.clear();  // a request is for the launch right after it
}


struct VerifyBufs {
    DevBuf io, dig, roots, scratch, spec;
    PinnedBuf staging;
};

void Engine::verify_many(int fam, int count, const VerifyJob* jobs, int* ok) {
    // chunks keep every launch's grid within its limits (the pre-image launch
    // has one row per (request, layer) plus one per 64 messages)
    unsigned dall = 1;
    for (int j = 0; j < count; ++j) dall = std::max(dall, jobs[j].d);
    const int kChunk = std::min(8192, (int)(65000 / (dall + 1)));
    if (count > kChunk) {
        for (int i = 0; i < count; i += kChunk) verify_many(fam, std::min(kChunk, count - i), jobs + i, ok + i);
        return;
    }
    Impl* I = impl_;
    if (!I->ver) I->ver = new VerifyBufs;
    VerifyBufs& V = *I->ver;
    FamOps* o = ops_of(fam);
    const unsigned n = fam_n(fam), len = fam_len(fam), pad = fam_pad(fam);
    keys_for(this, I, fam);
    const void* dk = I->kp();
    // deepest first: the requests that still have a layer L are a prefix
    std::vector<int> ord(count);
    for (int i = 0; i < count; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return jobs[a].d > jobs[b].d; });
    const unsigned dmax = count ? jobs[ord[0]].d : 0;
    if (dmax > 16) fatal("verify_many: more than 16 layers");
    std::vector<uint32_t> cnt(dmax, 0), base(dmax + 1, 0);
    unsigned thmax = 0;
    for (int j = 0; j < count; ++j) {
        for (unsigned L = 0; L < jobs[ord[j]].d; ++L) ++cnt[L];
        thmax = std::max(thmax, jobs[ord[j]].th);
    }
    for (unsigned L = 0; L < dmax; ++L) base[L + 1] = base[L] + cnt[L];
    const uint32_t items = base[dmax];
    // auth offsets (in nodes) from each layer's first auth path
    std::vector<uint32_t> auth_at(dmax + 1, 0);
    for (unsigned L = 0; L < dmax; ++L) {
        uint32_t s = 0;
        for (uint32_t j = 0; j < cnt[L]; ++j) s += jobs[ord[j]].th;
        auth_at[L + 1] = auth_at[L] + s;
    }
    const size_t pb = pad + 3 * n;
    std::vector<DMsg> md(count);
    size_t total = 0;
    for (int j = 0; j < count; ++j) {
        const VerifyJob& v = jobs[ord[j]];
        uint64_t idx = 0;
        for (unsigned b = 0; b < v.index_bytes; ++b) idx = (idx << 8) | v.sig[b];
        md[j] = {v.key, 0u, idx, (uint64_t)total, (uint32_t)(pb + v.mlen), 0};
        total += (pb + v.mlen + 15) & ~(size_t)15;
    }
    total += kMsgSlack;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_md = 0, o_blob = o_md + al(sizeof(DMsg) * count), o_it = o_blob + al(total),
                 o_sig = o_it + al(sizeof(DLeaf) * items), o_auth = o_sig + al((size_t)items * len * n),
                 io_bytes = o_auth + al((size_t)auth_at[dmax] * n);
    uint8_t* h = V.staging.need(std::max(io_bytes, (size_t)items * n));
    uint8_t* dv = V.io.need(io_bytes);
    memcpy(h + o_md, md.data(), sizeof(DMsg) * count);
    memset(h + o_blob, 0, total);
    DLeaf* it = reinterpret_cast<DLeaf*>(h + o_it);
    std::vector<uint32_t> fill(dmax, 0), afill(dmax, 0);
    for (int j = 0; j < count; ++j) {
        const VerifyJob& v = jobs[ord[j]];
        const uint8_t* R = v.sig + v.index_bytes;
        uint8_t* p = h + o_blob + md[j].off;
        p[pad - 1] = 2;  // toByte(2, pad) || R || root || toByte(idx, n) || M (hash.c:107-120)
        memcpy(p + pad, R, n);
        memcpy(p + pad + n, v.pk_root, n);
        uint64_t x = md[j].idx;
        for (int b = 0; b < 8; ++b) {
            p[pad + 3 * n - 1 - b] = (uint8_t)(x & 0xff);
            x >>= 8;
        }
        if (v.mlen) memcpy(p + pb, v.m, v.mlen);
        uint64_t idx = md[j].idx;
        const uint8_t* s = R + n;
        for (unsigned L = 0; L < v.d; ++L) {
            const uint32_t leaf = (uint32_t)(idx & ((1ull << v.th) - 1));
            idx >>= v.th;
            const uint32_t t = base[L] + fill[L]++;
            it[t] = {v.key, L, (uint32_t)(idx >> 32), (uint32_t)idx, leaf, v.th, LAT_ALL_STEPS,
                     auth_at[L] + afill[L]};
            memcpy(h + o_sig + (size_t)t * len * n, s, (size_t)len * n);
            s += (size_t)len * n;
            memcpy(h + o_auth + (size_t)(auth_at[L] + afill[L]) * n, s, (size_t)v.th * n);
            afill[L] += v.th;
            s += (size_t)v.th * n;
        }
    }
    I->h2d(dv, h, io_bytes);
    const DMsg* dmd = reinterpret_cast<const DMsg*>(dv + o_md);
    const DLeaf* dit = reinterpret_cast<const DLeaf*>(dv + o_it);
    uint8_t* roots = V.roots.need((size_t)items * n);
    uint8_t* dig = V.dig.need((size_t)count * n);
    uint8_t* scratch = V.scratch.need((size_t)items * o->lat_bytes(thmax));
    // two launches on one stream: H_msg with the address-only work of every
    // layer, then every layer of each request in one workgroup.  For few
    // XMSS^MT requests, three: the second walks layer 0 beside the
    // speculative chains of the upper layers (every start digit), the third
    // the upper layers, each without its chains.
    o->verify_pre(I->L, dk, dmd, (uint32_t)count, pad, dv + o_blob, dig, dit, items, thmax, scratch);
    if (dmax > 1 && count <= kSpecVerify) {
        const uint32_t nspec = (uint32_t)(items - base[1]);
        uint8_t* spec = V.spec.need((size_t)nspec * len * 16 * n);
        o->verify_layers(I->L, dit, (uint32_t)count, base.data(), dmax, 0, 1, dig, dv + o_sig, dv + o_auth, thmax,
                         scratch, roots, spec, nspec);
        o->verify_layers(I->L, dit, (uint32_t)count, base.data(), dmax, 1, dmax, dig, dv + o_sig, dv + o_auth,
                         thmax, scratch, roots, spec, 0);
    } else {
        o->verify_layers(I->L, dit, (uint32_t)count, base.data(), dmax, 0, dmax, dig, dv + o_sig, dv + o_auth,
                         thmax, scratch, roots, nullptr, 0);
    }
    I->d2h(h, roots, (size_t)items * n);
    I->sync();
    for (int j = 0; j < count; ++j) {
        const VerifyJob& v = jobs[ord[j]];
        ok[ord[j]] = !memcmp(h + (size_t)(base[v.d - 1] + j) * n, v.pk_root, n);
    }
}

// Layout of one verification in the static scratch (256-byte aligned parts).
struct StaticLayout {
    size_t keys, seed, md, it, blob, sig, auth, roots, dig, lat, spec, total;
};
static StaticLayout static_layout(FamOps* o, int fam, unsigned th, unsigned d, uint64_t mlen) {
    const unsigned n = fam_n(fam), len = fam_len(fam), pad = fam_pad(fam);
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    StaticLayout t;
    t.keys = 0;
    t.seed = al(o->key_bytes());
    t.md = t.seed + al(3 * n);
    t.it = t.md + al(sizeof(DMsg));
    t.blob = t.it + al(sizeof(DLeaf) * d);
    t.sig = t.blob + al(((pad + 3 * n + mlen + 15) & ~(uint64_t)15) + kMsgSlack);
    t.auth = t.sig + al((size_t)d * len * n);
    t.roots = t.auth + al((size_t)d * th * n);
    t.dig = t.roots + al((size_t)d * n);
    t.lat = t.dig + al(n);
    t.spec = t.lat + al((size_t)d * o->lat_bytes(th));
    t.total = t.spec + (d > 1 ? al((size_t)(d - 1) * len * 16 * n) : 0);
    return t;
}

size_t Engine::static_capacity() { return kStaticScratch; }

size_t Engine::static_bytes(int fam, unsigned th, unsigned d, uint64_t mlen) {
    return static_layout(ops_of(fam), fam, th, d, mlen).total;
}

// One verification (the verify_many sequence for a single request) with all
// of its device memory in the static scratch: the inputs are copied from the
// caller's (pageable) buffers, the key context is prepared into the scratch
// instead of the key pool, and no allocation is made.  The blob region is
// zeroed first by a kernel (the H_msg padding reads past the message).
int Engine::verify_static(int fam, const VerifyJob& v, int* ok) {
    Impl* I = impl_;
    FamOps* o = ops_of(fam);
    const unsigned n = fam_n(fam), len = fam_len(fam), pad = fam_pad(fam);
    const unsigned d = v.d, th = v.th;
    const StaticLayout t = static_layout(o, fam, th, d, v.mlen);
    if (t.total > kStaticScratch || d > 16) return -4;
    uint8_t* S = (uint8_t*)I->static_scratch;
    if (!S) fatal("verify_static: no static scratch on this device");
    I->L.s = I->L.s_main;
    const size_t pb = pad + 3 * n;
    zero_async(I->L, S + t.blob, t.sig - t.blob);
    // one header copy: PUB_SEED, the H_msg descriptor, the layer items and
    // the H_msg prefix (toByte(2, pad) || R || root || toByte(idx, n))
    uint8_t hdr[4096];
    const size_t hb = t.blob + pb - t.seed;
    if (hb > sizeof hdr) fatal("verify_static: header too large");
    memset(hdr, 0, hb);
    auto at = [&](size_t off) { return hdr + (off - t.seed); };  // the header byte of a scratch offset
    memcpy(at(t.seed) + 2 * n, v.pk_root + n, n);  // PUB_SEED: pk_root points at root || PUB_SEED
    uint64_t idx = 0;
    for (unsigned b = 0; b < v.index_bytes; ++b) idx = (idx << 8) | v.sig[b];
    const DMsg md{0u, 0u, idx, 0u, (uint32_t)(pb + v.mlen), 0u};
    memcpy(at(t.md), &md, sizeof md);
    uint64_t x = idx;
    for (unsigned L = 0; L < d; ++L) {
        const uint32_t leaf = (uint32_t)(x & ((1ull << th) - 1));
        x >>= th;
        const DLeaf q{0u, L, (uint32_t)(x >> 32), (uint32_t)x, leaf, th, LAT_ALL_STEPS, L * th};
        memcpy(at(t.it) + (size_t)L * sizeof(DLeaf), &q, sizeof q);
    }
    uint8_t* pre = at(t.blob);
    pre[pad - 1] = 2;
    memcpy(pre + pad, v.sig + v.index_bytes, n);
    memcpy(pre + pad + n, v.pk_root, n);
    x = idx;
    for (int b = 0; b < 8; ++b) {
        pre[pad + 3 * n - 1 - b] = (uint8_t)(x & 0xff);
        x >>= 8;
    }
    I->h2d(S + t.seed, hdr, hb);
    I->h2d(S + t.blob + pb, v.m, v.mlen);
    // WOTS signatures and auth paths of the d layers, straight out of sm
    const uint8_t* l0 = v.sig + v.index_bytes + n;
    const size_t lb = (size_t)(len + th) * n;
    HIPCK(hipMemcpy2DAsync(S + t.sig, (size_t)len * n, l0, lb, (size_t)len * n, d, hipMemcpyHostToDevice, I->L.s));
    if (th)
        HIPCK(hipMemcpy2DAsync(S + t.auth, (size_t)th * n, l0 + (size_t)len * n, lb, (size_t)th * n, d,
                               hipMemcpyHostToDevice, I->L.s));
    o->prepare(I->L, S + t.seed, 1, S + t.keys);
    uint32_t base[17] = {0};
    for (unsigned L = 0; L <= d; ++L) base[L] = L;
    o->verify_pre(I->L, S + t.keys, (const DMsg*)(S + t.md), 1, pad, S + t.blob, S + t.dig,
                  (const DLeaf*)(S + t.it), d, th, S + t.lat);
    const DLeaf* it = (const DLeaf*)(S + t.it);
    if (d > 1) {
        o->verify_layers(I->L, it, 1, base, d, 0, 1, S + t.dig, S + t.sig, S + t.auth, th, S + t.lat, S + t.roots,
                         S + t.spec, d - 1);
        o->verify_layers(I->L, it, 1, base, d, 1, d, S + t.dig, S + t.sig, S + t.auth, th, S + t.lat, S + t.roots,
                         S + t.spec, 0);
    } else {
        o->verify_layers(I->L, it, 1, base, d, 0, d, S + t.dig, S + t.sig, S + t.auth, th, S + t.lat, S + t.roots,
                         nullptr, 0);
    }
    uint8_t root[64];
    I->d2h(root, S + t.roots + (size_t)(d - 1) * n, n);
    I->sync();
    *ok = !memcmp(root, v.pk_root, n);
    return 0;
}

void Engine::release_workspace() {
    Impl* I = impl_;
    DevBuf* bufs[] = {&I->keys, &I->seeds, &I->desc, &I->ends, &I->tmp, &I->nodes, &I->a,
                      &I->b,    &I->c,     &I->d,    &I->idx,  &I->lat, &I->tpre};
    for (DevBuf* x : bufs) x->release();
    if (I->bds) {
        DevBuf* b[] = {&I->bds->io, &I->bds->leaf_arena, &I->bds->dig, &I->bds->scratch, &I->bds->lat_leaf,
                       &I->bds->merge_arena, &I->bds->vals, &I->bds->dagv, &I->bds->mid};
        for (DevBuf* x : b) x->release();
    }
    I->ahead.clear();
    I->ahead.at = nullptr;
    I->ahead.buf.release();
    if (I->ver) {
        DevBuf* v[] = {&I->ver->io, &I->ver->dig, &I->ver->roots, &I->ver->scratch, &I->ver->spec};
        for (DevBuf* x : v) x->release();
    }
    // the keys of set_keys are gone, and a caller's table is not kept
    I->kp_override = nullptr;
    I->external = false;
    I->key_fam = -1;
}

void Engine::trim() {
    Impl* I = impl_;
    release_workspace();
    for (Batch*& b : I->spare) {
        if (b) destroy_batch(b);  // a spare signing batch was scrubbed when it was freed
        b = nullptr;
    }
    for (int f = 0; f < F_COUNT; ++f) {
        KeyPool& P = I->pool[f];
        if (!P.p) continue;
        clear_device(P.p, P.cap * ops_of(f)->key_bytes());
        dev_free(P.p, P.cap * ops_of(f)->key_bytes());
        P = KeyPool{};
    }
}

// A request's share of a launch group's buffers: its own parts, plus a
// fixed allowance for the group's 256-byte region rounding, the message
// slack and the minimum sizes (at most 8 KB per group, so one per request
// covers it).
static constexpr size_t kGroupAllowance = 8192;

size_t Engine::bds_bytes(int fam, const BdsJob& j) {
    FamOps* o = ops_of(fam);
    const size_t n = fam_n(fam), len = fam_len(fam), pb = fam_pad(fam) + 3 * n;
    const BdsShape& s = j.shape;
    const size_t nw = 1 + (j.plan_wots ? j.nwots : 0);
    const size_t io = sizeof(DLeaf) * j.nplan + sizeof(DMsg) + ((pb + j.mlen + 15) & ~(size_t)15) +
                      sizeof(DLeaf) * nw + sizeof(DBds) + sizeof(DThash) * j.nmerge + 12 * (size_t)j.nmerge +
                      n * j.nlit + ((s.sk_bytes + 7) & ~7ull) + n + len * n +
                      ((size_t)s.sig_bytes - s.index_bytes - n - s.wots_sig_bytes);
    // (a group of one signature also walks its WOTS chains to every value,
    // keeps the merge values and H_msg's chaining value: bds_sign_chunk's
    // vals, dagv and mid)
    const size_t one = (size_t)len * 16 * n + (size_t)j.nmerge * n + sizeof(HmsgMid);
    return io + j.nplan * n + (size_t)j.nmerge * o->hpre_bytes() + (j.nplan + nw) * o->lat_bytes(0) + 5 * n + one +
           kGroupAllowance;
}

size_t Engine::verify_bytes(int fam, unsigned th, unsigned d, uint64_t mlen) {
    FamOps* o = ops_of(fam);
    const size_t n = fam_n(fam), len = fam_len(fam), pb = fam_pad(fam) + 3 * n;
    const size_t io = sizeof(DMsg) + ((pb + mlen + 15) & ~(uint64_t)15) + sizeof(DLeaf) * d + (size_t)d * len * n +
                      (size_t)d * th * n;
    const size_t spec = d > 1 ? (size_t)(d - 1) * len * 16 * n : 0;
    return io + (size_t)d * n + n + (size_t)d * o->lat_bytes(th) + spec + kGroupAllowance;
}

int Engine::pool_slot(int fam, const KeySeeds& s) {
    Impl* I = impl_;
    KeyPool& P = I->pool[fam];
    const Digest d = seed_digest(fam_n(fam), s);
    auto it = P.by_digest.find(d);
    if (it != P.by_digest.end()) {
        P.slot[it->second].used = I->epoch;
        return it->second;
    }
    if (P.cached >= kPoolCached) {  // evict the least recently used cached slot of an earlier call
        int victim = -1;
        for (size_t k = 0; k < P.slot.size(); ++k) {
            const KeyPool::Slot& S = P.slot[k];
            if (S.live && !S.pinned && S.used < I->epoch && (victim < 0 || S.used < P.slot[victim].used))
                victim = (int)k;
        }
        if (victim >= 0) pool_drop(I, fam, victim);
    }
    const int slot = pool_take(I, fam, s, false, d);
    P.by_digest[d] = slot;
    return slot;
}

}  // namespace xm
