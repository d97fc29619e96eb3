// serve.cpp -- a serving loop with continuous batching (include/xmss_mi355x.h,
// xmss_mi355x_server_*).
//
// Requests (BDS signatures, verifications) arrive at any time.  Each runs on
// a worker thread as ordinary BDS code (bds_sign, core_bds.cpp) over a
// backend whose hashing calls are handed to one dispatcher: the dispatcher
// waits until every running request is blocked on the device (or a short
// deadline passes), then executes all waiting calls as one launch per kind
// and hash family -- the leaves, merges, WOTS signatures, H_msg and
// verifications of many keys share launches, and a request that arrives while
// others are on the GPU joins the next launch.  Signatures of one key run in
// arrival order (its sk is a sequential state); other keys are independent.
//
// Device memory: the server holds exactly its HBM budget, as one arena
// (DeviceArena) allocated at creation, and nothing else.  In it live the key
// contexts (the PRF midstates of SK_SEED, SK_PRF, PUB_SEED: one table per
// hash family, indexed by the key's slot), each key's serialized sk (index,
// seeds, BDS states, XMSS^MT WOTS signatures), which the state-update kernel
// advances in place, and the buffers of every launch group: the lanes' engines
// allocate them inside the arena, exactly sized, while they execute a group,
// and free them after it.  The server keeps a byte-identical host mirror of
// every sk (each signature's result copy brings the updated bytes back): the
// host plans the next signature's leaves from it, serves
// xmss_mi355x_server_get_sk from it, and runs the last-index / exhausted path
// on it (then uploads it again).
// Admission: adding a key reserves its resident bytes; every request reserves
// the device bytes its launches will allocate (its share of a launch group,
// from its signature's leaf and merge plan or its verification's shape, by the
// engine's own sizing) and waits in FIFO order until the budget left by the
// keys and the running requests holds it.  A request larger than the budget
// left by the keys returns -3.
#include <string.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <thread>
#include <vector>

#include "../../include/xmss_core.h"
#include "../../include/xmss_mi355x.h"
#include "bds.h"
#include "core_common.h"

using namespace xm;

namespace {

using Clock = std::chrono::steady_clock;
// A lane waits at most this long for running requests to queue their device
// calls before it launches what it has.
constexpr std::chrono::microseconds kBatchWindow(200);
// The largest subtree height of any parameter set (XMSS-*_20_*).
constexpr unsigned kMaxTreeHeight = 20;
// Device bytes of one key context's preparation (its seeds staged in a 256-byte granule).
constexpr size_t kPrepareBytes = 4096;

// Key contexts of the server's keys: one table per hash family in the arena.
// Lanes hold `mu` shared while a launch group runs; a table that grows is
// moved under it held exclusively.
struct KeyTables {
    std::shared_mutex mu;
    struct T {
        void* p = nullptr;
        size_t cap = 0;  // slots
        int n = 0;
        int zero = -1;  // slot of the all-zero seeds (the wiped key of a last-index signature)
    } t[F_COUNT];
};

// ---------------------------------------------------------------- dispatcher
struct Call {
    enum Kind { LEAVES, THASH, WOTS, HMSG, VERIFY, BDS, PREPARE } kind;
    Engine::BdsJob* job = nullptr;  // BDS
    int fam = 0;
    int slot = 0;                   // key table slot
    unsigned n = 0;
    // LEAVES
    std::vector<LeafId> ids;
    // THASH
    uint32_t adrs[8];
    const unsigned char* in = nullptr;
    // WOTS
    uint32_t layer = 0;
    uint64_t tree = 0;
    uint32_t ots = 0;
    const unsigned char* msg = nullptr;
    // HMSG
    uint64_t idx = 0;
    const unsigned char* root = nullptr;
    const unsigned char* m = nullptr;
    uint64_t mlen = 0;
    unsigned char* out2 = nullptr;
    // VERIFY
    const xmss_params* p = nullptr;
    const unsigned char* pk = nullptr;  // OID-less
    const unsigned char* sm = nullptr;
    uint64_t smlen = 0;
    int ok = 0;
    // PREPARE: the key context of `seeds` into table slot `slot`
    KeySeeds seeds{};
    // device bytes the call's launches take (its request's reservation)
    size_t bytes = 0;
    // result
    unsigned char* out = nullptr;
    bool done = false;
    bool failed = false;  // the lane's device work failed (DeviceError)
    std::condition_variable cv;  // the caller waits here (no wake-ups of unrelated callers)
};

class Dispatcher {
public:
    // Lanes, each with its own thread and private engine (HIP streams): lane 0
    // takes every kind of call, the other signing lanes BDS signatures, the
    // verification lanes verifications.  A lane launches what its queue holds
    // as soon as it is idle and every running request is waiting on the device
    // (queued or executing anywhere), or 200 us after the oldest queued call --
    // so a request that arrives while a group is on the GPU starts in the next
    // idle lane instead of waiting for that group.  Every group runs with its
    // buffers in the arena and the keys of the key tables.
    Dispatcher(Engine* first, const std::vector<Engine*>& sign, const std::vector<Engine*>& verify,
               DeviceArena* arena, KeyTables* kt)
        : arena_(arena), kt_(kt) {
        lanes_.emplace_back(new Lane{first, (1u << Q_OTHER) | (1u << Q_SIGN)});
        for (Engine* e : sign) lanes_.emplace_back(new Lane{e, 1u << Q_SIGN});
        for (Engine* e : verify) lanes_.emplace_back(new Lane{e, 1u << Q_VERIFY});
        if (verify.empty()) lanes_[0]->queues |= 1u << Q_VERIFY;
        for (auto& l : lanes_) l->th = std::thread([this, p = l.get()] { loop(*p); });
    }
    ~Dispatcher() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& l : lanes_) {
            l->th.join();
            Engine::release_private(l->engine);
        }
    }
    void enter() {
        std::lock_guard<std::mutex> lk(mu_);
        ++active_;
    }
    void leave() {
        std::lock_guard<std::mutex> lk(mu_);
        --active_;
        cv_.notify_all();
    }
    // Blocks until a lane has executed the call.
    void run(Call* c) {
        std::unique_lock<std::mutex> lk(mu_);
        const int q = c->kind == Call::VERIFY ? Q_VERIFY : c->kind == Call::BDS ? Q_SIGN : Q_OTHER;
        if (queue_[q].empty()) first_[q] = Clock::now();
        queue_[q].push_back(c);
        ++waiting_;
        cv_.notify_all();
        c->cv.wait(lk, [c] { return c->done; });
    }
    uint64_t launches() const { return launches_; }
    uint64_t calls() const { return calls_; }

private:
    enum { Q_SIGN = 0, Q_VERIFY = 1, Q_OTHER = 2, NQ = 3 };
    struct Lane {
        Lane(Engine* e, uint32_t q) : engine(e), queues(q) {}
        Engine* engine;
        uint32_t queues;  // bit mask of the queues this lane serves
        std::thread th;
    };

    // The queue this lane should launch now, or -1.
    int ready(const Lane& l) const {
        for (int q = 0; q < NQ; ++q) {
            if (!(l.queues & (1u << q)) || queue_[q].empty()) continue;
            if (waiting_ >= active_ || Clock::now() - first_[q] >= kBatchWindow) return q;
        }
        return -1;
    }

    void loop(Lane& l) {
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            int q = -1;
            cv_.wait_for(lk, kBatchWindow, [&] { return stop_ || (q = ready(l)) >= 0; });
            if (q < 0) q = ready(l);
            if (q < 0) {
                if (stop_) return;
                continue;
            }
            std::vector<Call*> batch;
            batch.swap(queue_[q]);
            lk.unlock();
            // one contiguous block of the arena for everything this launch
            // takes: the sum of its calls' reservations (admission has left
            // that much of the budget free; while another lane's block splits
            // the free space, wait for it to come back)
            size_t need = 0;
            for (Call* c : batch) need += c->bytes;
            void* blk = nullptr;
            {
                std::unique_lock<std::mutex> ak(blk_mu_);
                while (!(blk = arena_->alloc(need)) && blocks_ > 0) blk_cv_.wait(ak);
                if (blk) ++blocks_;
            }
            if (!blk) {
                for (Call* c : batch) c->failed = true;
            } else {
                DeviceArena block(blk, (need + 255) & ~(size_t)255);
                ArenaScope scope(&block);
                try {
                    execute_on(l.engine, batch);
                } catch (const DeviceError&) {
                    // every call of the group fails; the lane's engine is made
                    // consistent again for the next group
                    for (Call* c : batch) c->failed = true;
                    try {
                        l.engine->recover();
                    } catch (...) {
                    }
                }
                try {
                    l.engine->release_workspace();  // nothing held between groups
                } catch (...) {
                }
                std::lock_guard<std::mutex> ak(blk_mu_);
                arena_->release(blk);
                --blocks_;
                blk_cv_.notify_all();
            }
            lk.lock();
            for (Call* c : batch) {
                c->done = true;
                c->cv.notify_one();
            }
            waiting_ -= (int)batch.size();
            calls_ += batch.size();
            cv_.notify_all();
        }
    }

    void execute_on(Engine* e_, std::vector<Call*>& batch) {
        std::shared_lock<std::shared_mutex> tl(kt_->mu);
        // (fam, kind): verifications of every shape of a family share a launch
        // sequence (their reservations count the scratch stride of the
        // largest tree height)
        std::map<std::pair<int, int>, std::vector<Call*>> groups;
        for (Call* c : batch) groups[{c->fam, (int)c->kind}].push_back(c);
        for (auto& g : groups) {
            const int fam = g.first.first;
            std::vector<Call*>& cs = g.second;
            const unsigned n = fam_n(fam);
            e_->use_keys(fam, kt_->t[fam].p);
            switch ((Call::Kind)g.first.second) {
                case Call::PREPARE: {
                    for (Call* c : cs) e_->prepare_keys(fam, &c->seeds, 1, kt_->t[fam].p, c->slot);
                    break;
                }
                case Call::LEAVES: {
                    std::vector<LeafRun> runs;
                    for (Call* c : cs)
                        for (const LeafId& id : c->ids) runs.push_back({(uint32_t)c->slot, id.layer, id.tree, id.leaf});
                    std::vector<unsigned char> out(runs.size() * n);
                    if (!runs.empty()) e_->leaves(fam, (int)runs.size(), runs.data(), 1, out.data());
                    size_t k = 0;
                    for (Call* c : cs) {
                        memcpy(c->out, &out[k * n], c->ids.size() * n);
                        k += c->ids.size();
                    }
                    break;
                }
                case Call::THASH: {
                    std::vector<ThashReq> q(cs.size());
                    std::vector<unsigned char> in(cs.size() * 2 * n), out(cs.size() * n);
                    for (size_t i = 0; i < cs.size(); ++i) {
                        q[i].key = (uint32_t)cs[i]->slot;
                        memcpy(q[i].adrs, cs[i]->adrs, 32);
                        memcpy(&in[i * 2 * n], cs[i]->in, 2 * n);
                    }
                    e_->thash_h(fam, (int)cs.size(), q.data(), in.data(), out.data());
                    for (size_t i = 0; i < cs.size(); ++i) memcpy(cs[i]->out, &out[i * n], n);
                    break;
                }
                case Call::WOTS: {
                    const size_t sb = (size_t)fam_len(fam) * n;
                    std::vector<WotsReq> q(cs.size());
                    std::vector<unsigned char> msgs(cs.size() * n), out(cs.size() * sb);
                    for (size_t i = 0; i < cs.size(); ++i) {
                        q[i] = {(uint32_t)cs[i]->slot, cs[i]->layer, cs[i]->tree, cs[i]->ots, WOTS_SIGN};
                        memcpy(&msgs[i * n], cs[i]->msg, n);
                    }
                    e_->wots(fam, (int)cs.size(), q.data(), msgs.data(), nullptr, out.data());
                    for (size_t i = 0; i < cs.size(); ++i) memcpy(cs[i]->out, &out[i * sb], sb);
                    break;
                }
                case Call::HMSG: {
                    std::vector<MsgReq> q(cs.size());
                    std::vector<unsigned char> R(cs.size() * n), D(cs.size() * n);
                    for (size_t i = 0; i < cs.size(); ++i)
                        q[i] = {(uint32_t)cs[i]->slot, 1, cs[i]->idx, nullptr, cs[i]->root, cs[i]->m, cs[i]->mlen};
                    e_->hmsg(fam, (int)cs.size(), q.data(), R.data(), D.data());
                    for (size_t i = 0; i < cs.size(); ++i) {
                        memcpy(cs[i]->out, &R[i * n], n);
                        memcpy(cs[i]->out2, &D[i * n], n);
                    }
                    break;
                }
                case Call::BDS: {
                    // signatures with the traversal on the device: one batch for all
                    // shapes of the family, key contexts from the key table
                    std::vector<Engine::BdsJob> jobs;
                    for (Call* c : cs) jobs.push_back(*c->job);
                    // a signature alone takes the one-call path's schedule (its
                    // reservation counts the memory that takes)
                    if (jobs.size() == 1) jobs[0].sig_values = true;
                    e_->bds_sign_batch(fam, (int)jobs.size(), jobs.data());
                    break;
                }
                case Call::VERIFY: {
                    // every key of one shape in one launch sequence (PUB_SEED's
                    // context from the key's table slot)
                    std::vector<Engine::VerifyJob> jobs;
                    for (Call* c : cs) {
                        const xmss_params* p = c->p;
                        jobs.push_back({(uint32_t)c->slot, p->tree_height, p->d, p->index_bytes, c->pk, c->sm,
                                        c->sm + p->sig_bytes, c->smlen - p->sig_bytes});
                    }
                    std::vector<int> ok(jobs.size());
                    e_->verify_many(fam, (int)jobs.size(), jobs.data(), ok.data());
                    for (size_t i = 0; i < cs.size(); ++i) cs[i]->ok = ok[i];
                    break;
                }
            }
            e_->release_workspace();  // the next group reuses the block
            ++launches_;
        }
    }

    DeviceArena* arena_;
    KeyTables* kt_;
    std::mutex blk_mu_;              // the lanes' blocks of the arena
    std::condition_variable blk_cv_;  // a block came back
    int blocks_ = 0;                 // blocks out
    std::mutex mu_;
    std::condition_variable cv_;  // lanes wait here
    std::vector<std::unique_ptr<Lane>> lanes_;
    std::vector<Call*> queue_[NQ];
    Clock::time_point first_[NQ];
    int active_ = 0;   // requests running on workers
    int waiting_ = 0;  // calls queued or executing in any lane
    bool stop_ = false;
    std::atomic<uint64_t> launches_{0};
    uint64_t calls_ = 0;
};

// BDS backend of one in-flight signature: hashing goes through the dispatcher.
struct ServeOps final : BdsOps {
    Dispatcher* d;
    int fam;
    unsigned n;
    int slot;
    size_t bytes;  // the request's reservation (one call of it at a time)
    std::function<int(const unsigned char*, const unsigned char*, const unsigned char*)> slot_for;
    std::vector<LeafId> plan;
    std::map<LeafId, std::vector<unsigned char>> cache;

    void bind_key(const unsigned char* a, const unsigned char* b, const unsigned char* c) override {
        slot = slot_for(a, b, c);
        if (plan.empty()) return;
        Call call;
        call.kind = Call::LEAVES;
        call.fam = fam;
        call.slot = slot;
        call.ids = plan;
        std::vector<unsigned char> out(plan.size() * n);
        call.out = out.data();
        run(call);
        for (size_t i = 0; i < plan.size(); ++i)
            cache[plan[i]] = std::vector<unsigned char>(out.begin() + i * n, out.begin() + (i + 1) * n);
    }
    void leaf(unsigned char* out, const LeafId& id) override {
        auto it = cache.find(id);
        if (it == cache.end()) {  // not planned (cannot happen: control flow is value-independent)
            Call call;
            call.kind = Call::LEAVES;
            call.fam = fam;
            call.slot = slot;
            call.ids = {id};
            call.out = out;
            run(call);
            return;
        }
        memcpy(out, it->second.data(), n);
    }
    void thash_h(unsigned char* out, const unsigned char* in, const uint32_t adrs[8]) override {
        Call call;
        call.kind = Call::THASH;
        call.fam = fam;
        call.slot = slot;
        memcpy(call.adrs, adrs, 32);
        call.in = in;
        std::vector<unsigned char> tmp(n);
        call.out = tmp.data();
        run(call);
        memcpy(out, tmp.data(), n);
    }
    void wots_sign(unsigned char* sig, const unsigned char* msg, uint32_t layer, uint64_t tree,
                   uint32_t ots) override {
        Call call;
        call.kind = Call::WOTS;
        call.fam = fam;
        call.slot = slot;
        call.layer = layer;
        call.tree = tree;
        call.ots = ots;
        std::vector<unsigned char> m(msg, msg + n);
        call.msg = m.data();
        call.out = sig;
        run(call);
    }
    void hmsg_sign(unsigned char* R, unsigned char* digest, uint64_t idx, const unsigned char* root,
                   const unsigned char* m, uint64_t mlen) override {
        Call call;
        call.kind = Call::HMSG;
        call.fam = fam;
        call.slot = slot;
        call.idx = idx;
        call.root = root;
        call.m = m;
        call.mlen = mlen;
        call.out = R;
        call.out2 = digest;
        run(call);
    }
    void run(Call& call) {
        call.bytes = bytes;
        d->run(&call);
        if (call.failed) throw DeviceError("serving lane: device call failed");
    }
};

struct Key {
    int mt = 0;
    xmss_params p{};
    int fam = -1;
    std::vector<unsigned char> sk, pk;  // OID-prefixed; sk is the host mirror of dsk
    unsigned char* dsk = nullptr;       // the resident OID-less sk in the arena
    int slot = -1;                      // key table slot of the key's own seeds
    bool wiped = false;                 // the slot holds PUB_SEED only (the key signed its last index)
    bool busy = false;
    bool failed = false;  // a device call of one of its signatures failed: its requests return -3
};

struct Request {
    uint64_t id = 0;
    bool sign = true;
    int key = 0;
    Key* k = nullptr;  // the key (stable address; taken under the server lock at submission)
    std::vector<unsigned char> data;  // sign: message; verify: sm
    std::vector<unsigned char> out;   // sign: sm; verify: m
    unsigned long long outlen = 0;
    int ret = 0;
    size_t reserve = 0;   // device bytes its launches take (sized when it is first considered)
    bool sized = false;
    bool too_big = false;  // more than the budget left by the keys: fails with -3
    Clock::time_point t0, t1;
    bool done = false;
    std::condition_variable cv;  // its result's waiter (no wake-ups of the others)
};

}  // namespace

struct xmss_mi355x_server {
    std::unique_ptr<DeviceArena> arena;
    KeyTables tables;
    std::unique_ptr<Dispatcher> disp;
    std::mutex mu;
    std::mutex add_mu;  // one add_key at a time (table growth, slot assignment)
    std::condition_variable cv_work;
    std::condition_variable cv_key;  // get_sk waiters: a key became free
    int key_waiters = 0;
    // keys live on the heap, so that a key added while requests run never
    // moves one that a worker or a lane is using
    std::vector<std::unique_ptr<Key>> keys;
    std::deque<Request*> queue;  // arrival order
    std::map<uint64_t, Request*> reqs;
    uint64_t next_id = 1;
    // resident: arena bytes of the key tables and the keys' sks; transient:
    // the reservations of running requests
    size_t budget = 0, resident = 0, transient = 0;
    uint64_t admission_waits = 0;
    bool stop = false;
    std::vector<std::thread> workers;

    // Device bytes a request's launches take: a signature's from its leaf and
    // merge plan (the dry run of its BDS update on the key's current sk), a
    // verification's from its shape and message length.
    size_t reservation(const Request* r) const {
        const Key& k = *r->k;
        const xmss_params& p = k.p;
        if (!r->sign) {
            const unsigned long long smlen = r->data.size();
            if (smlen < p.sig_bytes || smlen - p.sig_bytes > kMaxMsgBytes) return 0;
            // a launch sequence of several shapes strides its scratch by the
            // largest tree height among them: count the largest of any set
            return Engine::verify_bytes(k.fam, kMaxTreeHeight, p.d, smlen - p.sig_bytes);
        }
        if (r->data.size() > kMaxMsgBytes || k.failed) return 0;
        const unsigned char* sk = k.sk.data() + XMSS_OID_LEN;
        std::vector<LeafId> leaves;
        std::vector<uint32_t> merges, dag, wots;
        std::vector<unsigned char> lits;
        bds_plan(&p, sk, r->data.size(), leaves, merges, &dag, &lits, &wots);
        Engine::BdsJob j{};
        j.nplan = (uint32_t)leaves.size();
        j.nmerge = (uint32_t)(merges.size() / 8);
        j.nlit = (uint32_t)(lits.size() / p.n);
        j.plan_wots = wots.data();
        // the last index's host path signs again (a WOTS signature per layer)
        j.nwots = (uint32_t)(wots.size() / 4) + p.d;
        j.mlen = r->data.size();
        j.shape = bds_shape(&p);
        return Engine::bds_bytes(k.fam, j);
    }

    // The request to run next: the oldest one whose key is free, if the
    // budget admits it (FIFO: a request that does not fit holds back the rest).
    Request* pick() {
        for (auto it = queue.begin(); it != queue.end(); ++it) {
            Request* r = *it;
            if (r->sign && r->k->busy) continue;
            if (!r->sized) {  // the key is free: its sk is the one the request will sign with
                r->reserve = reservation(r);
                r->sized = true;
            }
            if (r->reserve > budget - std::min(budget, resident)) {
                r->too_big = true;
                r->reserve = 0;
            } else if (resident + transient + r->reserve > budget) {
                ++admission_waits;
                return nullptr;
            }
            queue.erase(it);
            return r;
        }
        return nullptr;
    }

    void work() {
        for (;;) {
            Request* r = nullptr;
            bool others;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stop || (r = pick()) != nullptr; });
                if (!r) return;
                transient += r->reserve;
                if (r->sign) r->k->busy = true;
                others = !queue.empty();
            }
            // Wake-ups go one worker at a time: a worker that takes a request
            // while others wait wakes the next one, which takes the next
            // admissible request or goes back to sleep -- instead of every
            // idle worker contending for the lock at once.
            if (others) cv_work.notify_one();
            if (r->too_big) {
                r->ret = XMSS_ENODEV;
                r->outlen = 0;
            } else {
                // the worker counts as active (more device work coming) from
                // here until its last device call has returned
                disp->enter();
                bool active = true;
                execute(r, active);
                if (active) disp->leave();
            }
            bool more, key_freed;
            {
                std::lock_guard<std::mutex> lk(mu);
                transient -= r->reserve;
                if (r->sign) r->k->busy = false;
                r->t1 = Clock::now();
                r->done = true;
                r->cv.notify_all();
                more = !queue.empty();  // freed budget or key may admit queued requests
                key_freed = r->sign && key_waiters > 0;
            }
            if (more) cv_work.notify_one();
            if (key_freed) cv_key.notify_all();
        }
    }

    // Writes a key context into a table slot through lane 0 (synchronous).
    void prepare_slot(int fam, int slot, const KeySeeds& ks) {
        Call call;
        call.kind = Call::PREPARE;
        call.fam = fam;
        call.slot = slot;
        call.seeds = ks;
        call.bytes = kPrepareBytes;
        disp->run(&call);
        if (call.failed) throw DeviceError("serving: key context preparation failed");
    }

    // The table slot for the seeds a signature asks for: the key's own slot,
    // or -- for the last index, which signs with the wiped sk -- the family's
    // slot of the all-zero seeds.
    int slot_for(Key& k, const unsigned char* a, const unsigned char* b, const unsigned char* c) {
        const unsigned n = k.p.n;
        bool zero = true;
        for (unsigned i = 0; i < n; ++i) zero = zero && !a[i] && !b[i] && !c[i];
        if (zero) return tables.t[k.fam].zero;
        if (!k.wiped) return k.slot;
        throw DeviceError("serving: seeds of a wiped key");
    }

    void execute(Request* r, bool& active) {
        Key& k = *r->k;
        const xmss_params& p = k.p;
        if (r->sign) {
            // the key's sk is owned by this request while busy
            unsigned char* sk = k.sk.data() + XMSS_OID_LEN;
            r->out.assign(p.sig_bytes + r->data.size(), 0);
            if (k.failed || r->data.size() > kMaxMsgBytes) {
                r->ret = k.failed ? XMSS_ENODEV : -1;
                r->outlen = 0;
                return;
            }
            // a failed device call leaves the mirror as it was before this
            // signature, and the key failed (its device copy is unknown)
            std::vector<unsigned char> before(k.sk);
            BdsDeviceSign ds;
            bool ok = true;
            try {
                if (bds_device_prepare(&p, sk, r->out.data(), r->data.data(), r->data.size(), ds)) {
                    ds.job.key = (uint32_t)k.slot;
                    ds.job.dsk = k.dsk;
                    Call call;
                    call.kind = Call::BDS;
                    call.fam = k.fam;
                    call.p = &p;
                    call.job = &ds.job;
                    call.bytes = r->reserve;
                    disp->run(&call);
                    disp->leave();
                    active = false;
                    if (call.failed) throw DeviceError("serving lane: signature failed");
                    bds_device_finish(&p, r->out.data(), &r->outlen, r->data.size(), ds);
                    r->ret = 0;
                    return;
                }
                // last index or exhausted key: host control, hashing through the dispatcher
                const bool last = !k.wiped;
                ServeOps ops;
                ops.d = disp.get();
                ops.fam = k.fam;
                ops.n = p.n;
                ops.slot = k.slot;
                ops.bytes = r->reserve;
                ops.slot_for = [&](const unsigned char* a, const unsigned char* b, const unsigned char* c) {
                    return slot_for(k, a, b, c);
                };
                ops.plan = bds_plan_leaves(&p, sk, r->data.size());
                r->ret = bds_sign(&p, sk, r->out.data(), &r->outlen, r->data.data(), r->data.size(), ops);
                Engine::upload(k.dsk, sk, p.sk_bytes);  // the mirror changed on the host
                if (last && r->ret == 0) {
                    // the last index was signed and the sk wiped: the key's slot
                    // keeps only PUB_SEED's context, for its verifications
                    std::vector<unsigned char> pub(k.pk.begin() + XMSS_OID_LEN + p.n, k.pk.end());
                    prepare_slot(k.fam, k.slot, {nullptr, nullptr, pub.data()});
                    k.wiped = true;
                }
            } catch (const DeviceError&) {
                ok = false;
            }
            if (!ok) {
                k.sk = before;
                k.failed = true;
                r->ret = XMSS_ENODEV;
                r->outlen = 0;
            }
            volatile unsigned char* v = before.data();
            for (size_t i = 0; i < before.size(); ++i) v[i] = 0;
        } else {
            const unsigned long long smlen = r->data.size();
            r->out.assign(smlen + 1, 0);
            if (smlen < p.sig_bytes || smlen - p.sig_bytes > kMaxMsgBytes) {
                r->ret = -1;
                r->outlen = 0;
                return;
            }
            Call call;
            call.kind = Call::VERIFY;
            call.fam = k.fam;
            call.slot = k.slot;
            call.bytes = r->reserve;
            call.p = &p;
            call.pk = k.pk.data() + XMSS_OID_LEN;
            call.sm = r->data.data();
            call.smlen = smlen;
            disp->run(&call);
            disp->leave();
            active = false;
            const unsigned long long mlen = smlen - p.sig_bytes;
            if (call.failed) {
                r->outlen = 0;
                r->ret = XMSS_ENODEV;
            } else if (call.ok) {
                memcpy(r->out.data(), r->data.data() + p.sig_bytes, mlen);
                r->outlen = mlen;
                r->ret = 0;
            } else {
                r->outlen = 0;
                r->ret = -1;
            }
        }
    }
};

extern "C" {

xmss_mi355x_server* xmss_mi355x_server_create(size_t hbm_budget_bytes, int workers) {
    {
        std::lock_guard<std::mutex> lk(engine_mutex());
        if (!Engine::get()) return nullptr;
    }
    std::unique_ptr<xmss_mi355x_server> s(new xmss_mi355x_server);
    s->budget = (hbm_budget_bytes + 255) & ~(size_t)255;
    try {
        s->arena.reset(new DeviceArena(s->budget));
    } catch (const DeviceError&) {
        return nullptr;  // the device (or the library's limit) cannot hold the budget
    }
    // four lanes on private engines: every kind of call, one more signing
    // lane, two verification lanes; each is stripped of any buffer an earlier
    // user left, so that the server's device memory is its arena alone
    std::vector<Engine*> eng;
    for (int i = 0; i < 4; ++i) {
        Engine* p = Engine::make_private();
        bool ok = p != nullptr;
        if (ok) {
            try {
                p->trim();
            } catch (const DeviceError&) {
                ok = false;
            }
        }
        if (!ok) {
            if (p) Engine::release_private(p);
            for (Engine* e : eng) Engine::release_private(e);
            return nullptr;
        }
        eng.push_back(p);
    }
    s->disp.reset(new Dispatcher(eng[0], {eng[1]}, {eng[2], eng[3]}, s->arena.get(), &s->tables));
    if (workers <= 0) workers = 256;
    xmss_mi355x_server* sp = s.release();
    for (int i = 0; i < workers; ++i) sp->workers.emplace_back([sp] { sp->work(); });
    return sp;
}

int xmss_mi355x_server_add_key(xmss_mi355x_server* s, int mt, const unsigned char* sk, const unsigned char* pk) {
    std::unique_ptr<Key> kp(new Key);
    Key& k = *kp;
    k.mt = mt;
    const uint32_t oid = ((uint32_t)pk[0] << 24) | ((uint32_t)pk[1] << 16) | ((uint32_t)pk[2] << 8) | pk[3];
    if (mt ? xmssmt_parse_oid(&k.p, oid) : xmss_parse_oid(&k.p, oid)) return -1;
    k.fam = fam_of(k.p.func, k.p.n);
    if (k.fam < 0) return -1;
    k.sk.assign(sk, sk + XMSS_OID_LEN + k.p.sk_bytes);
    k.pk.assign(pk, pk + XMSS_OID_LEN + k.p.pk_bytes);
    std::lock_guard<std::mutex> alk(s->add_mu);
    DeviceArena* A = s->arena.get();
    KeyTables::T& T = s->tables.t[k.fam];
    const size_t kb = Engine::pool_key_bytes_of(k.fam);
    const bool first = T.zero < 0;
    const int need = T.n + (first ? 2 : 1);
    const size_t sk_res = (k.p.sk_bytes + 255) & ~(size_t)255;
    {
        // the resident bytes: the sk, and the table's growth
        size_t cap = T.cap, grow = 0;
        while ((size_t)need > cap) cap = std::max<size_t>(16, 2 * cap);
        if (cap > T.cap) grow = ((cap * kb + 255) & ~(size_t)255) - ((T.cap * kb + 255) & ~(size_t)255);
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->resident + grow + sk_res > s->budget) return -1;  // the budget cannot hold the key
        if (cap > T.cap) {
            std::unique_lock<std::shared_mutex> tl(s->tables.mu);  // no group runs meanwhile
            void* np = A->alloc(cap * kb);
            if (!np) return -1;
            try {
                if (T.p) Engine::copy_device(np, T.p, (size_t)T.n * kb);
            } catch (const DeviceError&) {
                A->release(np);
                return -1;
            }
            A->release(T.p);
            T.p = np;
            T.cap = cap;
            s->resident += grow;
        }
        k.dsk = static_cast<unsigned char*>(A->alloc(k.p.sk_bytes));
        if (!k.dsk) return -1;
        s->resident += sk_res;
        if (first) T.zero = T.n++;
        k.slot = T.n++;
    }
    const unsigned n = k.p.n;
    const unsigned char* s0 = k.sk.data() + XMSS_OID_LEN + k.p.index_bytes;
    try {
        Engine::upload(k.dsk, k.sk.data() + XMSS_OID_LEN, k.p.sk_bytes);
        if (first) s->prepare_slot(k.fam, T.zero, {nullptr, nullptr, nullptr});
        // SK_SEED, SK_PRF, PUB_SEED of the sk (idx || SK_SEED || SK_PRF || root || PUB_SEED)
        s->prepare_slot(k.fam, k.slot, {s0, s0 + n, s0 + 3 * n});
    } catch (const DeviceError&) {
        return -1;  // its arena bytes stay reserved; the key is not added
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->keys.push_back(std::move(kp));
    return (int)s->keys.size() - 1;
}

static uint64_t submit(xmss_mi355x_server* s, bool sign, int key, const unsigned char* d, unsigned long long len) {
    Request* r = new Request;
    r->sign = sign;
    r->key = key;
    r->data.assign(d, d + len);
    r->t0 = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    if (key < 0 || key >= (int)s->keys.size()) {
        delete r;
        return 0;
    }
    r->k = s->keys[key].get();
    r->id = s->next_id++;
    s->reqs[r->id] = r;
    s->queue.push_back(r);
    s->cv_work.notify_one();
    return r->id;
}

uint64_t xmss_mi355x_server_sign(xmss_mi355x_server* s, int key, const unsigned char* m, unsigned long long mlen) {
    return submit(s, true, key, m, mlen);
}

uint64_t xmss_mi355x_server_verify(xmss_mi355x_server* s, int key, const unsigned char* sm,
                                   unsigned long long smlen) {
    return submit(s, false, key, sm, smlen);
}

int xmss_mi355x_server_result(xmss_mi355x_server* s, uint64_t id, unsigned char* out, unsigned long long* outlen,
                              double* latency_ms) {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return -4;
    Request* r = it->second;
    r->cv.wait(lk, [r] { return r->done; });
    if (out) memcpy(out, r->out.data(), r->outlen);
    if (outlen) *outlen = r->outlen;
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t1 - r->t0).count();
    const int ret = r->ret;
    s->reqs.erase(it);
    delete r;
    return ret;
}

int xmss_mi355x_server_get_sk(xmss_mi355x_server* s, int key, unsigned char* sk) {
    std::unique_lock<std::mutex> lk(s->mu);
    if (key < 0 || key >= (int)s->keys.size()) return -1;
    Key* k = s->keys[key].get();
    ++s->key_waiters;
    s->cv_key.wait(lk, [&] { return !k->busy; });
    --s->key_waiters;
    memcpy(sk, k->sk.data(), k->sk.size());
    return 0;
}

void xmss_mi355x_server_stats(xmss_mi355x_server* s, uint64_t* launches, uint64_t* calls, size_t* peak_bytes,
                              uint64_t* admission_waits) {
    std::lock_guard<std::mutex> lk(s->mu);
    if (launches) *launches = s->disp->launches();
    if (calls) *calls = s->disp->calls();
    if (peak_bytes) *peak_bytes = s->arena->peak();
    if (admission_waits) *admission_waits = s->admission_waits;
}

void xmss_mi355x_server_memory(xmss_mi355x_server* s, size_t* budget, size_t* resident, size_t* in_use,
                               size_t* peak) {
    std::lock_guard<std::mutex> lk(s->mu);
    if (budget) *budget = s->arena->size();
    if (resident) *resident = s->resident;
    if (in_use) *in_use = s->arena->in_use();
    if (peak) *peak = s->arena->peak();
}

void xmss_mi355x_server_destroy(xmss_mi355x_server* s) {
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    s->disp.reset();
    for (auto& kv : s->reqs) delete kv.second;
    // no key context or sk outlives the server in device memory
    try {
        Engine::clear_device(s->arena->base(), s->arena->size());
    } catch (const DeviceError&) {
    }
    delete s;  // the arena goes with it
}

}  // extern "C"
