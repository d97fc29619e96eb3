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
// Key state is resident in HBM for the lifetime of the key: the key context
// (the PRF midstates of SK_SEED, SK_PRF, PUB_SEED) in the device key pool and
// the serialized sk (index, seeds, BDS states, XMSS^MT WOTS signatures) in a
// buffer of its own, which the state-update kernel advances in place.  The
// server keeps a byte-identical host mirror of every sk (each signature's
// result copy brings the updated bytes back): the host plans the next
// signature's leaves from it, serves xmss_mi355x_server_get_sk from it, and
// runs the last-index / exhausted path on it (then uploads it again).
// Admission: adding a key reserves its resident bytes; every request
// reserves its device working set (message, signature, planned leaves)
// against the same fixed HBM budget and waits in FIFO order while the budget
// is exhausted.
#include <string.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
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

// ---------------------------------------------------------------- dispatcher
struct Call {
    enum Kind { LEAVES, THASH, WOTS, HMSG, VERIFY, BDS } kind;
    Engine::BdsJob* job = nullptr;  // BDS
    int fam = 0;
    int slot = 0;
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
    // result
    unsigned char* out = nullptr;
    bool done = false;
    bool failed = false;  // the lane's device work failed (DeviceError)
    std::condition_variable cv;  // the caller waits here (no wake-ups of unrelated callers)
};

class Dispatcher {
public:
    // Lanes, each with its own thread and engine (HIP streams and buffers):
    // lane 0 runs on the library's shared engine and takes every kind of call;
    // the other signing lanes take BDS signatures, the verification lanes
    // verifications, each on a private engine.  A lane launches what its
    // queue holds as soon as it is idle and every running request is waiting
    // on the device (queued or executing anywhere), or 200 us after the
    // oldest queued call -- so a request that arrives while a group is on
    // the GPU starts in the next idle lane instead of waiting for that group.
    Dispatcher(Engine* shared, const std::vector<Engine*>& sign, const std::vector<Engine*>& verify) {
        lanes_.emplace_back(new Lane{shared, true, (1u << Q_OTHER) | (1u << Q_SIGN)});
        for (Engine* e : sign) lanes_.emplace_back(new Lane{e, false, 1u << Q_SIGN});
        for (Engine* e : verify) lanes_.emplace_back(new Lane{e, false, 1u << Q_VERIFY});
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
            if (!l->shared) Engine::release_private(l->engine);
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
        Lane(Engine* e, bool s, uint32_t q) : engine(e), shared(s), queues(q) {}
        Engine* engine;
        bool shared;      // the library's engine: hold the library-wide engine lock
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
            try {
                execute_on(l.engine, batch, l.shared);
            } catch (const DeviceError&) {
                // every call of the group fails; the lane's engine is made
                // consistent again for the next group
                for (Call* c : batch) c->failed = true;
                std::unique_lock<std::mutex> elk(engine_mutex(), std::defer_lock);
                if (l.shared) elk.lock();
                try {
                    l.engine->recover();
                } catch (...) {
                }
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

    void execute_on(Engine* e_, std::vector<Call*>& batch, bool shared) {
        std::unique_lock<std::mutex> elk(engine_mutex(), std::defer_lock);
        if (shared) elk.lock();
        e_->pool_begin();  // the group's key contexts stay in the pool until it is done
        std::map<std::pair<int, int>, std::vector<Call*>> groups;  // (fam, kind)
        for (Call* c : batch) groups[{c->fam, (int)c->kind}].push_back(c);
        for (auto& g : groups) {
            const int fam = g.first.first;
            std::vector<Call*>& cs = g.second;
            const unsigned n = fam_n(fam);
            switch ((Call::Kind)g.first.second) {
                case Call::LEAVES: {
                    e_->pool_use(fam);
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
                    e_->pool_use(fam);
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
                    e_->pool_use(fam);
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
                    e_->pool_use(fam);
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
                    // signatures with the traversal on the device: one batch for all shapes,
                    // key contexts from this engine's own pool (cached by seed)
                    std::vector<Engine::BdsJob> jobs;
                    for (Call* c : cs) {
                        jobs.push_back(*c->job);
                        Engine::BdsJob& j = jobs.back();
                        j.key = (uint32_t)e_->pool_slot(fam, {j.sk_seed, j.sk_prf, j.pub_seed});
                    }
                    e_->bds_sign_batch(fam, (int)jobs.size(), jobs.data());
                    break;
                }
                case Call::VERIFY: {
                    // every parameter shape of the family in one launch sequence
                    // (PUB_SEED contexts in this engine's own pool, cached by seed)
                    std::vector<Engine::VerifyJob> jobs;
                    for (Call* c : cs) {
                        const xmss_params* p = c->p;
                        const uint32_t slot = (uint32_t)e_->pool_slot(fam, {nullptr, nullptr, c->pk + p->n});
                        jobs.push_back({slot, p->tree_height, p->d, p->index_bytes, c->pk, c->sm,
                                        c->sm + p->sig_bytes, c->smlen - p->sig_bytes});
                    }
                    std::vector<int> ok(jobs.size());
                    e_->verify_many(fam, (int)jobs.size(), jobs.data(), ok.data());
                    for (size_t i = 0; i < cs.size(); ++i) cs[i]->ok = ok[i];
                    break;
                }
            }
            ++launches_;
        }
    }

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
        d->run(&call);
        if (call.failed) throw DeviceError("serving lane: device call failed");
    }
};

struct Key {
    int mt = 0;
    xmss_params p{};
    int fam = -1;
    std::vector<unsigned char> sk, pk;  // OID-prefixed; sk is the host mirror of dsk
    unsigned char* dsk = nullptr;       // the resident OID-less sk in HBM
    int slot = -1;                      // pool slot of the key's own seeds
    std::vector<unsigned char> seeds;   // SK_SEED || SK_PRF || PUB_SEED the slot was made from
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
    size_t reserve = 0;
    Clock::time_point t0, t1;
    bool done = false;
    std::condition_variable cv;  // its result's waiter (no wake-ups of the others)
};

}  // namespace

struct xmss_mi355x_server {
    Engine* e = nullptr;
    std::unique_ptr<Dispatcher> disp;
    std::mutex mu;
    std::condition_variable cv_work;
    std::condition_variable cv_key;  // get_sk waiters: a key became free
    int key_waiters = 0;
    // keys live on the heap, so that a key added while requests run never
    // moves one that a worker or a lane is using
    std::vector<std::unique_ptr<Key>> keys;
    std::deque<Request*> queue;  // arrival order
    std::map<uint64_t, Request*> reqs;
    uint64_t next_id = 1;
    // resident: key contexts and sks of the added keys; transient: the
    // working sets of running requests; peak: of their sum
    size_t budget = 0, resident = 0, transient = 0, peak = 0;
    uint64_t admission_waits = 0;
    bool stop = false;
    std::vector<std::thread> workers;

    // The request to run next: the oldest one whose key is free, if the
    // budget admits it (FIFO: a request that does not fit holds back the rest;
    // with no request running, the oldest runs whatever its size).
    Request* pick() {
        for (auto it = queue.begin(); it != queue.end(); ++it) {
            Request* r = *it;
            if (r->sign && r->k->busy) continue;
            if (resident + transient + r->reserve > budget && transient > 0) {
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
                peak = std::max(peak, resident + transient);
                if (r->sign) r->k->busy = true;
                others = !queue.empty();
            }
            // Wake-ups go one worker at a time: a worker that takes a request
            // while others wait wakes the next one, which takes the next
            // admissible request or goes back to sleep -- instead of every
            // idle worker contending for the lock at once.
            if (others) cv_work.notify_one();
            // the worker counts as active (more device work coming) from
            // here until its last device call has returned
            disp->enter();
            bool active = true;
            execute(r, active);
            if (active) disp->leave();
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

    // The pool slot for the seeds a signature asks for: the key's own slot,
    // or -- for the last index, which signs with the wiped (zero) seeds -- the
    // pool's slot of those seeds.  Compared with the seeds the key's slot was
    // made from, not with the sk, which the signature has already wiped.
    int slot_for(Key& k, const unsigned char* a, const unsigned char* b, const unsigned char* c) {
        const unsigned n = k.p.n;
        std::lock_guard<std::mutex> elk(engine_mutex());
        if (k.slot >= 0 && !memcmp(a, &k.seeds[0], n) && !memcmp(b, &k.seeds[n], n) && !memcmp(c, &k.seeds[2 * n], n))
            return k.slot;
        return e->pool_slot(k.fam, {a, b, c});
    }
    std::vector<unsigned char> seeds_of(const Key& k) {
        const unsigned n = k.p.n, ib = k.p.index_bytes;
        const unsigned char* sk = k.sk.data() + XMSS_OID_LEN;
        std::vector<unsigned char> v(3 * n);
        memcpy(&v[0], sk + ib, n);
        memcpy(&v[n], sk + ib + n, n);
        memcpy(&v[2 * n], sk + ib + 3 * n, n);
        return v;
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
                    disp->run(&call);
                    disp->leave();
                    active = false;
                    if (call.failed) throw DeviceError("serving lane: signature failed");
                    bds_device_finish(&p, r->out.data(), &r->outlen, r->data.size(), ds);
                    r->ret = 0;
                    return;
                }
                // last index or exhausted key: host control, hashing through the dispatcher
                ServeOps ops;
                ops.d = disp.get();
                ops.fam = k.fam;
                ops.n = p.n;
                ops.slot = k.slot;
                ops.slot_for = [&](const unsigned char* a, const unsigned char* b, const unsigned char* c) {
                    return slot_for(k, a, b, c);
                };
                ops.plan = bds_plan_leaves(&p, sk, r->data.size());
                r->ret = bds_sign(&p, sk, r->out.data(), &r->outlen, r->data.data(), r->data.size(), ops);
                Engine::upload(k.dsk, sk, p.sk_bytes);  // the mirror changed on the host
                if (r->ret == 0 && k.slot >= 0) {
                    // the last index was signed and the sk wiped: the pool keeps
                    // no context of the old seeds
                    std::lock_guard<std::mutex> elk(engine_mutex());
                    e->pool_release(k.fam, k.slot);
                    k.slot = -1;
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
    Engine* e;
    {
        std::lock_guard<std::mutex> lk(engine_mutex());
        e = Engine::get();
    }
    if (!e) return nullptr;
    xmss_mi355x_server* s = new xmss_mi355x_server;
    s->e = e;
    s->budget = hbm_budget_bytes;
    // one more signing lane and two verification lanes on private engines
    std::vector<Engine*> sign_lanes, verify_lanes;
    for (int i = 0; i < 3; ++i) {
        Engine* p = Engine::make_private();
        if (p) (i == 0 ? sign_lanes : verify_lanes).push_back(p);
    }
    s->disp.reset(new Dispatcher(e, sign_lanes, verify_lanes));
    if (workers <= 0) workers = 256;
    for (int i = 0; i < workers; ++i) s->workers.emplace_back([s] { s->work(); });
    return s;
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
    const size_t res_bytes = s->e->pool_key_bytes(k.fam) + k.p.sk_bytes;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->resident + res_bytes > s->budget) return -1;  // the budget cannot hold the key
        s->resident += res_bytes;
        s->peak = std::max(s->peak, s->resident + s->transient);
    }
    k.dsk = static_cast<unsigned char*>(Engine::device_alloc(k.p.sk_bytes));
    bool ok = k.dsk != nullptr;
    if (ok) {
        try {
            Engine::upload(k.dsk, k.sk.data() + XMSS_OID_LEN, k.p.sk_bytes);
            k.seeds = s->seeds_of(k);
            const unsigned n = k.p.n;
            KeySeeds ks{&k.seeds[0], &k.seeds[n], &k.seeds[2 * n]};
            std::lock_guard<std::mutex> elk(engine_mutex());
            k.slot = s->e->pool_add(k.fam, ks);
        } catch (const DeviceError&) {
            ok = false;
        }
    }
    if (!ok) {
        Engine::device_free(k.dsk, k.p.sk_bytes);
        std::lock_guard<std::mutex> lk(s->mu);
        s->resident -= res_bytes;
        return -1;
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
    const xmss_params& p = r->k->p;
    // device working set: message + signature staging, planned leaves, H_msg prefix
    r->reserve = (size_t)p.sig_bytes + len + (size_t)(p.tree_height + 1) * p.n * 2 + 4 * p.n;
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
    if (peak_bytes) *peak_bytes = s->peak;
    if (admission_waits) *admission_waits = s->admission_waits;
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
    for (auto& k : s->keys) {
        Engine::device_free(k->dsk, k->p.sk_bytes);
        try {
            std::lock_guard<std::mutex> elk(engine_mutex());
            if (k->slot >= 0) s->e->pool_release(k->fam, k->slot);
        } catch (const DeviceError&) {
        }
    }
    delete s;
}

}  // extern "C"
