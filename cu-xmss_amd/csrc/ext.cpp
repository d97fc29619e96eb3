// ext.cpp -- additions beyond the reference API (include/xmss_mi355x.h):
// batches of independent keys, device check, kernel timing.
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/xmss_core.h"
#include "../../include/xmss_mi355x.h"
#include "core_common.h"

using namespace xm;

// One parameter set of a job: its batch on one engine.  A single-set job has
// one; a mixed job one per (XMSS / XMSS^MT, OID), so that their launches run
// side by side on the GPU: a signing job's sets each on an engine of its own,
// a verification job's on the three streams of the library's engine.
struct JobSet {
    int mt = 0;
    uint32_t oid = 0;
    xmss_params p{};
    int fam = -1;
    Engine* e = nullptr;
    bool shared = true;        // e is the library's engine (else a private one)
    Engine::Batch* b = nullptr;
    std::vector<size_t> live;  // job inputs that reach the device, in batch order
    std::vector<int> ok;       // verify: verdicts of the live inputs (collect)
};

struct xmss_mi355x_job {
    bool sign = false;
    bool failed = false;  // a device call failed: every input that reached the device returns -3
    size_t count = 0;
    std::vector<int> rets;
    std::vector<int> set;  // input -> index in sets (-1: unknown OID)
    std::vector<JobSet> sets;
    // sign (the pointer arrays are copied; the buffers they point to must
    // stay valid until collect)
    std::vector<unsigned char> idxbytes;  // 8 bytes per input
    std::vector<const unsigned char*> ms;
    std::vector<unsigned long long> mlens;
    // verify
    std::vector<const unsigned char*> sms;
    std::vector<unsigned long long> smlens;
    std::vector<const unsigned char*> pks;
    std::vector<uint32_t> oids;  // of each input's pk, for refills
};

namespace {
int batch_keypair(int mt, unsigned char* const* pks, unsigned char* const* sks, const unsigned char* seeds,
                  uint32_t oid, size_t count, const xmss_params& p, int fam);
uint32_t oid_of(const unsigned char* b) {
    return ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | b[3];
}
int parse(int mt, uint32_t oid, xmss_params* p) { return mt ? xmssmt_parse_oid(p, oid) : xmss_parse_oid(p, oid); }
// A signed message the device path verifies: at least a signature, and a
// message no longer than kMaxMsgBytes (others are rejected with -1).
bool verifiable(const xmss_params& p, unsigned long long smlen) {
    return smlen >= p.sig_bytes && smlen - p.sig_bytes <= kMaxMsgBytes;
}
}  // namespace

extern "C" {

int xmss_mi355x_device_ok(void) {
    std::lock_guard<std::mutex> lk(engine_mutex());
    return Engine::get() != nullptr;
}

int xmss_mi355x_valu_peak(unsigned blocks, unsigned iters, double* lane_ops_per_s, double* ms) {
    return xmss_mi355x_valu_probe(0, blocks, iters, lane_ops_per_s, ms);
}

int xmss_mi355x_valu_probe(int kind, unsigned blocks, unsigned iters, double* lane_ops_per_s, double* ms) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(engine_mutex());
        if (!Engine::get()) return XMSS_ENODEV;
        return valu_probe(kind, blocks, iters, lane_ops_per_s, ms);
    });
}

size_t xmss_mi355x_set_device_limit(size_t bytes) { return set_device_limit(bytes); }

size_t xmss_mi355x_device_bytes(void) { return device_bytes(); }

void xmss_mi355x_profile(int on) {
    std::lock_guard<std::mutex> lk(engine_mutex());
    if (Engine* e = Engine::get()) e->set_profiling(on != 0);
}

void xmss_mi355x_reset_stats(void) {
    std::lock_guard<std::mutex> lk(engine_mutex());
    if (Engine* e = Engine::get()) e->reset_stats();
}

int xmss_mi355x_set_verify_plan(int plan) { return xm::set_verify_plan(plan); }

size_t xmss_mi355x_verify_plans(unsigned* plans, size_t max) {
    std::vector<uint32_t> v(16);
    const size_t n = xm::verify_plans(v.data(), v.size());
    for (size_t i = 0; i < n && i < max && plans; ++i) plans[i] = v[i];
    return n;
}

size_t xmss_mi355x_tuned_plans(char* buf, size_t len) {
    const std::string s = xm::tuned_plans();
    if (buf && len) {
        const size_t m = std::min(len - 1, s.size());
        memcpy(buf, s.data(), m);
        buf[m] = 0;
    }
    return s.size() + 1;
}

size_t xmss_mi355x_kernel_stats(char* buf, size_t len) {
    std::lock_guard<std::mutex> lk(engine_mutex());
    std::string s;
    if (Engine* e = Engine::get()) {
        for (auto& k : e->stats()) {
            char line[256];
            snprintf(line, sizeof line, "%s %llu %.6f\n", k.name.c_str(), (unsigned long long)k.count, k.total_ms);
            s += line;
        }
    }
    if (buf && len) {
        const size_t m = std::min(len - 1, s.size());
        memcpy(buf, s.data(), m);
        buf[m] = 0;
    }
    return s.size() + 1;
}

}  // extern "C"

namespace {

// The set of input i (created on first use), or -1 for an unknown OID.  With
// `mixed` false every input must carry the first input's OID.
int set_of(xmss_mi355x_job* j, int mt, uint32_t oid, bool mixed) {
    for (size_t s = 0; s < j->sets.size(); ++s)
        if (j->sets[s].mt == mt && j->sets[s].oid == oid) return (int)s;
    if (!mixed && !j->sets.empty()) return -1;
    JobSet js;
    js.mt = mt;
    js.oid = oid;
    if (parse(mt, oid, &js.p) != 0) return -1;
    js.fam = fam_of(js.p.func, js.p.n);
    if (js.fam < 0) return -1;
    j->sets.push_back(js);
    return (int)j->sets.size() - 1;
}

// Engines for the sets that reach the device: the library's for the first,
// private ones for the others (one_engine: the library's for all).  Called
// with the engine lock held.
bool bind_engines(xmss_mi355x_job* j, bool one_engine = false) {
    Engine* shared = Engine::get();
    if (!shared) return false;
    bool first = true;
    for (JobSet& s : j->sets) {
        if (s.live.empty()) continue;
        s.e = first || one_engine ? shared : Engine::make_private();
        s.shared = first || one_engine;
        if (!s.e) {
            s.e = shared;
            s.shared = true;
        }
        first = false;
    }
    return true;
}

void fail_all(xmss_mi355x_job* j, int rc) {
    for (JobSet& s : j->sets) {
        for (size_t i : s.live) j->rets[i] = rc;
        s.live.clear();
    }
}

// After a device failure: every set's batch is dropped (its engine brought
// back to a consistent state) and every input that reached the device
// returns -3.  Called with the engine lock held.
void fail_job(xmss_mi355x_job* j) {
    for (JobSet& s : j->sets) {
        if (!s.e) continue;
        try {
            s.e->recover();
            if (s.b) s.e->free_batch(s.b);
        } catch (...) {
        }
        s.b = nullptr;
    }
    fail_all(j, XMSS_ENODEV);
    j->failed = true;
}

xmss_mi355x_job* sign_prepare(const int* mts, int mt_all, unsigned char* const* sks, const unsigned char* const* ms,
                              const unsigned long long* mlens, size_t count, bool mixed) {
    xmss_mi355x_job* j = new xmss_mi355x_job;
    j->sign = true;
    j->count = count;
    j->ms.assign(ms, ms + count);
    j->mlens.assign(mlens, mlens + count);
    j->rets.assign(count, -1);
    j->set.assign(count, -1);
    j->idxbytes.assign(count * 8, 0);
    std::vector<std::vector<Engine::SignItem>> items;
    for (size_t i = 0; i < count; ++i) {
        const int s = set_of(j, mts ? mts[i] : mt_all, oid_of(sks[i]), mixed);
        if (s < 0) continue;  // unknown OID (or, single-set job, another parameter set)
        j->set[i] = s;
        items.resize(j->sets.size());
        const xmss_params& p = j->sets[s].p;
        const unsigned n = p.n, ib = p.index_bytes;
        unsigned char* sk = sks[i] + XMSS_OID_LEN;
        if (mlens[i] > kMaxMsgBytes) continue;  // -1, sk unchanged
        const unsigned long long idx = (unsigned long)bytes_to_ull(sk, ib);
        if (index_check(&p, sk, idx)) {
            j->rets[i] = -2;
            continue;
        }
        memcpy(&j->idxbytes[i * 8], sk, ib);
        ull_to_bytes(sk, ib, idx + 1);
        items[s].push_back({sk + ib, sk + ib + n, sk + ib + 3 * n, sk + ib + 2 * n, idx, ms[i], mlens[i]});
        j->sets[s].live.push_back(i);
        j->rets[i] = 0;
    }
    std::lock_guard<std::mutex> lk(engine_mutex());
    try {
        if (!bind_engines(j)) {
            fail_all(j, XMSS_ENODEV);
            return j;
        }
        for (size_t s = 0; s < j->sets.size(); ++s) {
            JobSet& js = j->sets[s];
            if (js.live.empty()) continue;
            js.b = js.e->sign_prepare(js.fam, js.p.tree_height, js.p.d, (int)items[s].size(), items[s].data());
            js.e->set_graphed(js.b, true);  // its launches go to the GPU as one HIP graph
        }
    } catch (const DeviceError&) {
        fail_job(j);
    }
    return j;
}

xmss_mi355x_job* verify_prepare(const int* mts, int mt_all, const unsigned char* const* sms,
                                const unsigned long long* smlens, const unsigned char* const* pks, size_t count,
                                bool mixed) {
    xmss_mi355x_job* j = new xmss_mi355x_job;
    j->sign = false;
    j->count = count;
    j->sms.assign(sms, sms + count);
    j->smlens.assign(smlens, smlens + count);
    j->pks.assign(pks, pks + count);
    j->rets.assign(count, -1);
    j->set.assign(count, -1);
    j->oids.resize(count);
    std::vector<std::vector<Engine::VerifyItem>> items;
    for (size_t i = 0; i < count; ++i) {
        j->oids[i] = oid_of(pks[i]);
        const int s = set_of(j, mts ? mts[i] : mt_all, oid_of(pks[i]), mixed);
        if (s < 0) continue;
        j->set[i] = s;
        items.resize(j->sets.size());
        const xmss_params& p = j->sets[s].p;
        const unsigned n = p.n, ib = p.index_bytes;
        if (!verifiable(p, smlens[i])) continue;
        const unsigned char* sm = sms[i];
        items[s].push_back({pks[i] + XMSS_OID_LEN, bytes_to_ull(sm, ib), sm + ib, sm + ib + n, sm + p.sig_bytes,
                            smlens[i] - p.sig_bytes});
        j->sets[s].live.push_back(i);
    }
    std::lock_guard<std::mutex> lk(engine_mutex());
    try {
        // verification: every set on the library's engine, each on its own
        // stream of it (below)
        if (!bind_engines(j, true)) {
            fail_all(j, XMSS_ENODEV);
            return j;
        }
        size_t live_sets = 0, k = 0;
        for (JobSet& js : j->sets) live_sets += !js.live.empty();
        for (size_t s = 0; s < j->sets.size(); ++s) {
            JobSet& js = j->sets[s];
            if (js.live.empty()) continue;
            js.b = js.e->verify_prepare(js.fam, js.p.tree_height, js.p.d, (int)items[s].size(), items[s].data());
            // several sets: each keeps to one stream of the engine, the sets
            // on its three streams in turn, which HIP puts on three of the
            // process's hardware queues (four by default), so that they run
            // side by side.  (Sets on engines of their own had their main
            // streams on queues that depended on how many streams the process
            // had created before -- HIP's graph launches create some -- and two
            // sets sharing a queue ran one after another.)
            if (live_sets > 1) {
                js.e->set_one_stream(js.b, true);
                js.e->set_stream(js.b, (int)(k++ % 3));
            }
        }
    } catch (const DeviceError&) {
        fail_job(j);
    }
    return j;
}

}  // namespace

extern "C" {

xmss_mi355x_job* xmss_mi355x_sign_prepare(int mt, unsigned char* const* sks, const unsigned char* const* ms,
                                          const unsigned long long* mlens, size_t count) {
    return sign_prepare(nullptr, mt, sks, ms, mlens, count, false);
}

xmss_mi355x_job* xmss_mi355x_verify_prepare(int mt, const unsigned char* const* sms,
                                            const unsigned long long* smlens, const unsigned char* const* pks,
                                            size_t count) {
    return verify_prepare(nullptr, mt, sms, smlens, pks, count, false);
}

xmss_mi355x_job* xmss_mi355x_mixed_sign_prepare(const int* mts, unsigned char* const* sks,
                                                const unsigned char* const* ms, const unsigned long long* mlens,
                                                size_t count) {
    return sign_prepare(mts, 0, sks, ms, mlens, count, true);
}

xmss_mi355x_job* xmss_mi355x_mixed_verify_prepare(const int* mts, const unsigned char* const* sms,
                                                  const unsigned long long* smlens,
                                                  const unsigned char* const* pks, size_t count) {
    return verify_prepare(mts, 0, sms, smlens, pks, count, true);
}

int xmss_mi355x_verify_refill(xmss_mi355x_job* j, const unsigned char* const* sms,
                              const unsigned long long* smlens, const unsigned char* const* pks) {
    if (!j || j->sign) return -1;
    if (j->failed) return XMSS_ENODEV;
    // the same shape: each input keeps its public key's OID and whether it
    // reaches the device (a signed message at least sig_bytes long)
    std::vector<std::vector<Engine::VerifyItem>> items(j->sets.size());
    std::vector<size_t> at(j->sets.size(), 0);
    for (size_t i = 0; i < j->count; ++i) {
        if (oid_of(pks[i]) != j->oids[i]) return -1;
        const int s = j->set[i];
        if (s < 0) continue;
        JobSet& js = j->sets[s];
        const xmss_params& p = js.p;
        const bool live = at[s] < js.live.size() && js.live[at[s]] == i;
        if (live != verifiable(p, smlens[i])) return -1;
        if (!live) continue;
        ++at[s];
        const unsigned n = p.n, ib = p.index_bytes;
        const unsigned char* sm = sms[i];
        items[s].push_back({pks[i] + XMSS_OID_LEN, bytes_to_ull(sm, ib), sm + ib, sm + ib + n, sm + p.sig_bytes,
                            smlens[i] - p.sig_bytes});
    }
    j->sms.assign(sms, sms + j->count);
    j->smlens.assign(smlens, smlens + j->count);
    j->pks.assign(pks, pks + j->count);
    std::lock_guard<std::mutex> lk(engine_mutex());
    try {
        for (size_t s = 0; s < j->sets.size(); ++s)
            if (j->sets[s].b) j->sets[s].e->verify_refill(j->sets[s].b, items[s].data());
    } catch (const DeviceError&) {
        fail_job(j);
        return XMSS_ENODEV;
    }
    return 0;
}

int xmss_mi355x_job_launch(xmss_mi355x_job* j) {
    if (!j) return -1;
    if (j->failed) return XMSS_ENODEV;
    std::lock_guard<std::mutex> lk(engine_mutex());
    try {
        // every set's launches are queued on its own engine's streams, so the
        // parameter sets share the GPU
        for (JobSet& s : j->sets)
            if (s.b) s.e->launch(s.b);
    } catch (const DeviceError&) {
        fail_job(j);
        return XMSS_ENODEV;
    }
    return 0;
}

int xmss_mi355x_job_wait(xmss_mi355x_job* j) {
    if (!j) return -1;
    if (j->failed) return XMSS_ENODEV;
    try {
        bool prof = false;
        for (JobSet& s : j->sets) {
            if (!s.b) continue;
            s.e->wait_batch(s.b);  // without the engine lock: other callers go on meanwhile
            prof = prof || s.e->profiling();
        }
        if (prof) {  // per-kernel timings are read once the engines are idle
            std::lock_guard<std::mutex> lk(engine_mutex());
            for (JobSet& s : j->sets)
                if (s.b) s.e->wait();
        }
    } catch (const DeviceError&) {
        std::lock_guard<std::mutex> lk(engine_mutex());
        fail_job(j);
        return XMSS_ENODEV;
    }
    return 0;
}

int xmss_mi355x_job_run(xmss_mi355x_job* j) {
    const int rc = xmss_mi355x_job_launch(j);
    return rc ? rc : xmss_mi355x_job_wait(j);
}

int xmss_mi355x_sign_collect(xmss_mi355x_job* j, unsigned char* const* sms, unsigned long long* smlens,
                             int* rets) {
    if (!j || !j->sign) return -1;
    {
        std::lock_guard<std::mutex> lk(engine_mutex());
        try {
            for (JobSet& s : j->sets) {
                if (!s.b) continue;
                std::vector<unsigned char*> outs;
                for (size_t i : s.live) outs.push_back(sms[i] + s.p.index_bytes);
                s.e->collect_sign(s.b, outs.data());
            }
        } catch (const DeviceError&) {
            fail_job(j);
        }
    }
    parallel_for(j->count, 512, [&](size_t i) {
        rets[i] = j->rets[i];
        if (j->set[i] < 0) return;
        const xmss_params& p = j->sets[j->set[i]].p;
        if (j->rets[i] == 0 || j->rets[i] == -2) {
            // message after the signature, smlen set even for -2 (xmss_core.c:205-206)
            memmove(sms[i] + p.sig_bytes, j->ms[i], j->mlens[i]);
            smlens[i] = p.sig_bytes + j->mlens[i];
        }
        if (j->rets[i] == 0) memcpy(sms[i], &j->idxbytes[i * 8], p.index_bytes);
    });
    return 0;
}

int xmss_mi355x_verify_collect(xmss_mi355x_job* j, unsigned char* const* ms, unsigned long long* mlens,
                               int* rets) {
    if (!j || j->sign) return -1;
    if (!j->failed) {
        std::lock_guard<std::mutex> lk(engine_mutex());
        try {
            for (JobSet& s : j->sets) {
                s.ok.assign(s.live.size(), 0);
                if (s.b) s.e->collect_verify(s.b, s.ok.data());
            }
        } catch (const DeviceError&) {
            fail_job(j);
        }
    }
    for (size_t i = 0; i < j->count; ++i) {
        rets[i] = j->rets[i];
        if (rets[i] != 0) mlens[i] = 0;
    }
    for (JobSet& s : j->sets) {
        const xmss_params& p = s.p;
        const std::vector<int>& ok = s.ok;
        parallel_for(s.live.size(), 512, [&](size_t k) {
            const size_t i = s.live[k];
            const unsigned char* sm = j->sms[i];
            unsigned char* m = ms[i];
            const unsigned long long mlen = j->smlens[i] - p.sig_bytes;
            memmove(m + p.sig_bytes, sm + p.sig_bytes, mlen);
            unsigned char* pre = m + p.sig_bytes - p.padding_len - 3 * p.n;
            ull_to_bytes(pre, p.padding_len, 2);
            memcpy(pre + p.padding_len, sm + p.index_bytes, p.n);
            memcpy(pre + p.padding_len + p.n, j->pks[i] + XMSS_OID_LEN, p.n);
            ull_to_bytes(pre + p.padding_len + 2 * p.n, p.n, bytes_to_ull(sm, p.index_bytes));
            if (ok[k]) {
                memmove(m, sm + p.sig_bytes, mlen);
                mlens[i] = mlen;
                rets[i] = 0;
            } else {
                memset(m, 0, mlen);
                mlens[i] = 0;
                rets[i] = -1;
            }
        });
    }
    return 0;
}

void xmss_mi355x_job_free(xmss_mi355x_job* j) {
    if (!j) return;
    {
        std::lock_guard<std::mutex> lk(engine_mutex());
        for (JobSet& s : j->sets) {
            try {
                if (s.b) s.e->free_batch(s.b);
            } catch (const DeviceError&) {
                s.e->recover();
            }
            if (s.e && !s.shared) Engine::release_private(s.e);
        }
    }
    delete j;
}

int xmss_batch_keypair(int mt, unsigned char* const* pks, unsigned char* const* sks, const unsigned char* seeds,
                       uint32_t oid, size_t count) {
    xmss_params p;
    if (parse(mt, oid, &p)) return -1;
    const int fam = fam_of(p.func, p.n);
    if (fam < 0) return -1;
    if (count == 0) return 0;
    return guarded([&] { return batch_keypair(mt, pks, sks, seeds, oid, count, p, fam); });
}

}  // extern "C"

namespace {

int batch_keypair(int mt, unsigned char* const* pks, unsigned char* const* sks, const unsigned char* seeds,
                  uint32_t oid, size_t count, const xmss_params& p, int fam) {
    (void)mt;
    const unsigned n = p.n, ib = p.index_bytes;
    std::lock_guard<std::mutex> lk(engine_mutex());
    Engine* e = Engine::get();
    if (!e) return XMSS_ENODEV;
    std::vector<KeySeeds> ks(count);
    std::vector<LeafRun> runs(count);
    std::vector<NodeRef> refs(count);
    for (size_t i = 0; i < count; ++i) {
        const unsigned char* s = seeds + i * 3 * n;
        ks[i] = {s, s + n, s + 2 * n};
        runs[i] = {(uint32_t)i, p.d - 1, 0, 0};
        refs[i] = {(uint32_t)i, p.tree_height, 0};
    }
    e->set_keys(fam, (int)count, ks.data());
    e->build_trees(fam, p.tree_height, (int)count, runs.data());
    std::vector<unsigned char> roots(count * n);
    e->fetch_nodes(fam, p.tree_height, (int)count, refs.data(), roots.data());
    for (size_t i = 0; i < count; ++i) {
        const unsigned char* s = seeds + i * 3 * n;
        unsigned char* pk = pks[i];
        unsigned char* sk = sks[i];
        for (unsigned b = 0; b < XMSS_OID_LEN; ++b) pk[XMSS_OID_LEN - b - 1] = sk[XMSS_OID_LEN - b - 1] = (oid >> (8 * b)) & 0xFF;
        pk += XMSS_OID_LEN;
        sk += XMSS_OID_LEN;
        // xmss_core.c:146-158 layout
        memset(sk, 0, ib);
        memcpy(sk + ib, s, 2 * n);
        memcpy(sk + ib + 3 * n, s + 2 * n, n);
        memcpy(pk + n, s + 2 * n, n);
        memcpy(pk, &roots[i * n], n);
        memcpy(sk + ib + 2 * n, pk, n);
    }
    e->scrub_keys();
    return 0;
}

}  // namespace

extern "C" {

int xmss_batch_sign(int mt, unsigned char* const* sks, unsigned char* const* sms, unsigned long long* smlens,
                    const unsigned char* const* ms, const unsigned long long* mlens, int* rets, size_t count) {
    xmss_mi355x_job* j = xmss_mi355x_sign_prepare(mt, sks, ms, mlens, count);
    // collect also after a failed run: it reports -3 for the inputs that
    // reached the device
    const int rc = xmss_mi355x_job_run(j);
    const int rc2 = xmss_mi355x_sign_collect(j, sms, smlens, rets);
    xmss_mi355x_job_free(j);
    return rc ? rc : rc2;
}

int xmss_batch_sign_open(int mt, unsigned char* const* ms, unsigned long long* mlens,
                         const unsigned char* const* sms, const unsigned long long* smlens,
                         const unsigned char* const* pks, int* rets, size_t count) {
    xmss_mi355x_job* j = xmss_mi355x_verify_prepare(mt, sms, smlens, pks, count);
    const int rc = xmss_mi355x_job_run(j);
    const int rc2 = xmss_mi355x_verify_collect(j, ms, mlens, rets);
    xmss_mi355x_job_free(j);
    return rc ? rc : rc2;
}

}  // extern "C"
