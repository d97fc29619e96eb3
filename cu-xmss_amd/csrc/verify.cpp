// verify.cpp -- signature verification, shared by both cores
// (drop-in for cuxmss/xmss_commons.c:120-215).
#include <string.h>

#include <vector>

#include "../../include/xmss_core.h"
#include "core_common.h"

namespace xm {

std::mutex& engine_mutex() {
    static std::mutex mu;
    return mu;
}

// xmssmt_core_sign_open (xmss_commons.c:136-215).  m must hold smlen bytes:
// as in the reference, the message is first copied to m + sig_bytes and the
// H_msg prefix written in front of it.  Deviations: smlen < sig_bytes is
// rejected (-1, *mlen = 0) instead of underflowing *mlen (:159), and so is a
// message longer than kMaxMsgBytes.
int core_sign_open(const xmss_params* p, unsigned char* m, unsigned long long* mlen, const unsigned char* sm,
                   unsigned long long smlen, const unsigned char* pk) {
    if (smlen < p->sig_bytes || smlen - p->sig_bytes > kMaxMsgBytes) {
        *mlen = 0;
        return -1;
    }
    const int fam = fam_of(p->func, p->n);
    if (fam < 0) return -1;
    return guarded([&] {
        std::lock_guard<std::mutex> lk(engine_mutex());
        Engine* e = Engine::get();
        if (!e) return XMSS_ENODEV;
        *mlen = smlen - p->sig_bytes;
        const unsigned long long idx = bytes_to_ull(sm, p->index_bytes);
        memmove(m + p->sig_bytes, sm + p->sig_bytes, *mlen);
        // hash_message's in-place prefix (hash.c:114-117), kept for byte parity of m
        unsigned char* pre = m + p->sig_bytes - p->padding_len - 3 * p->n;
        ull_to_bytes(pre, p->padding_len, 2);
        memcpy(pre + p->padding_len, sm + p->index_bytes, p->n);
        memcpy(pre + p->padding_len + p->n, pk, p->n);
        ull_to_bytes(pre + p->padding_len + 2 * p->n, p->n, idx);

        // PUB_SEED's context stays in the device pool across verifications
        e->pool_begin();
        const uint32_t slot = (uint32_t)e->pool_slot(fam, {nullptr, nullptr, pk + p->n});
        Engine::VerifyJob job{slot, p->tree_height, p->d, p->index_bytes, pk, sm, m + p->sig_bytes, *mlen};
        int ok = 0;
        e->verify_many(fam, 1, &job, &ok);
        if (!ok) {
            memset(m, 0, *mlen);
            *mlen = 0;
            return -1;
        }
        memmove(m, sm + p->sig_bytes, *mlen);
        return 0;
    });
}

}  // namespace xm

extern "C" {

int xmss_core_sign_open(const xmss_params* params, unsigned char* m, unsigned long long* mlen,
                        const unsigned char* sm, unsigned long long smlen, const unsigned char* pk) {
    return xm::core_sign_open(params, m, mlen, sm, smlen, pk);
}

int xmssmt_core_sign_open(const xmss_params* params, unsigned char* m, unsigned long long* mlen,
                          const unsigned char* sm, unsigned long long smlen, const unsigned char* pk) {
    return xm::core_sign_open(params, m, mlen, sm, smlen, pk);
}

}  // extern "C"
