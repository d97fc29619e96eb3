// verify.cpp -- signature verification, shared by both cores
// (drop-in for cuxmss/xmss_commons.c:120-215).
#include <stdio.h>
#include <string.h>

#include <vector>

#include "../../include/xmss_core.h"
#include "../../include/xmss_mi355x.h"
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
                   unsigned long long smlen, const unsigned char* pk, bool no_alloc) {
    if (smlen < p->sig_bytes || smlen - p->sig_bytes > kMaxMsgBytes) {
        *mlen = 0;
        return -1;
    }
    const int fam = fam_of(p->func, p->n);
    if (fam < 0) return -1;
    if (no_alloc && Engine::static_bytes(fam, p->tree_height, p->d, smlen - p->sig_bytes) > Engine::static_capacity()) {
        fprintf(stderr,
                "xmss_mi355x_sign_open_static: a %llu-byte message of this parameter set needs %zu bytes of "
                "device scratch, more than the %zu the library holds for it\n",
                smlen - p->sig_bytes, Engine::static_bytes(fam, p->tree_height, p->d, smlen - p->sig_bytes),
                Engine::static_capacity());
        *mlen = 0;
        return XMSS_ETOOBIG;
    }
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

        int ok = 0;
        if (no_alloc) {
            // key context and every buffer in the device's static scratch
            Engine::VerifyJob job{0, p->tree_height, p->d, p->index_bytes, pk, sm, m + p->sig_bytes, *mlen};
            if (e->verify_static(fam, job, &ok)) fatal("verify_static: scratch too small");
        } else {
            // PUB_SEED's context stays in the device pool across verifications
            e->pool_begin();
            const uint32_t slot = (uint32_t)e->pool_slot(fam, {nullptr, nullptr, pk + p->n});
            Engine::VerifyJob job{slot, p->tree_height, p->d, p->index_bytes, pk, sm, m + p->sig_bytes, *mlen};
            e->verify_many(fam, 1, &job, &ok);
        }
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
    return xm::core_sign_open(params, m, mlen, sm, smlen, pk, false);
}

int xmssmt_core_sign_open(const xmss_params* params, unsigned char* m, unsigned long long* mlen,
                          const unsigned char* sm, unsigned long long smlen, const unsigned char* pk) {
    return xm::core_sign_open(params, m, mlen, sm, smlen, pk, false);
}

int xmss_mi355x_sign_open_static(int mt, unsigned char* m, unsigned long long* mlen, const unsigned char* sm,
                                 unsigned long long smlen, const unsigned char* pk) {
    xmss_params p;
    uint32_t oid = 0;
    for (unsigned i = 0; i < XMSS_OID_LEN; ++i) oid = (oid << 8) | pk[i];
    if ((mt ? xmssmt_parse_oid(&p, oid) : xmss_parse_oid(&p, oid)) != 0) return -1;
    return xm::core_sign_open(&p, m, mlen, sm, smlen, pk + XMSS_OID_LEN, true);
}

}  // extern "C"
