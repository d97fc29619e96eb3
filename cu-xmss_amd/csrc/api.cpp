// api.cpp -- the OID-prefixed C API (drop-in for cuxmss/xmss.c:10-109).
#include <stdint.h>

#include "../../include/params.h"
#include "../../include/xmss.h"
#include "../../include/xmss_core.h"

namespace {
uint32_t read_oid(const unsigned char* b) {
    uint32_t oid = 0;
    for (unsigned i = 0; i < XMSS_OID_LEN; ++i) oid |= (uint32_t)b[XMSS_OID_LEN - i - 1] << (i * 8);
    return oid;
}
void write_oid(unsigned char* b, uint32_t oid) {
    for (unsigned i = 0; i < XMSS_OID_LEN; ++i) b[XMSS_OID_LEN - i - 1] = (oid >> (8 * i)) & 0xFF;
}
}  // namespace

extern "C" {

int xmss_keypair(unsigned char* pk, unsigned char* sk, const uint32_t oid) {
    xmss_params p;
    if (xmss_parse_oid(&p, oid)) return -1;
    write_oid(pk, oid);
    write_oid(sk, oid);
    return xmss_core_keypair(&p, pk + XMSS_OID_LEN, sk + XMSS_OID_LEN);
}

int xmss_sign(unsigned char* sk, unsigned char* sm, unsigned long long* smlen, const unsigned char* m,
              unsigned long long mlen) {
    xmss_params p;
    if (xmss_parse_oid(&p, read_oid(sk))) return -1;
    return xmss_core_sign(&p, sk + XMSS_OID_LEN, sm, smlen, m, mlen);
}

int xmss_sign_open(unsigned char* m, unsigned long long* mlen, const unsigned char* sm, unsigned long long smlen,
                   const unsigned char* pk) {
    xmss_params p;
    if (xmss_parse_oid(&p, read_oid(pk))) return -1;
    return xmss_core_sign_open(&p, m, mlen, sm, smlen, pk + XMSS_OID_LEN);
}

int xmssmt_keypair(unsigned char* pk, unsigned char* sk, const uint32_t oid) {
    xmss_params p;
    if (xmssmt_parse_oid(&p, oid)) return -1;
    write_oid(pk, oid);
    write_oid(sk, oid);
    return xmssmt_core_keypair(&p, pk + XMSS_OID_LEN, sk + XMSS_OID_LEN);
}

int xmssmt_sign(unsigned char* sk, unsigned char* sm, unsigned long long* smlen, const unsigned char* m,
                unsigned long long mlen) {
    xmss_params p;
    if (xmssmt_parse_oid(&p, read_oid(sk))) return -1;
    return xmssmt_core_sign(&p, sk + XMSS_OID_LEN, sm, smlen, m, mlen);
}

int xmssmt_sign_open(unsigned char* m, unsigned long long* mlen, const unsigned char* sm,
                     unsigned long long smlen, const unsigned char* pk) {
    xmss_params p;
    if (xmssmt_parse_oid(&p, read_oid(pk))) return -1;
    return xmssmt_core_sign_open(&p, m, mlen, sm, smlen, pk + XMSS_OID_LEN);
}

}  // extern "C"
