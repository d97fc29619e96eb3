/*
 * xmss.h -- OID-prefixed XMSS / XMSS^MT API of libxmss_mi355x.
 *
 * Drop-in for cuxmss/xmss.h (wzh009888/CU-XMSS).  All hashing runs on the GPU
 * (HIP, gfx950); the signatures, byte layouts and return codes are those of
 * the reference:
 *   pk = OID(4, big-endian) || root || PUB_SEED
 *   sk = OID || idx || SK_SEED || SK_PRF || root || PUB_SEED [|| BDS state]
 *   sm = signature || message
 * Return 0 on success, -1 for an unknown OID or a failed verification,
 * -2 for an exhausted key.  Additionally -3 when no GPU is usable (the library
 * has no CPU fallback).
 *
 * libxmss_mi355x.so implements the plain core (cuxmss/xmss_core.c);
 * libxmss_mi355x_fast.so the BDS core (cuxmss/xmss_core_fast.c), as the
 * reference selects them at link time.
 */
#ifndef XMSS_H
#define XMSS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces cuxmss/xmss.h:15 (xmss.c:10-26). */
int xmss_keypair(unsigned char *pk, unsigned char *sk, const uint32_t oid);

/* Replaces cuxmss/xmss.h:23 (xmss.c:28-43).  sk is updated in place; its
 * index advances before the signature is produced. */
int xmss_sign(unsigned char *sk,
              unsigned char *sm, unsigned long long *smlen,
              const unsigned char *m, unsigned long long mlen);

/* Replaces cuxmss/xmss.h:34 (xmss.c:45-60).  m must hold smlen bytes (it is
 * used as scratch, as in the reference). */
int xmss_sign_open(unsigned char *m, unsigned long long *mlen,
                   const unsigned char *sm, unsigned long long smlen,
                   const unsigned char *pk);

/* Replaces cuxmss/xmss.h:43 (xmss.c:62-75). */
int xmssmt_keypair(unsigned char *pk, unsigned char *sk, const uint32_t oid);

/* Replaces cuxmss/xmss.h:51 (xmss.c:77-92). */
int xmssmt_sign(unsigned char *sk,
                unsigned char *sm, unsigned long long *smlen,
                const unsigned char *m, unsigned long long mlen);

/* Replaces cuxmss/xmss.h:62 (xmss.c:94-109). */
int xmssmt_sign_open(unsigned char *m, unsigned long long *mlen,
                     const unsigned char *sm, unsigned long long smlen,
                     const unsigned char *pk);

#ifdef __cplusplus
}
#endif

#endif
