/*
 * xmss_core.h -- core (OID-less) XMSS / XMSS^MT functions of libxmss_mi355x.
 *
 * Drop-in for cuxmss/xmss_core.h.  Keys here carry no OID prefix.
 */
#ifndef XMSS_CORE_H
#define XMSS_CORE_H

#include "params.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Secret-key size of the linked core.  Replaces cuxmss/xmss_core.h:17
 * (xmss_core.c:93-96 plain, xmss_core_fast.c:635-649 BDS). */
unsigned long long xmss_xmssmt_core_sk_bytes(const xmss_params *params);

/* Replaces cuxmss/xmss_core.h:24 (xmss_core.c:103-110, xmss_core_fast.c:656-701). */
int xmss_core_keypair(const xmss_params *params,
                      unsigned char *pk, unsigned char *sk);

/* Replaces cuxmss/xmss_core.h:31 (xmss_core.c:116-126, xmss_core_fast.c:710-851). */
int xmss_core_sign(const xmss_params *params,
                   unsigned char *sk,
                   unsigned char *sm, unsigned long long *smlen,
                   const unsigned char *m, unsigned long long mlen);

/* Replaces cuxmss/xmss_core.h:40 (xmss_commons.c:120-134). */
int xmss_core_sign_open(const xmss_params *params,
                        unsigned char *m, unsigned long long *mlen,
                        const unsigned char *sm, unsigned long long smlen,
                        const unsigned char *pk);

/* Replaces cuxmss/xmss_core.h:50 (xmss_core.c:168-177, xmss_core_fast.c:858-908). */
int xmssmt_core_keypair(const xmss_params *params,
                        unsigned char *pk, unsigned char *sk);

/* Plain core only, as in the reference.  seed = SK_SEED || SK_PRF || PUB_SEED.
 * Replaces cuxmss/xmss_core.h:59 (xmss_core.c:134-161). */
int xmssmt_core_seed_keypair(const xmss_params *params,
                             unsigned char *pk, unsigned char *sk,
                             unsigned char *seed);

/* Replaces cuxmss/xmss_core.h:68 (xmss_core.c:183-273, xmss_core_fast.c:917-1121). */
int xmssmt_core_sign(const xmss_params *params,
                     unsigned char *sk,
                     unsigned char *sm, unsigned long long *smlen,
                     const unsigned char *m, unsigned long long mlen);

/* Replaces cuxmss/xmss_core.h:77 (xmss_commons.c:136-215). */
int xmssmt_core_sign_open(const xmss_params *params,
                          unsigned char *m, unsigned long long *mlen,
                          const unsigned char *sm, unsigned long long smlen,
                          const unsigned char *pk);

#ifdef __cplusplus
}
#endif

#endif
