/*
 * randombytes.h -- randomness source used by keypair generation.
 * Drop-in for cuxmss/randombytes.h:11 (randombytes.c:10-41 reads /dev/urandom;
 * so does this library unless xmss_mi355x_rng_seed() selected a seeded stream).
 */
#ifndef XMSS_RANDOMBYTES_H
#define XMSS_RANDOMBYTES_H

#ifdef __cplusplus
extern "C" {
#endif

void randombytes(unsigned char *x, unsigned long long xlen);

#ifdef __cplusplus
}
#endif

#endif
