/*
 * xo.h -- CPU oracle for the XMSS / XMSS^MT hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (the C files of cuxmss/ in wzh009888/CU-XMSS, itself RFC 8391 + NIST
 * SP 800-208 parameter sets).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker.  The
 * product library (libxmss_mi355x*.so) never links or calls it.
 *
 * Every function carries the reference file:line it follows.  Hash primitives
 * follow FIPS 180-4 (SHA-256/512) and FIPS 202 (Keccak-f1600, SHAKE).
 *
 * Parity pins (which check fixes which function) are listed in the header of
 * xo.c.
 */
#ifndef XO_H
#define XO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XO_SHA2 0
#define XO_SHAKE128 1
#define XO_SHAKE256 2

/* Same field order and meaning as xmss_params (params.h:15-33). */
typedef struct {
    unsigned int func;
    unsigned int n;
    unsigned int padding_len;
    unsigned int wots_w;
    unsigned int wots_log_w;
    unsigned int wots_len1;
    unsigned int wots_len2;
    unsigned int wots_len;
    unsigned int wots_sig_bytes;
    unsigned int full_height;
    unsigned int tree_height;
    unsigned int d;
    unsigned int index_bytes;
    unsigned int sig_bytes;
    unsigned int pk_bytes;
    unsigned long long sk_bytes;
    unsigned int bds_k;
} xo_params;

/* ---- hash primitives (FIPS 180-4 / FIPS 202) ---- */
void xo_sha256(unsigned char out[32], const unsigned char *in, unsigned long long inlen);
void xo_sha512(unsigned char out[64], const unsigned char *in, unsigned long long inlen);
void xo_shake128(unsigned char *out, unsigned long long outlen,
                 const unsigned char *in, unsigned long long inlen);
void xo_shake256(unsigned char *out, unsigned long long outlen,
                 const unsigned char *in, unsigned long long inlen);
void xo_keccak_f1600(uint64_t s[25]);

/* ---- seeded randombytes stream (SplitMix64, little-endian bytes) ---- */
void xo_rng_seed(uint64_t seed);
void xo_randombytes(unsigned char *x, unsigned long long xlen);

/* ---- parameter sets (params.c) ---- */
int xo_xmss_str_to_oid(uint32_t *oid, const char *s);
int xo_xmssmt_str_to_oid(uint32_t *oid, const char *s);
/* core: 0 = plain core (xmss_core.c), 1 = BDS core (xmss_core_fast.c);
 * it only changes sk_bytes. */
int xo_xmss_parse_oid(xo_params *p, uint32_t oid, int core);
int xo_xmssmt_parse_oid(xo_params *p, uint32_t oid, int core);
int xo_initialize_params(xo_params *p, int core);
unsigned long long xo_sk_bytes(const xo_params *p, int core);

/* ---- tweakable hash layer (hash.c) ---- */
void xo_addr_to_bytes(unsigned char *bytes, const uint32_t addr[8]);
int xo_core_hash(const xo_params *p, unsigned char *out,
                 const unsigned char *in, unsigned long long inlen);
int xo_prf(const xo_params *p, unsigned char *out, const unsigned char in[32],
           const unsigned char *key);
int xo_prf_keygen(const xo_params *p, unsigned char *out,
                  const unsigned char *in, const unsigned char *key);
int xo_hash_message(const xo_params *p, unsigned char *out,
                    const unsigned char *R, const unsigned char *root,
                    unsigned long long idx, unsigned char *m_with_prefix,
                    unsigned long long mlen);
int xo_thash_h(const xo_params *p, unsigned char *out, const unsigned char *in,
               const unsigned char *pub_seed, uint32_t addr[8]);
int xo_thash_f(const xo_params *p, unsigned char *out, const unsigned char *in,
               const unsigned char *pub_seed, uint32_t addr[8]);

/* ---- WOTS+ (wots.c) ---- */
void xo_chain_lengths(const xo_params *p, int *lengths, const unsigned char *msg);
void xo_wots_pkgen(const xo_params *p, unsigned char *pk, const unsigned char *seed,
                   const unsigned char *pub_seed, uint32_t addr[8]);
void xo_wots_sign(const xo_params *p, unsigned char *sig, const unsigned char *msg,
                  const unsigned char *seed, const unsigned char *pub_seed,
                  uint32_t addr[8]);
void xo_wots_pk_from_sig(const xo_params *p, unsigned char *pk,
                         const unsigned char *sig, const unsigned char *msg,
                         const unsigned char *pub_seed, uint32_t addr[8]);

/* ---- Merkle helpers (xmss_commons.c) ---- */
void xo_l_tree(const xo_params *p, unsigned char *leaf, unsigned char *wots_pk,
               const unsigned char *pub_seed, uint32_t addr[8]);
void xo_compute_root(const xo_params *p, unsigned char *root,
                     const unsigned char *leaf, unsigned long leafidx,
                     const unsigned char *auth_path,
                     const unsigned char *pub_seed, uint32_t addr[8]);
void xo_gen_leaf_wots(const xo_params *p, unsigned char *leaf,
                      const unsigned char *sk_seed, const unsigned char *pub_seed,
                      uint32_t ltree_addr[8], uint32_t ots_addr[8]);
int xo_sign_open(const xo_params *p, unsigned char *m, unsigned long long *mlen,
                 const unsigned char *sm, unsigned long long smlen,
                 const unsigned char *pk);

/* ---- plain core (xmss_core.c) ---- */
void xo_treehash(const xo_params *p, unsigned char *root, unsigned char *auth_path,
                 const unsigned char *sk_seed, const unsigned char *pub_seed,
                 uint32_t leaf_idx, const uint32_t subtree_addr[8]);
void xo_subtree_root(const xo_params *p, unsigned char *root, const unsigned char *sk_seed,
                     const unsigned char *pub_seed, uint32_t leaf0, unsigned height,
                     const uint32_t subtree_addr[8]);
int xo_plain_seed_keypair(const xo_params *p, unsigned char *pk, unsigned char *sk,
                          const unsigned char *seed);
int xo_plain_keypair(const xo_params *p, unsigned char *pk, unsigned char *sk);
int xo_plain_sign(const xo_params *p, unsigned char *sk, unsigned char *sm,
                  unsigned long long *smlen, const unsigned char *m,
                  unsigned long long mlen);

/* ---- BDS core (xmss_core_fast.c) ---- */
int xo_bds_xmss_keypair(const xo_params *p, unsigned char *pk, unsigned char *sk);
int xo_bds_xmssmt_keypair(const xo_params *p, unsigned char *pk, unsigned char *sk);
int xo_bds_xmss_sign(const xo_params *p, unsigned char *sk, unsigned char *sm,
                     unsigned long long *smlen, const unsigned char *m,
                     unsigned long long mlen);
int xo_bds_xmssmt_sign(const xo_params *p, unsigned char *sk, unsigned char *sm,
                       unsigned long long *smlen, const unsigned char *m,
                       unsigned long long mlen);

/* ---- OID-prefixed API (xmss.c); mt selects xmssmt_*, core as above ---- */
int xo_api_keypair(unsigned char *pk, unsigned char *sk, uint32_t oid, int mt, int core);
int xo_api_sign(unsigned char *sk, unsigned char *sm, unsigned long long *smlen,
                const unsigned char *m, unsigned long long mlen, int mt, int core);
int xo_api_sign_open(unsigned char *m, unsigned long long *mlen,
                     const unsigned char *sm, unsigned long long smlen,
                     const unsigned char *pk, int mt);

#ifdef __cplusplus
}
#endif
#endif
