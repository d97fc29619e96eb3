/*
 * params.h -- XMSS / XMSS^MT parameter sets of libxmss_mi355x.
 *
 * Drop-in for cuxmss/params.h (wzh009888/CU-XMSS): same struct layout, macro
 * values and functions, so code written against the reference compiles and
 * links unchanged.
 */
#ifndef XMSS_PARAMS_H
#define XMSS_PARAMS_H

#include <stdint.h>

/* Hash function identifiers (cuxmss/params.h:7-9). */
#define XMSS_SHA2 0
#define XMSS_SHAKE128 1
#define XMSS_SHAKE256 2

/* Bytes of the OID prefix on keys (cuxmss/params.h:12). */
#define XMSS_OID_LEN 4

/* Field order and types as cuxmss/params.h:15-33. */
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
} xmss_params;

#ifdef __cplusplus
extern "C" {
#endif

/* "XMSS-SHA2_10_256" -> 0x00000001 etc.; -1 if unknown.
 * Replaces cuxmss/params.h:41 (params.c:7-67). */
int xmss_str_to_oid(uint32_t *oid, const char *s);

/* "XMSSMT-SHA2_20/2_256" -> 0x00000001 etc.; -1 if unknown.
 * Replaces cuxmss/params.h:48 (params.c:69-252). */
int xmssmt_str_to_oid(uint32_t *oid, const char *s);

/* Fills params for an XMSS OID; -1 if unknown.
 * Replaces cuxmss/params.h:54 (params.c:254-372). */
int xmss_parse_oid(xmss_params *params, const uint32_t oid);

/* Fills params for an XMSS^MT OID; -1 if unknown.
 * Replaces cuxmss/params.h:60 (params.c:374-698). */
int xmssmt_parse_oid(xmss_params *params, const uint32_t oid);

/* Derives the remaining fields from full_height, n, d, func, wots_w (and
 * bds_k).  Replaces cuxmss/params.h:72 (params.c:710-753). */
int xmss_xmssmt_initialize_params(xmss_params *params);

#ifdef __cplusplus
}
#endif

#endif
