/*
 * xo_xmss.c -- oracle restatement of the XMSS / XMSS^MT path.
 * TEST INFRASTRUCTURE ONLY (see xo.h): never linked into the product.
 *
 * Follows the reference C files function by function (file:line cited on each
 * function).  The reference is RFC 8391 plus SP 800-208 parameter sets; its
 * quirks are kept on purpose (index edge cases, BDS fields that are carried
 * through the caller's sk buffer untouched, the char-typed update counter).
 *
 * Parity pins (tests/test_oracle.py):
 *  - params / OIDs: every name of cuxmss/params.c, sizes against the
 *    reference's own formulas and golden sizes; oracle/_ref byte parity.
 *  - prf / thash / WOTS / L-tree / leaf: the reference's KAT emitter output
 *    (test/vectors.c, 7 WOTS+ and 14 XMSS/XMSSMT digests), golden vectors
 *    written by oracle/_ref, and wots_pkgen == pk_from_sig(sign).
 *  - treehash / plain keypair / plain sign / verify: golden pk/sk/sm bytes
 *    from oracle/_ref, brute-force root of tiny trees built from gen_leaf
 *    and thash_h directly, sign->verify round trips and bit-flip rejection.
 *  - BDS core: golden sk-state evolution from oracle/_ref (xmss_core_fast.c),
 *    and BDS signature bytes == plain signature bytes for the same seeds.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "xo.h"

#define PAD_F 0
#define PAD_H 1
#define PAD_HASH 2
#define PAD_PRF 3
#define PAD_PRF_KEYGEN 4

#define ADDR_OTS 0
#define ADDR_LTREE 1
#define ADDR_HASHTREE 2

/* ------------------------------------------------------------------ */
/* byte helpers: utils.c:6-30                                          */
/* ------------------------------------------------------------------ */

static void to_bytes(unsigned char *out, unsigned int outlen, unsigned long long v)
{
    int i;
    for (i = (int)outlen - 1; i >= 0; i--) {
        out[i] = (unsigned char)(v & 0xff);
        v >>= 8;
    }
}

static unsigned long long from_bytes(const unsigned char *in, unsigned int inlen)
{
    unsigned long long r = 0;
    unsigned int i;
    for (i = 0; i < inlen; i++) {
        r |= ((unsigned long long)in[i]) << (8 * (inlen - 1 - i));
    }
    return r;
}

/* ------------------------------------------------------------------ */
/* parameter sets: params.c:7-252 (names), :254-698 (OIDs), :710-753   */
/* ------------------------------------------------------------------ */

typedef struct {
    const char *name;
    uint32_t oid;
    unsigned char func, n, h, d;
} oid_entry;

/* XMSS: params.c:9-67 and :256-372 */
static const oid_entry XMSS_SETS[] = {
    {"XMSS-SHA2_10_256", 0x01, XO_SHA2, 32, 10, 1},
    {"XMSS-SHA2_16_256", 0x02, XO_SHA2, 32, 16, 1},
    {"XMSS-SHA2_20_256", 0x03, XO_SHA2, 32, 20, 1},
    {"XMSS-SHA2_10_512", 0x04, XO_SHA2, 64, 10, 1},
    {"XMSS-SHA2_16_512", 0x05, XO_SHA2, 64, 16, 1},
    {"XMSS-SHA2_20_512", 0x06, XO_SHA2, 64, 20, 1},
    {"XMSS-SHAKE_10_256", 0x07, XO_SHAKE128, 32, 10, 1},
    {"XMSS-SHAKE_16_256", 0x08, XO_SHAKE128, 32, 16, 1},
    {"XMSS-SHAKE_20_256", 0x09, XO_SHAKE128, 32, 20, 1},
    {"XMSS-SHAKE_10_512", 0x0a, XO_SHAKE256, 64, 10, 1},
    {"XMSS-SHAKE_16_512", 0x0b, XO_SHAKE256, 64, 16, 1},
    {"XMSS-SHAKE_20_512", 0x0c, XO_SHAKE256, 64, 20, 1},
    {"XMSS-SHA2_10_192", 0x0d, XO_SHA2, 24, 10, 1},
    {"XMSS-SHA2_16_192", 0x0e, XO_SHA2, 24, 16, 1},
    {"XMSS-SHA2_20_192", 0x0f, XO_SHA2, 24, 20, 1},
    {"XMSS-SHAKE256_10_256", 0x10, XO_SHAKE256, 32, 10, 1},
    {"XMSS-SHAKE256_16_256", 0x11, XO_SHAKE256, 32, 16, 1},
    {"XMSS-SHAKE256_20_256", 0x12, XO_SHAKE256, 32, 20, 1},
    {"XMSS-SHAKE256_10_192", 0x13, XO_SHAKE256, 24, 10, 1},
    {"XMSS-SHAKE256_16_192", 0x14, XO_SHAKE256, 24, 16, 1},
    {"XMSS-SHAKE256_20_192", 0x15, XO_SHAKE256, 24, 20, 1},
};

/* XMSS^MT: params.c:70-252 and :374-698.  OIDs run in blocks of eight
 * heights/layers (20/2, 20/4, 40/2, 40/4, 40/8, 60/3, 60/6, 60/12) over
 * seven (func, n) families. */
static const struct {
    const char *prefix;
    const char *suffix;
    unsigned char func, n;
    uint32_t first_oid;
} MT_FAMILIES[] = {
    {"XMSSMT-SHA2_", "_256", XO_SHA2, 32, 0x01},
    {"XMSSMT-SHA2_", "_512", XO_SHA2, 64, 0x09},
    {"XMSSMT-SHAKE_", "_256", XO_SHAKE128, 32, 0x11},
    {"XMSSMT-SHAKE_", "_512", XO_SHAKE256, 64, 0x19},
    {"XMSSMT-SHA2_", "_192", XO_SHA2, 24, 0x21},
    {"XMSSMT-SHAKE256_", "_256", XO_SHAKE256, 32, 0x29},
    {"XMSSMT-SHAKE256_", "_192", XO_SHAKE256, 24, 0x31},
};
static const struct {
    const char *hd;
    unsigned char h, d;
} MT_SHAPES[8] = {
    {"20/2", 20, 2}, {"20/4", 20, 4}, {"40/2", 40, 2}, {"40/4", 40, 4},
    {"40/8", 40, 8}, {"60/3", 60, 3}, {"60/6", 60, 6}, {"60/12", 60, 12},
};

static int mt_lookup_oid(uint32_t oid, oid_entry *e)
{
    unsigned int f;
    for (f = 0; f < sizeof(MT_FAMILIES) / sizeof(MT_FAMILIES[0]); f++) {
        if (oid >= MT_FAMILIES[f].first_oid && oid < MT_FAMILIES[f].first_oid + 8) {
            unsigned int s = oid - MT_FAMILIES[f].first_oid;
            e->name = NULL;
            e->oid = oid;
            e->func = MT_FAMILIES[f].func;
            e->n = MT_FAMILIES[f].n;
            e->h = MT_SHAPES[s].h;
            e->d = MT_SHAPES[s].d;
            return 0;
        }
    }
    return -1;
}

int xo_xmss_str_to_oid(uint32_t *oid, const char *s)
{
    unsigned int i;
    for (i = 0; i < sizeof(XMSS_SETS) / sizeof(XMSS_SETS[0]); i++) {
        if (!strcmp(s, XMSS_SETS[i].name)) {
            *oid = XMSS_SETS[i].oid;
            return 0;
        }
    }
    return -1;
}

int xo_xmssmt_str_to_oid(uint32_t *oid, const char *s)
{
    unsigned int f, k;
    char buf[64];
    for (f = 0; f < sizeof(MT_FAMILIES) / sizeof(MT_FAMILIES[0]); f++) {
        for (k = 0; k < 8; k++) {
            strcpy(buf, MT_FAMILIES[f].prefix);
            strcat(buf, MT_SHAPES[k].hd);
            strcat(buf, MT_FAMILIES[f].suffix);
            if (!strcmp(s, buf)) {
                *oid = MT_FAMILIES[f].first_oid + k;
                return 0;
            }
        }
    }
    return -1;
}

/* xmss_xmssmt_core_sk_bytes: xmss_core.c:93-96 (plain),
 * xmss_core_fast.c:635-649 (BDS). */
unsigned long long xo_sk_bytes(const xo_params *p, int core)
{
    if (!core) {
        return p->index_bytes + 4 * p->n;
    }
    return p->index_bytes + 4 * p->n
           + (2 * p->d - 1) * ((p->tree_height + 1) * p->n + 4 + p->tree_height + 1
                               + p->tree_height * p->n + (p->tree_height >> 1) * p->n
                               + (p->tree_height - p->bds_k) * (7 + p->n)
                               + ((1 << p->bds_k) - p->bds_k - 1) * p->n + 4)
           + (p->d - 1) * p->wots_sig_bytes;
}

/* xmss_xmssmt_initialize_params: params.c:710-753 */
int xo_initialize_params(xo_params *p, int core)
{
    p->tree_height = p->full_height / p->d;
    if (p->wots_w == 4) {
        p->wots_log_w = 2;
        p->wots_len2 = 5;
    } else if (p->wots_w == 16) {
        p->wots_log_w = 4;
        p->wots_len2 = 3;
    } else if (p->wots_w == 256) {
        p->wots_log_w = 8;
        p->wots_len2 = 2;
    } else {
        return -1;
    }
    p->wots_len1 = 8 * p->n / p->wots_log_w;
    p->wots_len = p->wots_len1 + p->wots_len2;
    p->wots_sig_bytes = p->wots_len * p->n;
    p->index_bytes = (p->d == 1) ? 4 : (p->full_height + 7) / 8;
    p->sig_bytes = p->index_bytes + p->n + p->d * p->wots_sig_bytes + p->full_height * p->n;
    p->pk_bytes = 2 * p->n;
    p->sk_bytes = xo_sk_bytes(p, core);
    return 0;
}

static int fill_params(xo_params *p, const oid_entry *e, int core)
{
    p->func = e->func;
    p->n = e->n;
    p->padding_len = (e->n == 24) ? 4 : e->n;
    p->full_height = e->h;
    p->d = e->d;
    p->wots_w = 16;
    p->bds_k = 0;
    return xo_initialize_params(p, core);
}

/* xmss_parse_oid: params.c:254-372 */
int xo_xmss_parse_oid(xo_params *p, uint32_t oid, int core)
{
    unsigned int i;
    for (i = 0; i < sizeof(XMSS_SETS) / sizeof(XMSS_SETS[0]); i++) {
        if (XMSS_SETS[i].oid == oid) {
            return fill_params(p, &XMSS_SETS[i], core);
        }
    }
    return -1;
}

/* xmssmt_parse_oid: params.c:374-698 */
int xo_xmssmt_parse_oid(xo_params *p, uint32_t oid, int core)
{
    oid_entry e;
    if (mt_lookup_oid(oid, &e)) {
        return -1;
    }
    return fill_params(p, &e, core);
}

/* ------------------------------------------------------------------ */
/* ADRS + tweakable hash: hash_address.c:3-65, hash.c:19-183           */
/* ------------------------------------------------------------------ */

void xo_addr_to_bytes(unsigned char *bytes, const uint32_t addr[8])
{
    int i;
    for (i = 0; i < 8; i++) {
        to_bytes(bytes + 4 * i, 4, addr[i]);
    }
}

static void set_tree_addr(uint32_t a[8], uint64_t tree)
{
    a[1] = (uint32_t)(tree >> 32);
    a[2] = (uint32_t)tree;
}

static void copy_subtree(uint32_t out[8], const uint32_t in[8])
{
    out[0] = in[0];
    out[1] = in[1];
    out[2] = in[2];
}

/* core_hash: hash.c:29-65 */
int xo_core_hash(const xo_params *p, unsigned char *out,
                 const unsigned char *in, unsigned long long inlen)
{
    unsigned char buf[64];
    if (p->n == 24 && p->func == XO_SHA2) {
        xo_sha256(buf, in, inlen);
        memcpy(out, buf, 24);
    } else if (p->n == 24 && p->func == XO_SHAKE256) {
        xo_shake256(out, 24, in, inlen);
    } else if (p->n == 32 && p->func == XO_SHA2) {
        xo_sha256(out, in, inlen);
    } else if (p->n == 32 && p->func == XO_SHAKE128) {
        xo_shake128(out, 32, in, inlen);
    } else if (p->n == 32 && p->func == XO_SHAKE256) {
        xo_shake256(out, 32, in, inlen);
    } else if (p->n == 64 && p->func == XO_SHA2) {
        xo_sha512(out, in, inlen);
    } else if (p->n == 64 && p->func == XO_SHAKE256) {
        xo_shake256(out, 64, in, inlen);
    } else {
        return -1;
    }
    return 0;
}

/* prf: hash.c:70-81 */
int xo_prf(const xo_params *p, unsigned char *out, const unsigned char in[32],
           const unsigned char *key)
{
    unsigned char buf[64 + 64 + 32];
    to_bytes(buf, p->padding_len, PAD_PRF);
    memcpy(buf + p->padding_len, key, p->n);
    memcpy(buf + p->padding_len + p->n, in, 32);
    return xo_core_hash(p, out, buf, p->padding_len + p->n + 32);
}

/* prf_keygen: hash.c:87-98 */
int xo_prf_keygen(const xo_params *p, unsigned char *out,
                  const unsigned char *in, const unsigned char *key)
{
    unsigned char buf[64 + 2 * 64 + 32];
    to_bytes(buf, p->padding_len, PAD_PRF_KEYGEN);
    memcpy(buf + p->padding_len, key, p->n);
    memcpy(buf + p->padding_len + p->n, in, p->n + 32);
    return xo_core_hash(p, out, buf, p->padding_len + 2 * p->n + 32);
}

/* hash_message: hash.c:107-120.  Writes its prefix in place before M. */
int xo_hash_message(const xo_params *p, unsigned char *out,
                    const unsigned char *R, const unsigned char *root,
                    unsigned long long idx, unsigned char *m_with_prefix,
                    unsigned long long mlen)
{
    to_bytes(m_with_prefix, p->padding_len, PAD_HASH);
    memcpy(m_with_prefix + p->padding_len, R, p->n);
    memcpy(m_with_prefix + p->padding_len + p->n, root, p->n);
    to_bytes(m_with_prefix + p->padding_len + 2 * p->n, p->n, idx);
    return xo_core_hash(p, out, m_with_prefix, mlen + p->padding_len + 3 * p->n);
}

/* thash_h: hash.c:125-155.  Sets addr[7] (key_and_mask) as a side effect. */
int xo_thash_h(const xo_params *p, unsigned char *out, const unsigned char *in,
               const unsigned char *pub_seed, uint32_t addr[8])
{
    unsigned char buf[64 + 3 * 64], mask[2 * 64], ab[32];
    unsigned int i;
    to_bytes(buf, p->padding_len, PAD_H);
    addr[7] = 0;
    xo_addr_to_bytes(ab, addr);
    xo_prf(p, buf + p->padding_len, ab, pub_seed);
    addr[7] = 1;
    xo_addr_to_bytes(ab, addr);
    xo_prf(p, mask, ab, pub_seed);
    addr[7] = 2;
    xo_addr_to_bytes(ab, addr);
    xo_prf(p, mask + p->n, ab, pub_seed);
    for (i = 0; i < 2 * p->n; i++) {
        buf[p->padding_len + p->n + i] = in[i] ^ mask[i];
    }
    return xo_core_hash(p, out, buf, p->padding_len + 3 * p->n);
}

/* thash_f: hash.c:157-183.  out may alias in. */
int xo_thash_f(const xo_params *p, unsigned char *out, const unsigned char *in,
               const unsigned char *pub_seed, uint32_t addr[8])
{
    unsigned char buf[64 + 2 * 64], mask[64], ab[32];
    unsigned int i;
    to_bytes(buf, p->padding_len, PAD_F);
    addr[7] = 0;
    xo_addr_to_bytes(ab, addr);
    xo_prf(p, buf + p->padding_len, ab, pub_seed);
    addr[7] = 1;
    xo_addr_to_bytes(ab, addr);
    xo_prf(p, mask, ab, pub_seed);
    for (i = 0; i < p->n; i++) {
        buf[p->padding_len + p->n + i] = in[i] ^ mask[i];
    }
    return xo_core_hash(p, out, buf, p->padding_len + 2 * p->n);
}

/* ------------------------------------------------------------------ */
/* WOTS+: wots.c:14-175                                                */
/* ------------------------------------------------------------------ */

/* expand_seed: wots.c:14-29 */
static void expand_seed(const xo_params *p, unsigned char *outseeds,
                        const unsigned char *inseed, const unsigned char *pub_seed,
                        uint32_t addr[8])
{
    unsigned char buf[64 + 32];
    uint32_t i;
    addr[6] = 0;
    addr[7] = 0;
    memcpy(buf, pub_seed, p->n);
    for (i = 0; i < p->wots_len; i++) {
        addr[5] = i;
        xo_addr_to_bytes(buf + p->n, addr);
        xo_prf_keygen(p, outseeds + i * p->n, buf, inseed);
    }
}

/* gen_chain: wots.c:38-53 */
static void gen_chain(const xo_params *p, unsigned char *out, const unsigned char *in,
                      unsigned int start, unsigned int steps,
                      const unsigned char *pub_seed, uint32_t addr[8])
{
    uint32_t i;
    memmove(out, in, p->n);
    for (i = start; i < start + steps && i < p->wots_w; i++) {
        addr[6] = i;
        xo_thash_f(p, out, out, pub_seed, addr);
    }
}

/* base_w: wots.c:60-79 */
static void base_w(const xo_params *p, int *output, int out_len,
                   const unsigned char *input)
{
    int in = 0, out, bits = 0;
    unsigned char total = 0;
    for (out = 0; out < out_len; out++) {
        if (bits == 0) {
            total = input[in++];
            bits += 8;
        }
        bits -= (int)p->wots_log_w;
        output[out] = (total >> bits) & (int)(p->wots_w - 1);
    }
}

/* wots_checksum + chain_lengths: wots.c:82-107 */
void xo_chain_lengths(const xo_params *p, int *lengths, const unsigned char *msg)
{
    int csum = 0;
    unsigned char csum_bytes[8];
    unsigned int i, nbytes = (p->wots_len2 * p->wots_log_w + 7) / 8;
    base_w(p, lengths, (int)p->wots_len1, msg);
    for (i = 0; i < p->wots_len1; i++) {
        csum += (int)p->wots_w - 1 - lengths[i];
    }
    csum = csum << (8 - ((p->wots_len2 * p->wots_log_w) % 8));
    to_bytes(csum_bytes, nbytes, (unsigned long long)csum);
    base_w(p, lengths + p->wots_len1, (int)p->wots_len2, csum_bytes);
}

/* wots_pkgen: wots.c:117-131 */
void xo_wots_pkgen(const xo_params *p, unsigned char *pk, const unsigned char *seed,
                   const unsigned char *pub_seed, uint32_t addr[8])
{
    uint32_t i;
    expand_seed(p, pk, seed, pub_seed, addr);
    for (i = 0; i < p->wots_len; i++) {
        addr[5] = i;
        gen_chain(p, pk + i * p->n, pk + i * p->n, 0, p->wots_w - 1, pub_seed, addr);
    }
}

/* wots_sign: wots.c:136-154 */
void xo_wots_sign(const xo_params *p, unsigned char *sig, const unsigned char *msg,
                  const unsigned char *seed, const unsigned char *pub_seed,
                  uint32_t addr[8])
{
    int lengths[256];
    uint32_t i;
    xo_chain_lengths(p, lengths, msg);
    expand_seed(p, sig, seed, pub_seed, addr);
    for (i = 0; i < p->wots_len; i++) {
        addr[5] = i;
        gen_chain(p, sig + i * p->n, sig + i * p->n, 0, (unsigned int)lengths[i], pub_seed, addr);
    }
}

/* wots_pk_from_sig: wots.c:161-175 */
void xo_wots_pk_from_sig(const xo_params *p, unsigned char *pk,
                         const unsigned char *sig, const unsigned char *msg,
                         const unsigned char *pub_seed, uint32_t addr[8])
{
    int lengths[256];
    uint32_t i;
    xo_chain_lengths(p, lengths, msg);
    for (i = 0; i < p->wots_len; i++) {
        addr[5] = i;
        gen_chain(p, pk + i * p->n, sig + i * p->n, (unsigned int)lengths[i],
                  p->wots_w - 1 - (unsigned int)lengths[i], pub_seed, addr);
    }
}

/* ------------------------------------------------------------------ */
/* Merkle helpers: xmss_commons.c:16-215                               */
/* ------------------------------------------------------------------ */

/* l_tree: xmss_commons.c:16-49 (destroys wots_pk) */
void xo_l_tree(const xo_params *p, unsigned char *leaf, unsigned char *wots_pk,
               const unsigned char *pub_seed, uint32_t addr[8])
{
    unsigned int l = p->wots_len, parents, i;
    uint32_t height = 0;
    addr[5] = height;
    while (l > 1) {
        parents = l >> 1;
        for (i = 0; i < parents; i++) {
            addr[6] = i;
            xo_thash_h(p, wots_pk + i * p->n, wots_pk + 2 * i * p->n, pub_seed, addr);
        }
        if (l & 1) {
            memcpy(wots_pk + (l >> 1) * p->n, wots_pk + (l - 1) * p->n, p->n);
            l = (l >> 1) + 1;
        } else {
            l >>= 1;
        }
        height++;
        addr[5] = height;
    }
    memcpy(leaf, wots_pk, p->n);
}

/* compute_root: xmss_commons.c:54-96 */
void xo_compute_root(const xo_params *p, unsigned char *root,
                     const unsigned char *leaf, unsigned long leafidx,
                     const unsigned char *auth_path,
                     const unsigned char *pub_seed, uint32_t addr[8])
{
    unsigned char buffer[2 * 64];
    uint32_t i;
    if (leafidx & 1) {
        memcpy(buffer + p->n, leaf, p->n);
        memcpy(buffer, auth_path, p->n);
    } else {
        memcpy(buffer, leaf, p->n);
        memcpy(buffer + p->n, auth_path, p->n);
    }
    auth_path += p->n;
    for (i = 0; i < p->tree_height - 1; i++) {
        addr[5] = i;
        leafidx >>= 1;
        addr[6] = (uint32_t)leafidx;
        if (leafidx & 1) {
            xo_thash_h(p, buffer + p->n, buffer, pub_seed, addr);
            memcpy(buffer, auth_path, p->n);
        } else {
            xo_thash_h(p, buffer, buffer, pub_seed, addr);
            memcpy(buffer + p->n, auth_path, p->n);
        }
        auth_path += p->n;
    }
    addr[5] = p->tree_height - 1;
    leafidx >>= 1;
    addr[6] = (uint32_t)leafidx;
    xo_thash_h(p, root, buffer, pub_seed, addr);
}

/* gen_leaf_wots: xmss_commons.c:104-113 */
void xo_gen_leaf_wots(const xo_params *p, unsigned char *leaf,
                      const unsigned char *sk_seed, const unsigned char *pub_seed,
                      uint32_t ltree_addr[8], uint32_t ots_addr[8])
{
    unsigned char pk[256 * 64];
    xo_wots_pkgen(p, pk, sk_seed, pub_seed, ots_addr);
    xo_l_tree(p, leaf, pk, pub_seed, ltree_addr);
}

/* xmssmt_core_sign_open: xmss_commons.c:136-215.  As in the reference, m must
 * hold sig_bytes + mlen bytes (it is used as scratch for the H_msg prefix). */
int xo_sign_open(const xo_params *p, unsigned char *m, unsigned long long *mlen,
                 const unsigned char *sm, unsigned long long smlen,
                 const unsigned char *pk)
{
    const unsigned char *pub_root = pk, *pub_seed = pk + p->n;
    unsigned char wots_pk[256 * 64], leaf[64], root[64];
    unsigned long long idx;
    unsigned int i;
    uint32_t idx_leaf;
    uint32_t ots_addr[8] = {0}, ltree_addr[8] = {0}, node_addr[8] = {0};

    ots_addr[3] = ADDR_OTS;
    ltree_addr[3] = ADDR_LTREE;
    node_addr[3] = ADDR_HASHTREE;

    *mlen = smlen - p->sig_bytes;
    idx = from_bytes(sm, p->index_bytes);
    memcpy(m + p->sig_bytes, sm + p->sig_bytes, *mlen);
    xo_hash_message(p, root, sm + p->index_bytes, pk, idx,
                    m + p->sig_bytes - p->padding_len - 3 * p->n, *mlen);
    sm += p->index_bytes + p->n;

    for (i = 0; i < p->d; i++) {
        idx_leaf = (uint32_t)(idx & ((1u << p->tree_height) - 1));
        idx >>= p->tree_height;
        ots_addr[0] = ltree_addr[0] = node_addr[0] = i;
        set_tree_addr(ltree_addr, idx);
        set_tree_addr(ots_addr, idx);
        set_tree_addr(node_addr, idx);
        ots_addr[4] = idx_leaf;
        xo_wots_pk_from_sig(p, wots_pk, sm, root, pub_seed, ots_addr);
        sm += p->wots_sig_bytes;
        ltree_addr[4] = idx_leaf;
        xo_l_tree(p, leaf, wots_pk, pub_seed, ltree_addr);
        xo_compute_root(p, root, leaf, idx_leaf, sm, pub_seed, node_addr);
        sm += p->tree_height * p->n;
    }
    if (memcmp(root, pub_root, p->n)) {
        memset(m, 0, *mlen);
        *mlen = 0;
        return -1;
    }
    memcpy(m, sm, *mlen);
    return 0;
}

/* ------------------------------------------------------------------ */
/* plain core: xmss_core.c                                             */
/* ------------------------------------------------------------------ */

/* treehash: xmss_core.c:19-86 */
void xo_treehash(const xo_params *p, unsigned char *root, unsigned char *auth_path,
                 const unsigned char *sk_seed, const unsigned char *pub_seed,
                 uint32_t leaf_idx, const uint32_t subtree_addr[8])
{
    unsigned char *stack = (unsigned char *)malloc((p->tree_height + 1) * p->n);
    unsigned int heights[64];
    unsigned int offset = 0;
    uint32_t idx, tree_idx;
    uint32_t ots_addr[8] = {0}, ltree_addr[8] = {0}, node_addr[8] = {0};

    copy_subtree(ots_addr, subtree_addr);
    copy_subtree(ltree_addr, subtree_addr);
    copy_subtree(node_addr, subtree_addr);
    ots_addr[3] = ADDR_OTS;
    ltree_addr[3] = ADDR_LTREE;
    node_addr[3] = ADDR_HASHTREE;

    for (idx = 0; idx < (uint32_t)(1u << p->tree_height); idx++) {
        ltree_addr[4] = idx;
        ots_addr[4] = idx;
        xo_gen_leaf_wots(p, stack + offset * p->n, sk_seed, pub_seed, ltree_addr, ots_addr);
        offset++;
        heights[offset - 1] = 0;
        if ((leaf_idx ^ 0x1) == idx) {
            memcpy(auth_path, stack + (offset - 1) * p->n, p->n);
        }
        while (offset >= 2 && heights[offset - 1] == heights[offset - 2]) {
            tree_idx = idx >> (heights[offset - 1] + 1);
            node_addr[5] = heights[offset - 1];
            node_addr[6] = tree_idx;
            xo_thash_h(p, stack + (offset - 2) * p->n, stack + (offset - 2) * p->n,
                       pub_seed, node_addr);
            offset--;
            heights[offset - 1]++;
            if (((leaf_idx >> heights[offset - 1]) ^ 0x1) == tree_idx) {
                memcpy(auth_path + heights[offset - 1] * p->n, stack + (offset - 1) * p->n, p->n);
            }
        }
    }
    memcpy(root, stack, p->n);
    free(stack);
}

/* The root of the 2^height leaves leaf0 .. leaf0 + 2^height - 1 (leaf0 a
 * multiple of 2^height) of the subtree at subtree_addr: treehash
 * (xmss_core.c:19-86) over that aligned range, with the same node
 * addresses the whole tree's treehash gives them.  Lets a CPU keygen split
 * its leaves across processes; with leaf0 = 0 and height = tree_height it is
 * xo_treehash's root. */
void xo_subtree_root(const xo_params *p, unsigned char *root, const unsigned char *sk_seed,
                     const unsigned char *pub_seed, uint32_t leaf0, unsigned height,
                     const uint32_t subtree_addr[8])
{
    unsigned char *stack = (unsigned char *)malloc((height + 1) * p->n);
    unsigned int heights[64];
    unsigned int offset = 0;
    uint32_t idx;
    uint32_t ots_addr[8] = {0}, ltree_addr[8] = {0}, node_addr[8] = {0};

    copy_subtree(ots_addr, subtree_addr);
    copy_subtree(ltree_addr, subtree_addr);
    copy_subtree(node_addr, subtree_addr);
    ots_addr[3] = ADDR_OTS;
    ltree_addr[3] = ADDR_LTREE;
    node_addr[3] = ADDR_HASHTREE;
    for (idx = leaf0; idx < leaf0 + (1u << height); idx++) {
        ltree_addr[4] = idx;
        ots_addr[4] = idx;
        xo_gen_leaf_wots(p, stack + offset * p->n, sk_seed, pub_seed, ltree_addr, ots_addr);
        offset++;
        heights[offset - 1] = 0;
        while (offset >= 2 && heights[offset - 1] == heights[offset - 2]) {
            node_addr[5] = heights[offset - 1];
            node_addr[6] = idx >> (heights[offset - 1] + 1);
            xo_thash_h(p, stack + (offset - 2) * p->n, stack + (offset - 2) * p->n, pub_seed, node_addr);
            offset--;
            heights[offset - 1]++;
        }
    }
    memcpy(root, stack, p->n);
    free(stack);
}

/* xmssmt_core_seed_keypair: xmss_core.c:134-161 */
int xo_plain_seed_keypair(const xo_params *p, unsigned char *pk, unsigned char *sk,
                          const unsigned char *seed)
{
    unsigned char *auth = (unsigned char *)malloc(p->tree_height * p->n);
    uint32_t top[8] = {0};
    top[0] = p->d - 1;
    memset(sk, 0, p->index_bytes);
    sk += p->index_bytes;
    memcpy(sk, seed, 2 * p->n);
    memcpy(sk + 3 * p->n, seed + 2 * p->n, p->n);
    memcpy(pk + p->n, sk + 3 * p->n, p->n);
    xo_treehash(p, pk, auth, sk, pk + p->n, 0, top);
    memcpy(sk + 2 * p->n, pk, p->n);
    free(auth);
    return 0;
}

/* xmssmt_core_keypair: xmss_core.c:168-177 (randombytes(3n) then seed keypair) */
int xo_plain_keypair(const xo_params *p, unsigned char *pk, unsigned char *sk)
{
    unsigned char seed[3 * 64];
    xo_randombytes(seed, 3 * p->n);
    return xo_plain_seed_keypair(p, pk, sk, seed);
}

/* Index check shared by both cores: xmss_core.c:209-231,
 * xmss_core_fast.c:742-751,970-979.  Returns -2 for an exhausted key,
 * 0 otherwise; wipes sk when idx >= 2^h - 1. */
static int check_index(const xo_params *p, unsigned char *sk, unsigned long long idx)
{
    if (idx >= ((1ULL << p->full_height) - 1)) {
        memset(sk, 0xFF, p->index_bytes);
        memset(sk + p->index_bytes, 0, (size_t)(p->sk_bytes - p->index_bytes));
        if (idx > ((1ULL << p->full_height) - 1)) {
            return -2;
        }
        if (p->full_height == 64 && idx == ((1ULL << p->full_height) - 1)) {
            return -2;
        }
    }
    return 0;
}

/* xmssmt_core_sign: xmss_core.c:183-273 (also xmss_core_sign, :116-126) */
int xo_plain_sign(const xo_params *p, unsigned char *sk, unsigned char *sm,
                  unsigned long long *smlen, const unsigned char *m,
                  unsigned long long mlen)
{
    const unsigned char *sk_seed = sk + p->index_bytes;
    const unsigned char *sk_prf = sk + p->index_bytes + p->n;
    const unsigned char *pub_root = sk + p->index_bytes + 2 * p->n;
    const unsigned char *pub_seed = sk + p->index_bytes + 3 * p->n;
    unsigned char root[64], idx32[32];
    unsigned long long idx;
    unsigned int i;
    uint32_t idx_leaf, ots_addr[8] = {0};

    memmove(sm + p->sig_bytes, m, mlen);
    *smlen = p->sig_bytes + mlen;
    idx = (unsigned long)from_bytes(sk, p->index_bytes);
    if (check_index(p, sk, idx)) {
        return -2;
    }
    memcpy(sm, sk, p->index_bytes);
    to_bytes(sk, p->index_bytes, idx + 1);
    to_bytes(idx32, 32, idx);
    xo_prf(p, sm + p->index_bytes, idx32, sk_prf);
    xo_hash_message(p, root, sm + p->index_bytes, pub_root, idx,
                    sm + p->sig_bytes - p->padding_len - 3 * p->n, mlen);
    sm += p->index_bytes + p->n;
    ots_addr[3] = ADDR_OTS;
    for (i = 0; i < p->d; i++) {
        idx_leaf = (uint32_t)(idx & ((1u << p->tree_height) - 1));
        idx >>= p->tree_height;
        ots_addr[0] = i;
        set_tree_addr(ots_addr, idx);
        ots_addr[4] = idx_leaf;
        xo_wots_sign(p, sm, root, sk_seed, pub_seed, ots_addr);
        sm += p->wots_sig_bytes;
        xo_treehash(p, root, sm, sk_seed, pub_seed, idx_leaf, ots_addr);
        sm += p->tree_height * p->n;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* BDS core: xmss_core_fast.c                                          */
/*                                                                     */
/* The state is a view into the sk bytes, exactly as the reference's   */
/* (de)serialize pair maps it (xmss_core_fast.c:38-136): node arrays    */
/* are pointers into sk, integer fields are copied out on load and      */
/* written back on store.  Fields the reference never assigns therefore */
/* keep whatever the caller's sk buffer held.                           */
/* ------------------------------------------------------------------ */

typedef struct {
    unsigned char h;
    unsigned long next_idx;
    unsigned char stackusage;
    unsigned char completed;
    unsigned char *node;
} th_inst;

typedef struct {
    unsigned char *stack;
    unsigned int stackoffset;
    unsigned char *stacklevels;
    unsigned char *auth;
    unsigned char *keep;
    th_inst *th;
    unsigned char *retain;
    unsigned int next_leaf;
} bds_t;

#define MAX_TH 64
#define MAX_STATES (2 * 12 - 1)

/* xmssmt_deserialize_state: xmss_core_fast.c:80-136 */
static void bds_load(const xo_params *p, bds_t *st, th_inst *ths,
                     unsigned char **wots_sigs, unsigned char *sk)
{
    unsigned int i, j, nth = p->tree_height - p->bds_k;
    sk += p->index_bytes + 4 * p->n;
    for (i = 0; i < 2 * p->d - 1; i++) {
        st[i].th = ths + i * nth;
        st[i].stack = sk;
        sk += (p->tree_height + 1) * p->n;
        st[i].stackoffset = (unsigned int)from_bytes(sk, 4);
        sk += 4;
        st[i].stacklevels = sk;
        sk += p->tree_height + 1;
        st[i].auth = sk;
        sk += p->tree_height * p->n;
        st[i].keep = sk;
        sk += (p->tree_height >> 1) * p->n;
        for (j = 0; j < nth; j++) {
            st[i].th[j].h = (unsigned char)from_bytes(sk, 1);
            st[i].th[j].next_idx = (unsigned long)from_bytes(sk + 1, 4);
            st[i].th[j].stackusage = (unsigned char)from_bytes(sk + 5, 1);
            st[i].th[j].completed = (unsigned char)from_bytes(sk + 6, 1);
            st[i].th[j].node = sk + 7;
            sk += 7 + p->n;
        }
        st[i].retain = sk;
        sk += ((1 << p->bds_k) - p->bds_k - 1) * p->n;
        st[i].next_leaf = (unsigned int)from_bytes(sk, 4);
        sk += 4;
    }
    if (wots_sigs) {
        *wots_sigs = sk;
    }
}

/* xmssmt_serialize_state: xmss_core_fast.c:38-78 */
static void bds_store(const xo_params *p, unsigned char *sk, const bds_t *st)
{
    unsigned int i, j, nth = p->tree_height - p->bds_k;
    sk += p->index_bytes + 4 * p->n;
    for (i = 0; i < 2 * p->d - 1; i++) {
        sk += (p->tree_height + 1) * p->n;
        to_bytes(sk, 4, st[i].stackoffset);
        sk += 4;
        sk += p->tree_height + 1;
        sk += p->tree_height * p->n;
        sk += (p->tree_height >> 1) * p->n;
        for (j = 0; j < nth; j++) {
            to_bytes(sk, 1, st[i].th[j].h);
            to_bytes(sk + 1, 4, st[i].th[j].next_idx);
            to_bytes(sk + 5, 1, st[i].th[j].stackusage);
            to_bytes(sk + 6, 1, st[i].th[j].completed);
            sk += 7 + p->n;
        }
        sk += ((1 << p->bds_k) - p->bds_k - 1) * p->n;
        to_bytes(sk, 4, st[i].next_leaf);
        sk += 4;
    }
}

/* deep_state_swap: xmss_core_fast.c:165-193 (swaps the bytes, not pointers) */
static void swap_bytes(unsigned char *a, unsigned char *b, size_t len)
{
    size_t i;
    for (i = 0; i < len; i++) {
        unsigned char t = a[i];
        a[i] = b[i];
        b[i] = t;
    }
}

static void bds_swap(const xo_params *p, bds_t *a, bds_t *b)
{
    unsigned int i, t;
    unsigned long tl;
    unsigned char tc;
    swap_bytes(a->stack, b->stack, (p->tree_height + 1) * p->n);
    t = a->stackoffset; a->stackoffset = b->stackoffset; b->stackoffset = t;
    swap_bytes(a->stacklevels, b->stacklevels, p->tree_height + 1);
    swap_bytes(a->auth, b->auth, p->tree_height * p->n);
    swap_bytes(a->keep, b->keep, (p->tree_height >> 1) * p->n);
    for (i = 0; i < p->tree_height - p->bds_k; i++) {
        tc = a->th[i].h; a->th[i].h = b->th[i].h; b->th[i].h = tc;
        tl = a->th[i].next_idx; a->th[i].next_idx = b->th[i].next_idx; b->th[i].next_idx = tl;
        tc = a->th[i].stackusage; a->th[i].stackusage = b->th[i].stackusage; b->th[i].stackusage = tc;
        tc = a->th[i].completed; a->th[i].completed = b->th[i].completed; b->th[i].completed = tc;
        swap_bytes(a->th[i].node, b->th[i].node, p->n);
    }
    swap_bytes(a->retain, b->retain, ((1 << p->bds_k) - p->bds_k - 1) * p->n);
    t = a->next_leaf; a->next_leaf = b->next_leaf; b->next_leaf = t;
}

/* treehash_minheight_on_stack: xmss_core_fast.c:195-207 */
static unsigned int minheight_on_stack(const xo_params *p, const bds_t *st, const th_inst *t)
{
    unsigned int r = p->tree_height, i;
    for (i = 0; i < t->stackusage; i++) {
        if (st->stacklevels[st->stackoffset - i - 1] < r) {
            r = st->stacklevels[st->stackoffset - i - 1];
        }
    }
    return r;
}

static void bds_addrs(const uint32_t addr[8], uint32_t ots[8], uint32_t lt[8], uint32_t nd[8])
{
    memset(ots, 0, 32);
    memset(lt, 0, 32);
    memset(nd, 0, 32);
    copy_subtree(ots, addr);
    ots[3] = ADDR_OTS;
    copy_subtree(lt, addr);
    lt[3] = ADDR_LTREE;
    copy_subtree(nd, addr);
    nd[3] = ADDR_HASHTREE;
}

/* Stack-merge bookkeeping shared by treehash_init (xmss_core_fast.c:260-276)
 * and bds_state_update (:532-549): before hashing the two top nodes of
 * height nodeh, record the right one as an auth / treehash / retain node. */
static void bds_record(const xo_params *p, bds_t *st, unsigned int nodeh,
                       unsigned int i, const unsigned char *node)
{
    if (i >> nodeh == 1) {
        memcpy(st->auth + nodeh * p->n, node, p->n);
    } else if (nodeh < p->tree_height - p->bds_k && i >> nodeh == 3) {
        memcpy(st->th[nodeh].node, node, p->n);
    } else if (nodeh >= p->tree_height - p->bds_k) {
        memcpy(st->retain + ((1 << (p->tree_height - 1 - nodeh)) + nodeh - p->tree_height
                             + (((i >> nodeh) - 3) >> 1)) * p->n,
               node, p->n);
    }
}

/* treehash_init: xmss_core_fast.c:216-285 (the BDS keygen of one tree).
 * The reference's copy of the empty slot above the stack into treehash[0]
 * at leaf 3 (:257-259) is overwritten by the merge at the same leaf, so it
 * has no observable effect and is not restated. */
static void bds_treehash_init(const xo_params *p, unsigned char *node, bds_t *st,
                              const unsigned char *sk_seed, const unsigned char *pub_seed,
                              const uint32_t addr[8])
{
    uint32_t ots[8], lt[8], nd[8];
    unsigned int height = p->tree_height, idx, lastnode, i, nodeh;
    unsigned char *stack = (unsigned char *)malloc((height + 1) * p->n);
    unsigned int levels[MAX_TH + 1];
    unsigned int off = 0;

    bds_addrs(addr, ots, lt, nd);
    lastnode = 1u << height;
    for (i = 0; i < p->tree_height - p->bds_k; i++) {
        st->th[i].h = (unsigned char)i;
        st->th[i].completed = 1;
        st->th[i].stackusage = 0;
    }
    for (idx = 0, i = 0; idx < lastnode; idx++, i++) {
        lt[4] = idx;
        ots[4] = idx;
        xo_gen_leaf_wots(p, stack + off * p->n, sk_seed, pub_seed, lt, ots);
        levels[off] = 0;
        off++;
        while (off > 1 && levels[off - 1] == levels[off - 2]) {
            nodeh = levels[off - 1];
            bds_record(p, st, nodeh, i, stack + (off - 1) * p->n);
            nd[5] = levels[off - 1];
            nd[6] = idx >> (levels[off - 1] + 1);
            xo_thash_h(p, stack + (off - 2) * p->n, stack + (off - 2) * p->n, pub_seed, nd);
            levels[off - 2]++;
            off--;
        }
    }
    memcpy(node, stack, p->n);
    free(stack);
}

/* treehash_update: xmss_core_fast.c:406-452 */
static void bds_th_update(const xo_params *p, th_inst *t, bds_t *st,
                          const unsigned char *sk_seed, const unsigned char *pub_seed,
                          const uint32_t addr[8])
{
    uint32_t ots[8], lt[8], nd[8];
    unsigned char nb[2 * 64];
    unsigned int nodeheight = 0;
    bds_addrs(addr, ots, lt, nd);
    lt[4] = (uint32_t)t->next_idx;
    ots[4] = (uint32_t)t->next_idx;
    xo_gen_leaf_wots(p, nb, sk_seed, pub_seed, lt, ots);
    while (t->stackusage > 0 && st->stacklevels[st->stackoffset - 1] == nodeheight) {
        memcpy(nb + p->n, nb, p->n);
        memcpy(nb, st->stack + (st->stackoffset - 1) * p->n, p->n);
        nd[5] = nodeheight;
        nd[6] = (uint32_t)(t->next_idx >> (nodeheight + 1));
        xo_thash_h(p, nb, nb, pub_seed, nd);
        nodeheight++;
        t->stackusage--;
        st->stackoffset--;
    }
    if (nodeheight == t->h) {
        memcpy(t->node, nb, p->n);
        t->completed = 1;
    } else {
        memcpy(st->stack + st->stackoffset * p->n, nb, p->n);
        t->stackusage++;
        st->stacklevels[st->stackoffset] = (unsigned char)nodeheight;
        st->stackoffset++;
        t->next_idx++;
    }
}

/* bds_treehash_update: xmss_core_fast.c:458-491.  The reference returns
 * `char`; the value is narrowed and sign-extended exactly the same way here. */
static unsigned int bds_th_updates(const xo_params *p, bds_t *st, unsigned int updates,
                                   const unsigned char *sk_seed,
                                   const unsigned char *pub_seed, const uint32_t addr[8])
{
    unsigned int i, j, level, l_min, low, used = 0;
    unsigned int nth = p->tree_height - p->bds_k;
    for (j = 0; j < updates; j++) {
        l_min = p->tree_height;
        level = nth;
        for (i = 0; i < nth; i++) {
            if (st->th[i].completed) {
                low = p->tree_height;
            } else if (st->th[i].stackusage == 0) {
                low = i;
            } else {
                low = minheight_on_stack(p, st, &st->th[i]);
            }
            if (low < l_min) {
                level = i;
                l_min = low;
            }
        }
        if (level == nth) {
            break;
        }
        bds_th_update(p, &st->th[level], st, sk_seed, pub_seed, addr);
        used++;
    }
    return (unsigned int)(int)(signed char)(updates - used);
}

/* bds_state_update: xmss_core_fast.c:497-552 */
static int bds_state_update(const xo_params *p, bds_t *st,
                            const unsigned char *sk_seed, const unsigned char *pub_seed,
                            const uint32_t addr[8])
{
    uint32_t ots[8], lt[8], nd[8];
    unsigned int nodeh;
    int idx = (int)st->next_leaf;
    if (idx == 1 << p->tree_height) {
        return -1;
    }
    bds_addrs(addr, ots, lt, nd);
    ots[4] = (uint32_t)idx;
    lt[4] = (uint32_t)idx;
    xo_gen_leaf_wots(p, st->stack + st->stackoffset * p->n, sk_seed, pub_seed, lt, ots);
    st->stacklevels[st->stackoffset] = 0;
    st->stackoffset++;
    while (st->stackoffset > 1
           && st->stacklevels[st->stackoffset - 1] == st->stacklevels[st->stackoffset - 2]) {
        nodeh = st->stacklevels[st->stackoffset - 1];
        bds_record(p, st, nodeh, (unsigned int)idx, st->stack + (st->stackoffset - 1) * p->n);
        nd[5] = st->stacklevels[st->stackoffset - 1];
        nd[6] = (uint32_t)(idx >> (st->stacklevels[st->stackoffset - 1] + 1));
        xo_thash_h(p, st->stack + (st->stackoffset - 2) * p->n,
                   st->stack + (st->stackoffset - 2) * p->n, pub_seed, nd);
        st->stacklevels[st->stackoffset - 2]++;
        st->stackoffset--;
    }
    st->next_leaf++;
    return 0;
}

/* bds_round: xmss_core_fast.c:559-626 */
static void bds_round(const xo_params *p, bds_t *st, unsigned long leaf_idx,
                      const unsigned char *sk_seed, const unsigned char *pub_seed,
                      const uint32_t addr[8])
{
    unsigned int i, tau = p->tree_height, startidx, offset, rowidx, lim;
    unsigned char buf[2 * 64];
    uint32_t ots[8], lt[8], nd[8];
    unsigned int nth = p->tree_height - p->bds_k;
    bds_addrs(addr, ots, lt, nd);
    for (i = 0; i < p->tree_height; i++) {
        if (!((leaf_idx >> i) & 1)) {
            tau = i;
            break;
        }
    }
    if (tau > 0) {
        memcpy(buf, st->auth + (tau - 1) * p->n, p->n);
        memcpy(buf + p->n, st->keep + ((tau - 1) >> 1) * p->n, p->n);
    }
    if (!((leaf_idx >> (tau + 1)) & 1) && (tau < p->tree_height - 1)) {
        memcpy(st->keep + (tau >> 1) * p->n, st->auth + tau * p->n, p->n);
    }
    if (tau == 0) {
        lt[4] = (uint32_t)leaf_idx;
        ots[4] = (uint32_t)leaf_idx;
        xo_gen_leaf_wots(p, st->auth, sk_seed, pub_seed, lt, ots);
    } else {
        nd[5] = tau - 1;
        nd[6] = (uint32_t)(leaf_idx >> tau);
        xo_thash_h(p, st->auth + tau * p->n, buf, pub_seed, nd);
        for (i = 0; i < tau; i++) {
            if (i < nth) {
                memcpy(st->auth + i * p->n, st->th[i].node, p->n);
            } else {
                offset = (1 << (p->tree_height - 1 - i)) + i - p->tree_height;
                rowidx = (unsigned int)(((leaf_idx >> i) - 1) >> 1);
                memcpy(st->auth + i * p->n, st->retain + (offset + rowidx) * p->n, p->n);
            }
        }
        lim = tau < nth ? tau : nth;
        for (i = 0; i < lim; i++) {
            startidx = (unsigned int)(leaf_idx + 1 + 3 * (1u << i));
            if (startidx < 1u << p->tree_height) {
                st->th[i].h = (unsigned char)i;
                st->th[i].next_idx = startidx;
                st->th[i].completed = 0;
                st->th[i].stackusage = 0;
            }
        }
    }
}

/* xmss_core_keypair (BDS): xmss_core_fast.c:656-701.  The reference
 * ignores randombytes here and uses SK_SEED = SK_PRF = 0x02.., PUB_SEED =
 * 0x04.. (:678-687). */
int xo_bds_xmss_keypair(const xo_params *p, unsigned char *pk, unsigned char *sk)
{
    uint32_t addr[8] = {0};
    bds_t st;
    th_inst ths[MAX_TH];
    unsigned int i;
    bds_load(p, &st, ths, NULL, sk);
    st.stackoffset = 0;
    st.next_leaf = 0;
    sk[0] = sk[1] = sk[2] = sk[3] = 0;
    for (i = 0; i < 2 * p->n; i++) {
        sk[p->index_bytes + i] = 2;
    }
    for (i = 0; i < p->n; i++) {
        sk[p->index_bytes + 3 * p->n + i] = 4;
    }
    memcpy(pk + p->n, sk + p->index_bytes + 3 * p->n, p->n);
    bds_treehash_init(p, pk, &st, sk + p->index_bytes, sk + p->index_bytes + 3 * p->n, addr);
    memcpy(sk + p->index_bytes + 2 * p->n, pk, p->n);
    bds_store(p, sk, &st);
    return 0;
}

/* xmss_core_sign (BDS): xmss_core_fast.c:710-851 */
int xo_bds_xmss_sign(const xo_params *p, unsigned char *sk, unsigned char *sm,
                     unsigned long long *smlen, const unsigned char *m,
                     unsigned long long mlen)
{
    const unsigned char *pub_root = sk + p->index_bytes + 2 * p->n;
    bds_t st;
    th_inst ths[MAX_TH];
    unsigned char sk_seed[64], sk_prf[64], pub_seed[64], idx32[32], R[64], mh[64];
    uint32_t ots_addr[8] = {0};
    unsigned long idx;

    bds_load(p, &st, ths, NULL, sk);
    idx = ((unsigned long)sk[0] << 24) | ((unsigned long)sk[1] << 16) |
          ((unsigned long)sk[2] << 8) | sk[3];
    if (check_index(p, sk, idx)) {
        return -2;
    }
    memcpy(sk_seed, sk + p->index_bytes, p->n);
    memcpy(sk_prf, sk + p->index_bytes + p->n, p->n);
    memcpy(pub_seed, sk + p->index_bytes + 3 * p->n, p->n);
    to_bytes(idx32, 32, idx);
    to_bytes(sk, 4, idx + 1);
    xo_prf(p, R, idx32, sk_prf);
    memmove(sm + p->sig_bytes, m, mlen);
    xo_hash_message(p, mh, R, pub_root, idx, sm + p->sig_bytes - p->padding_len - 3 * p->n, mlen);
    to_bytes(sm, 4, idx);
    memcpy(sm + 4, R, p->n);
    *smlen = 4 + p->n;
    sm += 4 + p->n;
    ots_addr[3] = ADDR_OTS;
    ots_addr[4] = (uint32_t)idx;
    xo_wots_sign(p, sm, mh, sk_seed, pub_seed, ots_addr);
    sm += p->wots_sig_bytes;
    *smlen += p->wots_sig_bytes;
    memcpy(sm, st.auth, p->tree_height * p->n);
    if (idx < (1U << p->tree_height) - 1) {
        bds_round(p, &st, idx, sk_seed, pub_seed, ots_addr);
        bds_th_updates(p, &st, (p->tree_height - p->bds_k) >> 1, sk_seed, pub_seed, ots_addr);
    }
    sm += p->tree_height * p->n;
    *smlen += p->tree_height * p->n;
    memmove(sm, m, mlen);
    *smlen += mlen;
    bds_store(p, sk, &st);
    return 0;
}

/* xmssmt_core_keypair (BDS): xmss_core_fast.c:858-908 */
int xo_bds_xmssmt_keypair(const xo_params *p, unsigned char *pk, unsigned char *sk)
{
    uint32_t addr[8] = {0};
    unsigned int i;
    unsigned char *wots_sigs;
    bds_t st[MAX_STATES];
    th_inst ths[MAX_STATES * MAX_TH];

    bds_load(p, st, ths, &wots_sigs, sk);
    for (i = 0; i < 2 * p->d - 1; i++) {
        st[i].stackoffset = 0;
        st[i].next_leaf = 0;
    }
    for (i = 0; i < p->index_bytes; i++) {
        sk[i] = 0;
    }
    xo_randombytes(sk + p->index_bytes, 2 * p->n);
    xo_randombytes(sk + p->index_bytes + 3 * p->n, p->n);
    memcpy(pk + p->n, sk + p->index_bytes + 3 * p->n, p->n);
    addr[0] = 0;
    for (i = 0; i < p->d - 1; i++) {
        bds_treehash_init(p, pk, st + i, sk + p->index_bytes, pk + p->n, addr);
        addr[0] = i + 1;
        xo_wots_sign(p, wots_sigs + i * p->wots_sig_bytes, pk, sk + p->index_bytes, pk + p->n, addr);
    }
    bds_treehash_init(p, pk, st + i, sk + p->index_bytes, pk + p->n, addr);
    memcpy(sk + p->index_bytes + 2 * p->n, pk, p->n);
    bds_store(p, sk, st);
    return 0;
}

/* xmssmt_core_sign (BDS): xmss_core_fast.c:917-1121 */
int xo_bds_xmssmt_sign(const xo_params *p, unsigned char *sk, unsigned char *sm,
                       unsigned long long *smlen, const unsigned char *m,
                       unsigned long long mlen)
{
    const unsigned char *pub_root = sk + p->index_bytes + 2 * p->n;
    uint64_t idx_tree, i, j;
    uint32_t idx_leaf;
    int needswap_upto = -1;
    unsigned int updates;
    unsigned char sk_seed[64], sk_prf[64], pub_seed[64], R[64], mh[64], idx32[32];
    uint32_t addr[8] = {0}, ots_addr[8] = {0};
    unsigned char *wots_sigs;
    bds_t st[MAX_STATES];
    th_inst ths[MAX_STATES * MAX_TH];
    unsigned long long idx = 0;
    unsigned int th = p->tree_height;

    bds_load(p, st, ths, &wots_sigs, sk);
    for (i = 0; i < p->index_bytes; i++) {
        idx |= ((unsigned long long)sk[i]) << 8 * (p->index_bytes - 1 - i);
    }
    if (check_index(p, sk, idx)) {
        return -2;
    }
    memcpy(sk_seed, sk + p->index_bytes, p->n);
    memcpy(sk_prf, sk + p->index_bytes + p->n, p->n);
    memcpy(pub_seed, sk + p->index_bytes + 3 * p->n, p->n);
    for (i = 0; i < p->index_bytes; i++) {
        sk[i] = ((idx + 1) >> 8 * (p->index_bytes - 1 - i)) & 255;
    }
    to_bytes(idx32, 32, idx);
    xo_prf(p, R, idx32, sk_prf);
    memmove(sm + p->sig_bytes, m, mlen);
    xo_hash_message(p, mh, R, pub_root, idx, sm + p->sig_bytes - p->padding_len - 3 * p->n, mlen);
    for (i = 0; i < p->index_bytes; i++) {
        sm[i] = (idx >> 8 * (p->index_bytes - 1 - i)) & 255;
    }
    memcpy(sm + p->index_bytes, R, p->n);
    sm += p->index_bytes + p->n;
    *smlen = p->index_bytes + p->n;

    ots_addr[3] = ADDR_OTS;
    idx_tree = idx >> th;
    idx_leaf = (uint32_t)(idx & ((1u << th) - 1));
    ots_addr[0] = 0;
    set_tree_addr(ots_addr, idx_tree);
    ots_addr[4] = idx_leaf;
    xo_wots_sign(p, sm, mh, sk_seed, pub_seed, ots_addr);
    sm += p->wots_sig_bytes;
    *smlen += p->wots_sig_bytes;
    memcpy(sm, st[0].auth, th * p->n);
    sm += th * p->n;
    *smlen += th * p->n;
    for (i = 1; i < p->d; i++) {
        memcpy(sm, wots_sigs + (i - 1) * p->wots_sig_bytes, p->wots_sig_bytes);
        sm += p->wots_sig_bytes;
        *smlen += p->wots_sig_bytes;
        memcpy(sm, st[i].auth, th * p->n);
        sm += th * p->n;
        *smlen += th * p->n;
    }

    updates = (th - p->bds_k) >> 1;
    set_tree_addr(addr, idx_tree + 1);
    if ((1 + idx_tree) * (1ULL << th) + idx_leaf < (1ULL << p->full_height)) {
        bds_state_update(p, &st[p->d], sk_seed, pub_seed, addr);
    }
    for (i = 0; i < p->d; i++) {
        if (!(((idx + 1) & ((1ULL << ((i + 1) * th)) - 1)) == 0)) {
            idx_leaf = (uint32_t)((idx >> (th * i)) & ((1u << th) - 1));
            idx_tree = idx >> (th * (i + 1));
            addr[0] = (uint32_t)i;
            set_tree_addr(addr, idx_tree);
            if (i == (uint64_t)(needswap_upto + 1)) {
                bds_round(p, &st[i], idx_leaf, sk_seed, pub_seed, addr);
            }
            updates = bds_th_updates(p, &st[i], updates, sk_seed, pub_seed, addr);
            set_tree_addr(addr, idx_tree + 1);
            if ((1 + idx_tree) * (1ULL << th) + idx_leaf < (1ULL << (p->full_height - th * i))) {
                if (i > 0 && updates > 0 && st[p->d + i].next_leaf < (1ULL << p->full_height)) {
                    bds_state_update(p, &st[p->d + i], sk_seed, pub_seed, addr);
                    updates--;
                }
            }
        } else if (idx < (1ULL << p->full_height) - 1) {
            bds_swap(p, st + p->d + i, st + i);
            ots_addr[0] = (uint32_t)(i + 1);
            set_tree_addr(ots_addr, (idx + 1) >> ((i + 2) * th));
            ots_addr[4] = (uint32_t)(((idx >> ((i + 1) * th)) + 1) & ((1u << th) - 1));
            xo_wots_sign(p, wots_sigs + i * p->wots_sig_bytes, st[i].stack, sk_seed, pub_seed, ots_addr);
            st[p->d + i].stackoffset = 0;
            st[p->d + i].next_leaf = 0;
            updates--;
            needswap_upto = (int)i;
            for (j = 0; j < th - p->bds_k; j++) {
                st[i].th[j].completed = 1;
            }
        }
    }
    memmove(sm, m, mlen);
    *smlen += mlen;
    bds_store(p, sk, st);
    return 0;
}

/* ------------------------------------------------------------------ */
/* OID-prefixed API: xmss.c:10-109                                     */
/* ------------------------------------------------------------------ */

static uint32_t read_oid(const unsigned char *b)
{
    return ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | b[3];
}

int xo_api_keypair(unsigned char *pk, unsigned char *sk, uint32_t oid, int mt, int core)
{
    xo_params p;
    unsigned int i;
    if (mt ? xo_xmssmt_parse_oid(&p, oid, core) : xo_xmss_parse_oid(&p, oid, core)) {
        return -1;
    }
    for (i = 0; i < 4; i++) {
        pk[3 - i] = (oid >> (8 * i)) & 0xFF;
        sk[3 - i] = (oid >> (8 * i)) & 0xFF;
    }
    if (core) {
        return (mt ? xo_bds_xmssmt_keypair : xo_bds_xmss_keypair)(&p, pk + 4, sk + 4);
    }
    return xo_plain_keypair(&p, pk + 4, sk + 4);
}

int xo_api_sign(unsigned char *sk, unsigned char *sm, unsigned long long *smlen,
                const unsigned char *m, unsigned long long mlen, int mt, int core)
{
    xo_params p;
    uint32_t oid = read_oid(sk);
    if (mt ? xo_xmssmt_parse_oid(&p, oid, core) : xo_xmss_parse_oid(&p, oid, core)) {
        return -1;
    }
    if (core) {
        return (mt ? xo_bds_xmssmt_sign : xo_bds_xmss_sign)(&p, sk + 4, sm, smlen, m, mlen);
    }
    return xo_plain_sign(&p, sk + 4, sm, smlen, m, mlen);
}

int xo_api_sign_open(unsigned char *m, unsigned long long *mlen,
                     const unsigned char *sm, unsigned long long smlen,
                     const unsigned char *pk, int mt)
{
    xo_params p;
    uint32_t oid = read_oid(pk);
    if (mt ? xo_xmssmt_parse_oid(&p, oid, 0) : xo_xmss_parse_oid(&p, oid, 0)) {
        return -1;
    }
    return xo_sign_open(&p, m, mlen, sm, smlen, pk + 4);
}
