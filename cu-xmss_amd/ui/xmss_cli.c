/*
 * xmss_cli.c -- command-line tools with the interface of the reference's
 * cuxmss/ui/{keypair,sign,open}.c, built on the MI355X library:
 *
 *   <p>_keypair <parameter set>        writes pk || sk (both OID-prefixed) to stdout
 *   <p>_sign <keypair file> <message>  signs with the sk in the keypair file,
 *                                      writes the advanced sk back into the
 *                                      file, then sm (signature || message)
 *                                      to stdout
 *   <p>_open <keypair file> <sm file>  exit status 0 when sm verifies under
 *                                      the file's pk; prints the verdict
 *
 * <p> is xmss, or xmssmt when compiled with -DCLI_MT; CLI_TOOL names the tool.
 * Linked against libxmss_mi355x.so (plain core) or libxmss_mi355x_fast.so
 * (BDS core), as the reference links xmss_core.c or xmss_core_fast.c.
 *
 * Beyond the reference: every read and write is checked; sign persists the
 * advanced secret key (fflush + fsync of the keypair file) before any byte
 * of the signature is emitted, and emits nothing when signing fails;
 * XMSS_MI355X_SEED in the environment (decimal or 0x-hex) seeds the
 * library's randombytes stream so that keys can be reproduced.
 */
#define _POSIX_C_SOURCE 200809L
#include <errno.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "../../include/params.h"
#include "../../include/xmss.h"
#include "../../include/xmss_mi355x.h"

#ifndef CLI_TOOL
#error "CLI_TOOL must name the tool: keypair, sign or open"
#endif

#ifdef CLI_MT
#define PREFIX "xmssmt"
#define STR_TO_OID xmssmt_str_to_oid
#define PARSE_OID xmssmt_parse_oid
#define KEYPAIR xmssmt_keypair
#define SIGN xmssmt_sign
#define SIGN_OPEN xmssmt_sign_open
#else
#define PREFIX "xmss"
#define STR_TO_OID xmss_str_to_oid
#define PARSE_OID xmss_parse_oid
#define KEYPAIR xmss_keypair
#define SIGN xmss_sign
#define SIGN_OPEN xmss_sign_open
#endif

static uint32_t be32(const unsigned char *b) {
    return ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | b[3];
}

/* Whole file into a fresh buffer (one spare byte so that empty files work). */
static unsigned char *slurp(const char *path, size_t *len) {
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    size_t cap = 1 << 16, n = 0;
    unsigned char *buf = malloc(cap);
    for (;;) {
        if (!buf) break;
        n += fread(buf + n, 1, cap - n, f);
        if (n < cap) break;
        unsigned char *nb = realloc(buf, cap * 2);
        if (!nb) {
            free(buf);
            buf = NULL;
            break;
        }
        buf = nb;
        cap *= 2;
    }
    if (buf && ferror(f)) {
        free(buf);
        buf = NULL;
    }
    fclose(f);
    *len = n;
    return buf;
}

static int write_all(FILE *f, const unsigned char *p, size_t n) {
    return fwrite(p, 1, n, f) == n ? 0 : -1;
}

static void seed_from_env(void) {
    const char *s = getenv("XMSS_MI355X_SEED");
    if (s && *s) xmss_mi355x_rng_seed((uint64_t)strtoull(s, NULL, 0));
}

/* Parameters of the OID at b (4 bytes, big-endian); -1 if unknown. */
static int params_at(xmss_params *p, const unsigned char *b, size_t avail) {
    if (avail < XMSS_OID_LEN) return -1;
    return PARSE_OID(p, be32(b));
}

static int tool_keypair(int argc, char **argv) {
    if (argc != 2) {
        fprintf(stderr, "usage: %s_keypair <parameter set, e.g. %s>\n"
                        "The keypair (pk || sk) is written to stdout.\n",
                PREFIX, PREFIX[4] ? "XMSSMT-SHA2_20/2_256" : "XMSS-SHA2_10_256");
        return -1;
    }
    uint32_t oid = 0;
    xmss_params p;
    if (STR_TO_OID(&oid, argv[1]) != 0 || PARSE_OID(&p, oid) != 0) {
        fprintf(stderr, "unknown parameter set: %s\n", argv[1]);
        return -1;
    }
    seed_from_env();
    const size_t pkl = XMSS_OID_LEN + p.pk_bytes, skl = XMSS_OID_LEN + (size_t)p.sk_bytes;
    unsigned char *pk = calloc(1, pkl), *sk = calloc(1, skl);
    int rc = (pk && sk) ? KEYPAIR(pk, sk, oid) : -1;
    if (rc == 0 && (write_all(stdout, pk, pkl) || write_all(stdout, sk, skl) || fflush(stdout))) rc = -1;
    if (rc != 0) fprintf(stderr, "keypair failed (%d)\n", rc);
    free(pk);
    free(sk);
    return rc;
}

static int tool_sign(int argc, char **argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: %s_sign <keypair file> <message file>\n"
                        "The secret key in the keypair file is advanced in place; the signed\n"
                        "message (signature || message) is written to stdout.\n",
                PREFIX);
        return -1;
    }
    size_t kl = 0, ml = 0;
    unsigned char *kp = slurp(argv[1], &kl);
    if (!kp) {
        fprintf(stderr, "cannot read keypair file %s\n", argv[1]);
        return -1;
    }
    unsigned char *m = slurp(argv[2], &ml);
    if (!m) {
        fprintf(stderr, "cannot read message file %s\n", argv[2]);
        free(kp);
        return -1;
    }
    xmss_params p;
    int rc = -1;
    unsigned char *sm = NULL;
    FILE *kf = NULL;
    if (params_at(&p, kp, kl) != 0) {
        fprintf(stderr, "unknown public key OID\n");
        goto out;
    }
    const size_t sk_at = XMSS_OID_LEN + p.pk_bytes;
    /* the secret key's own OID decides its size, as in the reference */
    if (params_at(&p, kp + sk_at, kl > sk_at ? kl - sk_at : 0) != 0) {
        fprintf(stderr, "unknown secret key OID\n");
        goto out;
    }
    if (kl < sk_at + XMSS_OID_LEN + p.sk_bytes) {
        fprintf(stderr, "keypair file too short\n");
        goto out;
    }
    unsigned char *sk = kp + sk_at;
    unsigned long long smlen = 0;
    sm = malloc(p.sig_bytes + ml + 1);
    if (!sm) goto out;
    rc = SIGN(sk, sm, &smlen, m, ml);
    /* persist the (possibly advanced or wiped) key before anything else */
    kf = fopen(argv[1], "r+b");
    if (!kf || fseek(kf, (long)(sk_at + XMSS_OID_LEN), SEEK_SET) != 0 ||
        write_all(kf, sk + XMSS_OID_LEN, p.sk_bytes) || fflush(kf) || fsync(fileno(kf))) {
        fprintf(stderr, "cannot write the secret key back: %s\n", strerror(errno));
        rc = -1;
        goto out;
    }
    if (rc != 0) {
        fprintf(stderr, "signing failed (%d)\n", rc);
        goto out;
    }
    if (write_all(stdout, sm, smlen) || fflush(stdout)) rc = -1;
out:
    if (kf) fclose(kf);
    free(sm);
    free(m);
    free(kp);
    return rc;
}

static int tool_open(int argc, char **argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: %s_open <keypair file> <signed message file>\n"
                        "The keypair file needs only the public key; exit status 0 means the\n"
                        "signature verifies.\n",
                PREFIX);
        return -1;
    }
    size_t kl = 0, sl = 0;
    unsigned char *kp = slurp(argv[1], &kl);
    if (!kp) {
        fprintf(stderr, "cannot read keypair file %s\n", argv[1]);
        return -1;
    }
    unsigned char *sm = slurp(argv[2], &sl);
    if (!sm) {
        fprintf(stderr, "cannot read signed message file %s\n", argv[2]);
        free(kp);
        return -1;
    }
    xmss_params p;
    int rc = -1;
    if (params_at(&p, kp, kl) != 0 || kl < XMSS_OID_LEN + p.pk_bytes) {
        fprintf(stderr, "unknown or truncated public key\n");
    } else {
        unsigned char *m = malloc(sl + 1);
        unsigned long long mlen = 0;
        rc = m ? SIGN_OPEN(m, &mlen, sm, sl, kp) : -1;
        printf(rc ? "Verification failed!\n" : "Verification succeeded.\n");
        free(m);
    }
    free(sm);
    free(kp);
    return rc;
}

int main(int argc, char **argv) {
    if (!strcmp(CLI_TOOL, "keypair")) return tool_keypair(argc, argv);
    if (!strcmp(CLI_TOOL, "sign")) return tool_sign(argc, argv);
    return tool_open(argc, argv);
}
