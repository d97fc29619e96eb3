/*
 * xo_hash.c -- oracle hash primitives.  TEST INFRASTRUCTURE ONLY (see xo.h).
 *
 * SHA-256 / SHA-512 follow FIPS 180-4 section 6; Keccak-f1600 and the SHAKE
 * sponge follow FIPS 202 (the reference's vendored copies are
 * cuxmss/sha512.c:241-336,442-453 and cuxmss/fips202.c:65-417; n=24/32 SHA-256
 * is OpenSSL's SHA256() in the reference, cuxmss/hash.c:37,45).
 * The Keccak rotation offsets and round constants are derived at first use from
 * the FIPS 202 definitions (rho walk, 3.2.2; rc() LFSR, 3.2.5) instead of
 * being typed in.
 *
 * Pinned by: FIPS 180-4 / FIPS 202 known answers and Python hashlib over
 * random lengths (tests/test_oracle.py).
 */
#include <stdint.h>
#include <string.h>

#include "xo.h"

/* ------------------------------------------------------------------ */
/* SHA-256 (FIPS 180-4 6.2)                                            */
/* ------------------------------------------------------------------ */

static const uint32_t K256[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

static uint32_t ror32(uint32_t x, int r) { return (x >> r) | (x << (32 - r)); }

static void sha256_block(uint32_t h[8], const unsigned char *blk)
{
    uint32_t w[64], a, b, c, d, e, f, g, hh, t1, t2;
    int t;
    for (t = 0; t < 16; t++) {
        w[t] = ((uint32_t)blk[4 * t] << 24) | ((uint32_t)blk[4 * t + 1] << 16) |
               ((uint32_t)blk[4 * t + 2] << 8) | (uint32_t)blk[4 * t + 3];
    }
    for (t = 16; t < 64; t++) {
        uint32_t s0 = ror32(w[t - 15], 7) ^ ror32(w[t - 15], 18) ^ (w[t - 15] >> 3);
        uint32_t s1 = ror32(w[t - 2], 17) ^ ror32(w[t - 2], 19) ^ (w[t - 2] >> 10);
        w[t] = w[t - 16] + s0 + w[t - 7] + s1;
    }
    a = h[0]; b = h[1]; c = h[2]; d = h[3]; e = h[4]; f = h[5]; g = h[6]; hh = h[7];
    for (t = 0; t < 64; t++) {
        t1 = hh + (ror32(e, 6) ^ ror32(e, 11) ^ ror32(e, 25)) + ((e & f) ^ (~e & g)) + K256[t] + w[t];
        t2 = (ror32(a, 2) ^ ror32(a, 13) ^ ror32(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
        hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
}

void xo_sha256(unsigned char out[32], const unsigned char *in, unsigned long long inlen)
{
    uint32_t h[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                     0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    unsigned char tail[128];
    unsigned long long bits = inlen * 8, rem, i, tl;
    while (inlen >= 64) {
        sha256_block(h, in);
        in += 64;
        inlen -= 64;
    }
    rem = inlen;
    memset(tail, 0, sizeof tail);
    memcpy(tail, in, rem);
    tail[rem] = 0x80;
    tl = (rem + 9 <= 64) ? 64 : 128;
    for (i = 0; i < 8; i++) {
        tail[tl - 1 - i] = (unsigned char)(bits >> (8 * i));
    }
    sha256_block(h, tail);
    if (tl == 128) {
        sha256_block(h, tail + 64);
    }
    for (i = 0; i < 8; i++) {
        out[4 * i] = (unsigned char)(h[i] >> 24);
        out[4 * i + 1] = (unsigned char)(h[i] >> 16);
        out[4 * i + 2] = (unsigned char)(h[i] >> 8);
        out[4 * i + 3] = (unsigned char)h[i];
    }
}

/* ------------------------------------------------------------------ */
/* SHA-512 (FIPS 180-4 6.4)                                            */
/* ------------------------------------------------------------------ */

static const uint64_t K512[80] = {
    0x428a2f98d728ae22ULL, 0x7137449123ef65cdULL, 0xb5c0fbcfec4d3b2fULL, 0xe9b5dba58189dbbcULL,
    0x3956c25bf348b538ULL, 0x59f111f1b605d019ULL, 0x923f82a4af194f9bULL, 0xab1c5ed5da6d8118ULL,
    0xd807aa98a3030242ULL, 0x12835b0145706fbeULL, 0x243185be4ee4b28cULL, 0x550c7dc3d5ffb4e2ULL,
    0x72be5d74f27b896fULL, 0x80deb1fe3b1696b1ULL, 0x9bdc06a725c71235ULL, 0xc19bf174cf692694ULL,
    0xe49b69c19ef14ad2ULL, 0xefbe4786384f25e3ULL, 0x0fc19dc68b8cd5b5ULL, 0x240ca1cc77ac9c65ULL,
    0x2de92c6f592b0275ULL, 0x4a7484aa6ea6e483ULL, 0x5cb0a9dcbd41fbd4ULL, 0x76f988da831153b5ULL,
    0x983e5152ee66dfabULL, 0xa831c66d2db43210ULL, 0xb00327c898fb213fULL, 0xbf597fc7beef0ee4ULL,
    0xc6e00bf33da88fc2ULL, 0xd5a79147930aa725ULL, 0x06ca6351e003826fULL, 0x142929670a0e6e70ULL,
    0x27b70a8546d22ffcULL, 0x2e1b21385c26c926ULL, 0x4d2c6dfc5ac42aedULL, 0x53380d139d95b3dfULL,
    0x650a73548baf63deULL, 0x766a0abb3c77b2a8ULL, 0x81c2c92e47edaee6ULL, 0x92722c851482353bULL,
    0xa2bfe8a14cf10364ULL, 0xa81a664bbc423001ULL, 0xc24b8b70d0f89791ULL, 0xc76c51a30654be30ULL,
    0xd192e819d6ef5218ULL, 0xd69906245565a910ULL, 0xf40e35855771202aULL, 0x106aa07032bbd1b8ULL,
    0x19a4c116b8d2d0c8ULL, 0x1e376c085141ab53ULL, 0x2748774cdf8eeb99ULL, 0x34b0bcb5e19b48a8ULL,
    0x391c0cb3c5c95a63ULL, 0x4ed8aa4ae3418acbULL, 0x5b9cca4f7763e373ULL, 0x682e6ff3d6b2b8a3ULL,
    0x748f82ee5defb2fcULL, 0x78a5636f43172f60ULL, 0x84c87814a1f0ab72ULL, 0x8cc702081a6439ecULL,
    0x90befffa23631e28ULL, 0xa4506cebde82bde9ULL, 0xbef9a3f7b2c67915ULL, 0xc67178f2e372532bULL,
    0xca273eceea26619cULL, 0xd186b8c721c0c207ULL, 0xeada7dd6cde0eb1eULL, 0xf57d4f7fee6ed178ULL,
    0x06f067aa72176fbaULL, 0x0a637dc5a2c898a6ULL, 0x113f9804bef90daeULL, 0x1b710b35131c471bULL,
    0x28db77f523047d84ULL, 0x32caab7b40c72493ULL, 0x3c9ebe0a15c9bebcULL, 0x431d67c49c100d4cULL,
    0x4cc5d4becb3e42b6ULL, 0x597f299cfc657e2aULL, 0x5fcb6fab3ad6faecULL, 0x6c44198c4a475817ULL};

static uint64_t ror64(uint64_t x, int r) { return (x >> r) | (x << (64 - r)); }

static void sha512_block(uint64_t h[8], const unsigned char *blk)
{
    uint64_t w[80], a, b, c, d, e, f, g, hh, t1, t2;
    int t, j;
    for (t = 0; t < 16; t++) {
        w[t] = 0;
        for (j = 0; j < 8; j++) {
            w[t] = (w[t] << 8) | blk[8 * t + j];
        }
    }
    for (t = 16; t < 80; t++) {
        uint64_t s0 = ror64(w[t - 15], 1) ^ ror64(w[t - 15], 8) ^ (w[t - 15] >> 7);
        uint64_t s1 = ror64(w[t - 2], 19) ^ ror64(w[t - 2], 61) ^ (w[t - 2] >> 6);
        w[t] = w[t - 16] + s0 + w[t - 7] + s1;
    }
    a = h[0]; b = h[1]; c = h[2]; d = h[3]; e = h[4]; f = h[5]; g = h[6]; hh = h[7];
    for (t = 0; t < 80; t++) {
        t1 = hh + (ror64(e, 14) ^ ror64(e, 18) ^ ror64(e, 41)) + ((e & f) ^ (~e & g)) + K512[t] + w[t];
        t2 = (ror64(a, 28) ^ ror64(a, 34) ^ ror64(a, 39)) + ((a & b) ^ (a & c) ^ (b & c));
        hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
}

void xo_sha512(unsigned char out[64], const unsigned char *in, unsigned long long inlen)
{
    uint64_t h[8] = {0x6a09e667f3bcc908ULL, 0xbb67ae8584caa73bULL, 0x3c6ef372fe94f82bULL,
                     0xa54ff53a5f1d36f1ULL, 0x510e527fade682d1ULL, 0x9b05688c2b3e6c1fULL,
                     0x1f83d9abfb41bd6bULL, 0x5be0cd19137e2179ULL};
    unsigned char tail[256];
    unsigned long long bits = inlen * 8, rem, i, tl;
    while (inlen >= 128) {
        sha512_block(h, in);
        in += 128;
        inlen -= 128;
    }
    rem = inlen;
    memset(tail, 0, sizeof tail);
    memcpy(tail, in, rem);
    tail[rem] = 0x80;
    tl = (rem + 17 <= 128) ? 128 : 256;
    /* 128-bit length field; messages here are < 2^61 bytes, high half 0. */
    for (i = 0; i < 8; i++) {
        tail[tl - 1 - i] = (unsigned char)(bits >> (8 * i));
    }
    sha512_block(h, tail);
    if (tl == 256) {
        sha512_block(h, tail + 128);
    }
    for (i = 0; i < 8; i++) {
        for (int j = 0; j < 8; j++) {
            out[8 * i + j] = (unsigned char)(h[i] >> (56 - 8 * j));
        }
    }
}

/* ------------------------------------------------------------------ */
/* Keccak-f1600 and SHAKE (FIPS 202)                                   */
/* ------------------------------------------------------------------ */

static int keccak_ready;
static unsigned int keccak_rho[25]; /* rotation of lane x + 5y */
static uint64_t keccak_rc[24];

/* rc(t) of FIPS 202 Algorithm 5: the LFSR x^8 + x^6 + x^5 + x^4 + 1. */
static int keccak_rc_bit(int t)
{
    unsigned int r = 1, i;
    if (t % 255 == 0) {
        return 1;
    }
    for (i = 1; i <= (unsigned int)(t % 255); i++) {
        r <<= 1;
        if (r & 0x100) {
            r ^= 0x171; /* x^8 = x^6 + x^5 + x^4 + 1 */
        }
    }
    return r & 1;
}

static void keccak_init_tables(void)
{
    unsigned int x = 1, y = 0, t, j, nx;
    keccak_rho[0] = 0;
    for (t = 0; t < 24; t++) {
        keccak_rho[x + 5 * y] = ((t + 1) * (t + 2) / 2) % 64;
        nx = y;
        y = (2 * x + 3 * y) % 5;
        x = nx;
    }
    for (t = 0; t < 24; t++) {
        uint64_t rc = 0;
        for (j = 0; j <= 6; j++) {
            if (keccak_rc_bit((int)(j + 7 * t))) {
                rc |= (uint64_t)1 << ((1u << j) - 1);
            }
        }
        keccak_rc[t] = rc;
    }
    keccak_ready = 1;
}

static uint64_t rol64(uint64_t v, unsigned int r) { return r ? (v << r) | (v >> (64 - r)) : v; }

void xo_keccak_f1600(uint64_t a[25])
{
    uint64_t c[5], d[5], b[25];
    int r, x, y;
    if (!keccak_ready) {
        keccak_init_tables();
    }
    for (r = 0; r < 24; r++) {
        for (x = 0; x < 5; x++) {
            c[x] = a[x] ^ a[x + 5] ^ a[x + 10] ^ a[x + 15] ^ a[x + 20];
        }
        for (x = 0; x < 5; x++) {
            d[x] = c[(x + 4) % 5] ^ rol64(c[(x + 1) % 5], 1);
        }
        for (y = 0; y < 5; y++) {
            for (x = 0; x < 5; x++) {
                /* rho then pi: B[y][2x+3y] = rot(A[x][y]) */
                b[y + 5 * ((2 * x + 3 * y) % 5)] = rol64(a[x + 5 * y] ^ d[x], keccak_rho[x + 5 * y]);
            }
        }
        for (y = 0; y < 5; y++) {
            for (x = 0; x < 5; x++) {
                a[x + 5 * y] = b[x + 5 * y] ^ (~b[(x + 1) % 5 + 5 * y] & b[(x + 2) % 5 + 5 * y]);
            }
        }
        a[0] ^= keccak_rc[r];
    }
}

static void shake(unsigned int rate, unsigned char *out, unsigned long long outlen,
                  const unsigned char *in, unsigned long long inlen)
{
    uint64_t s[25];
    unsigned char blk[200];
    unsigned long long i;
    memset(s, 0, sizeof s);
    /* absorb full blocks */
    while (inlen >= rate) {
        for (i = 0; i < rate; i++) {
            s[i / 8] ^= (uint64_t)in[i] << (8 * (i % 8));
        }
        xo_keccak_f1600(s);
        in += rate;
        inlen -= rate;
    }
    /* last block with SHAKE domain bits 1111 and pad10*1 */
    memset(blk, 0, sizeof blk);
    memcpy(blk, in, inlen);
    blk[inlen] ^= 0x1F;
    blk[rate - 1] ^= 0x80;
    for (i = 0; i < rate; i++) {
        s[i / 8] ^= (uint64_t)blk[i] << (8 * (i % 8));
    }
    /* squeeze */
    while (outlen > 0) {
        unsigned long long take = outlen < rate ? outlen : rate;
        xo_keccak_f1600(s);
        for (i = 0; i < take; i++) {
            out[i] = (unsigned char)(s[i / 8] >> (8 * (i % 8)));
        }
        out += take;
        outlen -= take;
    }
}

void xo_shake128(unsigned char *out, unsigned long long outlen,
                 const unsigned char *in, unsigned long long inlen)
{
    shake(168, out, outlen, in, inlen);
}

void xo_shake256(unsigned char *out, unsigned long long outlen,
                 const unsigned char *in, unsigned long long inlen)
{
    shake(136, out, outlen, in, inlen);
}

/* ------------------------------------------------------------------ */
/* Seeded randombytes: SplitMix64 outputs, each emitted as 8 little-    */
/* endian bytes; bytes are consumed in order across calls.  Stands in   */
/* for the reference's /dev/urandom reader (randombytes.c:10-41) so     */
/* that keypairs are reproducible.                                      */
/* ------------------------------------------------------------------ */

static uint64_t rng_state;
static unsigned char rng_buf[8];
static unsigned int rng_left;

void xo_rng_seed(uint64_t seed)
{
    rng_state = seed;
    rng_left = 0;
}

static uint64_t splitmix64(void)
{
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void xo_randombytes(unsigned char *x, unsigned long long xlen)
{
    while (xlen > 0) {
        if (rng_left == 0) {
            uint64_t v = splitmix64();
            for (int i = 0; i < 8; i++) {
                rng_buf[i] = (unsigned char)(v >> (8 * i));
            }
            rng_left = 8;
        }
        *x++ = rng_buf[8 - rng_left];
        rng_left--;
        xlen--;
    }
}
