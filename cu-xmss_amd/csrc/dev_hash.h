// dev_hash.h -- hash compression functions for gfx950 (wave64, integer VALU).
//
// SHA-256 / SHA-512 follow FIPS 180-4, Keccak-f1600 follows FIPS 202.
// Everything is written so that hipcc fully unrolls it: message words that are
// compile-time constants (padding, zero prefixes, address words fixed by the
// caller) fold away, and the remaining work maps onto v_alignbit_b32 (rotate),
// v_bitop3_b32 (three-input XOR and majority in one op) and v_add3_u32.
// The rounds keep the whole state and message schedule in VGPRs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xm {

// ---------------------------------------------------------------------------
// 32-bit helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t rotr32(uint32_t x, int n) {
    return __builtin_amdgcn_alignbit(x, x, n);
}
// a ^ b ^ c as one v_bitop3_b32 (truth table 0x96, symmetric in its inputs)
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}
// majority(a, b, c) as one v_bitop3_b32 (truth table 0xE8, symmetric)
__device__ __forceinline__ uint32_t maj3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0xE8);
}
__device__ __forceinline__ uint32_t bswap32(uint32_t x) { return __builtin_bswap32(x); }

// ---------------------------------------------------------------------------
// SHA-256 (FIPS 180-4 6.2).  State: 8 words; block: 16 big-endian words.
// ---------------------------------------------------------------------------
__device__ __constant__ static const uint32_t kSha256K[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

// The same constants as an immediate-friendly constexpr table, so that fully
// unrolled rounds take K[t] as a literal instead of a memory load.
struct Sha256Consts {
    static constexpr uint32_t K[64] = {
        0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
        0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
        0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
        0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
        0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
        0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
        0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
        0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};
    static constexpr uint32_t IV[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                                       0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
};

// Message-schedule word t >= 16 in the rolling 16-word buffer.
__device__ __forceinline__ uint32_t sha256_sched(const uint32_t w[16], int t) {
    const uint32_t w15 = w[(t + 1) & 15], w2 = w[(t + 14) & 15];
    const uint32_t s0 = xor3(rotr32(w15, 7), rotr32(w15, 18), w15 >> 3);
    const uint32_t s1 = xor3(rotr32(w2, 17), rotr32(w2, 19), w2 >> 10);
    return w[t & 15] + s0 + s1 + w[(t + 9) & 15];
}

// Rounds T0 .. T1-1 on the working variables v = {a, ..., h}.  Rounds from
// SCHED on extend the schedule in the rolling buffer first; rounds before
// SCHED (and >= 16) take w[t & 15] as already holding W[t], which lets a
// caller supply schedule words it computed once for many compressions.
template <int T0, int T1, int SCHED = 16>
__device__ __forceinline__ void sha256_rounds(uint32_t v[8], uint32_t w[16]) {
    uint32_t a = v[0], b = v[1], c = v[2], d = v[3];
    uint32_t e = v[4], f = v[5], g = v[6], h = v[7];
#pragma unroll
    for (int t = T0; t < T1; ++t) {
        if (t >= 16 && t >= SCHED) w[t & 15] = sha256_sched(w, t);
        const uint32_t S1 = xor3(rotr32(e, 6), rotr32(e, 11), rotr32(e, 25));
        const uint32_t ch = (e & f) ^ (~e & g);
        const uint32_t t1 = h + S1 + ch + Sha256Consts::K[t] + w[t & 15];
        const uint32_t S0 = xor3(rotr32(a, 2), rotr32(a, 13), rotr32(a, 22));
        const uint32_t t2 = S0 + maj3(a, b, c);
        h = g; g = f; f = e; e = d + t1;
        d = c; c = b; b = a; a = t1 + t2;
    }
    v[0] = a; v[1] = b; v[2] = c; v[3] = d; v[4] = e; v[5] = f; v[6] = g; v[7] = h;
}

// Block b of a len-byte string at p (p 16-byte aligned, readable kMsgSlack
// bytes past the end): sha256_block_load issues the four 16-byte loads,
// sha256_block_words turns them into the 16 message words (big-endian
// values) with the FIPS 180-4 padding of the last one or two blocks.
__device__ __forceinline__ void sha256_block_load(uint4 (&raw)[4], const uint8_t* p, uint32_t b) {
    const uint4* qb = reinterpret_cast<const uint4*>(p + 64 * (size_t)b);
#pragma unroll
    for (int j = 0; j < 4; ++j) raw[j] = qb[j];
}

__device__ __forceinline__ void sha256_block_words(uint32_t w[16], const uint4 (&raw)[4], uint32_t len, uint32_t b) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        w[4 * j] = bswap32(raw[j].x);
        w[4 * j + 1] = bswap32(raw[j].y);
        w[4 * j + 2] = bswap32(raw[j].z);
        w[4 * j + 3] = bswap32(raw[j].w);
    }
    if (64 * (b + 1) <= len) return;  // message bytes only
    const int nb = (int)((len + 9 + 63) / 64);
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int rem = (int)len - (int)(64 * b + 4 * j);  // message bytes from this word on
        if (rem >= 4) continue;
        if (rem > 0) w[j] = (w[j] & (0xFFFFFFFFu << (32 - 8 * rem))) | (0x80u << (24 - 8 * rem));
        else if (rem == 0) w[j] = 0x80000000u;
        else w[j] = 0;
    }
    if ((int)b == nb - 1) {
        w[14] = 0;
        w[15] = len * 8u;
    }
}

// The 64 round inputs W[t] + K[t] of one block from its 16 message words:
// out(t, value) receives them in order (the message schedule and the
// constant, everything of a compression that does not depend on the state).
template <class Out>
__device__ __forceinline__ void sha256_schedule_k(uint32_t w[16], Out&& out) {
#pragma unroll
    for (int t = 0; t < 64; ++t) {
        if (t >= 16) w[t & 15] = sha256_sched(w, t);
        out(t, w[t & 15] + Sha256Consts::K[t]);
    }
}

// The rounds of one compression from precomputed W[t] + K[t] (wk(t)):
// st <- st + F(st, w), with no schedule work on this thread.
template <class In>
__device__ __forceinline__ void sha256_compress_wk(uint32_t st[8], In&& wk) {
    uint32_t a = st[0], b = st[1], c = st[2], d = st[3];
    uint32_t e = st[4], f = st[5], g = st[6], h = st[7];
#pragma unroll
    for (int t = 0; t < 64; ++t) {
        const uint32_t S1 = xor3(rotr32(e, 6), rotr32(e, 11), rotr32(e, 25));
        const uint32_t ch = (e & f) ^ (~e & g);
        const uint32_t t1 = h + S1 + ch + wk(t);
        const uint32_t S0 = xor3(rotr32(a, 2), rotr32(a, 13), rotr32(a, 22));
        const uint32_t t2 = S0 + maj3(a, b, c);
        h = g; g = f; f = e; e = d + t1;
        d = c; c = b; b = a; a = t1 + t2;
    }
    st[0] += a; st[1] += b; st[2] += c; st[3] += d;
    st[4] += e; st[5] += f; st[6] += g; st[7] += h;
}

// One compression: st <- st + F(st, w).  w is clobbered (rolling schedule).
__device__ __forceinline__ void sha256_compress(uint32_t st[8], uint32_t w[16]) {
    uint32_t v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = st[i];
    sha256_rounds<0, 64>(v, w);
#pragma unroll
    for (int i = 0; i < 8; ++i) st[i] += v[i];
}

__device__ __forceinline__ void sha256_iv(uint32_t st[8]) {
#pragma unroll
    for (int i = 0; i < 8; ++i) st[i] = Sha256Consts::IV[i];
}

// Absorbs NW message words (already big-endian values) that start on a
// block boundary, appends the FIPS 180-4 padding for a message of
// `total_bytes` bytes in all (prefix blocks already in st included), and
// finishes.  NW*4 must equal total_bytes minus a whole number of blocks.
template <int NW>
__device__ __forceinline__ void sha256_tail(uint32_t st[8], const uint32_t (&m)[NW],
                                            uint32_t total_bytes) {
    constexpr int nblocks = (NW * 4 + 9 + 63) / 64;
    uint32_t w[16];
#pragma unroll
    for (int blk = 0; blk < nblocks; ++blk) {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int k = blk * 16 + j;
            uint32_t v;
            if (k < NW) v = m[k];
            else if (k == NW) v = 0x80000000u;
            else v = 0;
            if (blk == nblocks - 1 && j == 14) v = 0;  // length high word (< 2^32 bits)
            if (blk == nblocks - 1 && j == 15) v = total_bytes * 8u;
            w[j] = v;
        }
        sha256_compress(st, w);
    }
}

// ---------------------------------------------------------------------------
// SHA-512 (FIPS 180-4 6.4).  64-bit words as uint64_t; the compiler splits
// them into v_alignbit pairs (rotates), v_lshl_add_u64 (adds) and bitop3
// halves (logic).
// ---------------------------------------------------------------------------
struct Sha512Consts {
    static constexpr uint64_t K[80] = {
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
    static constexpr uint64_t IV[8] = {0x6a09e667f3bcc908ULL, 0xbb67ae8584caa73bULL,
                                       0x3c6ef372fe94f82bULL, 0xa54ff53a5f1d36f1ULL,
                                       0x510e527fade682d1ULL, 0x9b05688c2b3e6c1fULL,
                                       0x1f83d9abfb41bd6bULL, 0x5be0cd19137e2179ULL};
};

__device__ __forceinline__ uint64_t rotr64(uint64_t x, int n) {
    const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    uint32_t rlo, rhi;
    if (n < 32) {
        rlo = __builtin_amdgcn_alignbit(hi, lo, n);
        rhi = __builtin_amdgcn_alignbit(lo, hi, n);
    } else {
        rlo = __builtin_amdgcn_alignbit(lo, hi, n - 32);
        rhi = __builtin_amdgcn_alignbit(hi, lo, n - 32);
    }
    return ((uint64_t)rhi << 32) | rlo;
}
// Two 32-bit halves as one 64-bit register pair.  The empty asm hides the
// halves from the optimizer, so a following 64-bit add stays one
// v_lshl_add_u64 instead of being split into an add of the low half (with a
// zero high half) and a separate add into the high half.
__device__ __forceinline__ uint64_t pair64(uint32_t lo, uint32_t hi) {
    uint64_t x = ((uint64_t)hi << 32) | lo;
    asm("" : "+v"(x));
    return x;
}
__device__ __forceinline__ uint64_t xor3_64(uint64_t a, uint64_t b, uint64_t c) {
    const uint32_t lo = xor3((uint32_t)a, (uint32_t)b, (uint32_t)c);
    const uint32_t hi = xor3((uint32_t)(a >> 32), (uint32_t)(b >> 32), (uint32_t)(c >> 32));
    return pair64(lo, hi);
}
__device__ __forceinline__ uint64_t maj3_64(uint64_t a, uint64_t b, uint64_t c) {
    const uint32_t lo = maj3((uint32_t)a, (uint32_t)b, (uint32_t)c);
    const uint32_t hi = maj3((uint32_t)(a >> 32), (uint32_t)(b >> 32), (uint32_t)(c >> 32));
    return pair64(lo, hi);
}
// (e & f) ^ (~e & g) per half (one v_bitop3 each)
__device__ __forceinline__ uint64_t ch_64(uint64_t e, uint64_t f, uint64_t g) {
    const uint32_t lo = __builtin_amdgcn_bitop3_b32((uint32_t)e, (uint32_t)f, (uint32_t)g, 0xCA);
    const uint32_t hi = __builtin_amdgcn_bitop3_b32((uint32_t)(e >> 32), (uint32_t)(f >> 32), (uint32_t)(g >> 32), 0xCA);
    return pair64(lo, hi);
}

__device__ __constant__ static const uint64_t kSha512K[80] = {
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

template <bool SCHED>
__device__ __forceinline__ void sha512_round(uint64_t* v, uint64_t* w, int t, int j, uint64_t k) {
    // v holds a..h rotated so that round j of an 8-round group reads fixed slots
    uint64_t& a = v[(8 - j) & 7];
    uint64_t& b = v[(9 - j) & 7];
    uint64_t& c = v[(10 - j) & 7];
    uint64_t& d = v[(11 - j) & 7];
    uint64_t& e = v[(12 - j) & 7];
    uint64_t& f = v[(13 - j) & 7];
    uint64_t& g = v[(14 - j) & 7];
    uint64_t& h = v[(15 - j) & 7];
    (void)t;
    const int s = j & 15;
    if (SCHED) {
        const uint64_t w15 = w[(s + 1) & 15], w2 = w[(s + 14) & 15];
        const uint64_t s0 = xor3_64(rotr64(w15, 1), rotr64(w15, 8), w15 >> 7);
        const uint64_t s1 = xor3_64(rotr64(w2, 19), rotr64(w2, 61), w2 >> 6);
        w[s] = w[s] + s0 + s1 + w[(s + 9) & 15];
    }
    const uint64_t S1 = xor3_64(rotr64(e, 14), rotr64(e, 18), rotr64(e, 41));
    const uint64_t ch = ch_64(e, f, g);
    const uint64_t t1 = h + S1 + ch + k + w[s];
    const uint64_t S0 = xor3_64(rotr64(a, 28), rotr64(a, 34), rotr64(a, 39));
    d += t1;
    h = t1 + S0 + maj3_64(a, b, c);
}

// One SHA-512 compression.  Rounds 0..15 are unrolled (constant message
// words fold); rounds 16..79 run as a loop of 16-round groups with K read
// through the scalar cache, which keeps the scheduling window -- and so the
// number of live 64-bit values -- small enough to avoid spills.
__device__ __forceinline__ void sha512_compress(uint64_t st[8], uint64_t w[16]) {
    uint64_t v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = st[i];
#pragma unroll
    for (int j = 0; j < 16; ++j) sha512_round<false>(v, w, j, j, Sha512Consts::K[j]);
#pragma unroll 1
    for (int t0 = 16; t0 < 80; t0 += 16) {
#pragma unroll
        for (int j = 0; j < 16; ++j) sha512_round<true>(v, w, t0 + j, j, kSha512K[t0 + j]);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) st[i] += v[i];
}

__device__ __forceinline__ void sha512_iv(uint64_t st[8]) {
#pragma unroll
    for (int i = 0; i < 8; ++i) st[i] = Sha512Consts::IV[i];
}

// As sha256_tail, for 128-byte blocks and a 128-bit length field.
template <int NW>
__device__ __forceinline__ void sha512_tail(uint64_t st[8], const uint64_t (&m)[NW],
                                            uint32_t total_bytes) {
    constexpr int nblocks = (NW * 8 + 17 + 127) / 128;
    uint64_t w[16];
#pragma unroll
    for (int blk = 0; blk < nblocks; ++blk) {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int k = blk * 16 + j;
            uint64_t v;
            if (k < NW) v = m[k];
            else if (k == NW) v = 0x8000000000000000ULL;
            else v = 0;
            if (blk == nblocks - 1 && j == 14) v = 0;
            if (blk == nblocks - 1 && j == 15) v = (uint64_t)total_bytes * 8u;
            w[j] = v;
        }
        sha512_compress(st, w);
    }
}

// ---------------------------------------------------------------------------
// Keccak-f1600 (FIPS 202 3.3).  Lane A[x][y] is s[x + 5y], little-endian.
// ---------------------------------------------------------------------------
struct KeccakConsts {
    static constexpr uint64_t RC[24] = {
        0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808aULL, 0x8000000080008000ULL,
        0x000000000000808bULL, 0x0000000080000001ULL, 0x8000000080008081ULL, 0x8000000000008009ULL,
        0x000000000000008aULL, 0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000aULL,
        0x000000008000808bULL, 0x800000000000008bULL, 0x8000000000008089ULL, 0x8000000000008003ULL,
        0x8000000000008002ULL, 0x8000000000000080ULL, 0x000000000000800aULL, 0x800000008000000aULL,
        0x8000000080008081ULL, 0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL};
    // rho offsets r[x + 5y]
    static constexpr int RHO[25] = {0, 1, 62, 28, 27, 36, 44, 6, 55, 20, 3, 10, 43,
                                    25, 39, 41, 45, 15, 21, 8, 18, 2, 61, 56, 14};
};

__device__ __forceinline__ uint64_t rotl64(uint64_t x, int n) {
    return n == 0 ? x : rotr64(x, 64 - n);
}

// One 64-bit lane as two 32-bit halves: every step of the round maps onto
// one 32-bit instruction per half.  Per round: theta 20 (column parities,
// two v_bitop3 xor3 per half) + 10 (rotl 1 of the parities, v_alignbit) +
// 50 (one xor3 per half-lane), rho 48 (two v_alignbit per lane, lane 0 is
// not rotated), chi 50 (one v_bitop3 a ^ (~b & c) per half-lane), iota 2:
// 180 + 2 instructions; pi is register renaming.
template <int N>
__device__ __forceinline__ void rotl64_halves(uint32_t& lo, uint32_t& hi, uint32_t l, uint32_t h) {
    if constexpr (N == 0) {
        lo = l;
        hi = h;
    } else if constexpr (N < 32) {
        hi = __builtin_amdgcn_alignbit(h, l, 32 - N);
        lo = __builtin_amdgcn_alignbit(l, h, 32 - N);
    } else if constexpr (N == 32) {
        lo = h;
        hi = l;
    } else {
        hi = __builtin_amdgcn_alignbit(l, h, 64 - N);
        lo = __builtin_amdgcn_alignbit(h, l, 64 - N);
    }
}

// a ^ (~b & c) as one v_bitop3_b32: the truth table is the function applied
// to the patterns 0xF0, 0xCC, 0xAA of (src0, src1, src2).
__device__ __forceinline__ uint32_t chi3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0xD2);
}

struct KeccakRC32 {
    static constexpr uint32_t LO[24] = {
        0x00000001u, 0x00008082u, 0x0000808au, 0x80008000u, 0x0000808bu, 0x80000001u, 0x80008081u, 0x00008009u,
        0x0000008au, 0x00000088u, 0x80008009u, 0x8000000au, 0x8000808bu, 0x0000008bu, 0x00008089u, 0x00008003u,
        0x00008002u, 0x00000080u, 0x0000800au, 0x8000000au, 0x80008081u, 0x00008080u, 0x80000001u, 0x80008008u};
    static constexpr uint32_t HI[24] = {
        0x00000000u, 0x00000000u, 0x80000000u, 0x80000000u, 0x00000000u, 0x00000000u, 0x80000000u, 0x80000000u,
        0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x80000000u, 0x80000000u, 0x80000000u,
        0x80000000u, 0x80000000u, 0x00000000u, 0x80000000u, 0x80000000u, 0x80000000u, 0x00000000u, 0x80000000u};
};

template <int X, int Y>
__device__ __forceinline__ void keccak_theta_rho_pi(uint32_t (&BL)[25], uint32_t (&BH)[25], const uint32_t (&L)[25],
                                                    const uint32_t (&H)[25], const uint32_t (&cl)[5],
                                                    const uint32_t (&ch)[5], const uint32_t (&rl)[5],
                                                    const uint32_t (&rh)[5]) {
    constexpr int i = X + 5 * Y;
    constexpr int j = Y + 5 * ((2 * X + 3 * Y) % 5);
    const uint32_t l = xor3(L[i], cl[(X + 4) % 5], rl[(X + 1) % 5]);
    const uint32_t h = xor3(H[i], ch[(X + 4) % 5], rh[(X + 1) % 5]);
    rotl64_halves<KeccakConsts::RHO[i]>(BL[j], BH[j], l, h);
}

template <int Y>
__device__ __forceinline__ void keccak_row(uint32_t (&BL)[25], uint32_t (&BH)[25], const uint32_t (&L)[25],
                                           const uint32_t (&H)[25], const uint32_t (&cl)[5], const uint32_t (&ch)[5],
                                           const uint32_t (&rl)[5], const uint32_t (&rh)[5]) {
    keccak_theta_rho_pi<0, Y>(BL, BH, L, H, cl, ch, rl, rh);
    keccak_theta_rho_pi<1, Y>(BL, BH, L, H, cl, ch, rl, rh);
    keccak_theta_rho_pi<2, Y>(BL, BH, L, H, cl, ch, rl, rh);
    keccak_theta_rho_pi<3, Y>(BL, BH, L, H, cl, ch, rl, rh);
    keccak_theta_rho_pi<4, Y>(BL, BH, L, H, cl, ch, rl, rh);
}

__device__ __forceinline__ void keccak_f1600(uint64_t s[25]) {
    uint32_t L[25], H[25];
#pragma unroll
    for (int i = 0; i < 25; ++i) {
        L[i] = (uint32_t)s[i];
        H[i] = (uint32_t)(s[i] >> 32);
    }
#pragma unroll 2
    for (int r = 0; r < 24; ++r) {
        uint32_t cl[5], ch[5], rl[5], rh[5];
#pragma unroll
        for (int x = 0; x < 5; ++x) {
            cl[x] = xor3(xor3(L[x], L[x + 5], L[x + 10]), L[x + 15], L[x + 20]);
            ch[x] = xor3(xor3(H[x], H[x + 5], H[x + 10]), H[x + 15], H[x + 20]);
        }
#pragma unroll
        for (int x = 0; x < 5; ++x) {  // rotl64(c[x], 1)
            rl[x] = __builtin_amdgcn_alignbit(cl[x], ch[x], 31);
            rh[x] = __builtin_amdgcn_alignbit(ch[x], cl[x], 31);
        }
        uint32_t BL[25], BH[25];
        keccak_row<0>(BL, BH, L, H, cl, ch, rl, rh);
        keccak_row<1>(BL, BH, L, H, cl, ch, rl, rh);
        keccak_row<2>(BL, BH, L, H, cl, ch, rl, rh);
        keccak_row<3>(BL, BH, L, H, cl, ch, rl, rh);
        keccak_row<4>(BL, BH, L, H, cl, ch, rl, rh);
#pragma unroll
        for (int y = 0; y < 5; ++y)
#pragma unroll
            for (int x = 0; x < 5; ++x) {
                L[x + 5 * y] = chi3(BL[x + 5 * y], BL[(x + 1) % 5 + 5 * y], BL[(x + 2) % 5 + 5 * y]);
                H[x + 5 * y] = chi3(BH[x + 5 * y], BH[(x + 1) % 5 + 5 * y], BH[(x + 2) % 5 + 5 * y]);
            }
        L[0] ^= KeccakRC32::LO[r];
        H[0] ^= KeccakRC32::HI[r];
    }
#pragma unroll
    for (int i = 0; i < 25; ++i) s[i] = ((uint64_t)H[i] << 32) | L[i];
}

}  // namespace xm
