// dev_family.h -- the tweakable-hash layer of XMSS on the device, one policy
// class per hash family (the (func, n) pairs that core_hash accepts,
// cuxmss/hash.c:29-65).
//
// Each policy provides, on nodes held in registers:
//   prepare(key)                 per-key precomputation (SHA-2 midstates)
//   prf_keygen(key, adrs)        hash.c:87-98   PRF_keygen(SK_SEED, PUB_SEED || ADRS)
//   thash_f(key, adrs, x)        hash.c:157-183 F with key and bitmask from PRF
//   thash_h(key, adrs, l, r)     hash.c:125-155 H with key and two bitmasks
//   prf_idx(key, idx)            hash.c:70-81   R = PRF(SK_PRF, toByte(idx, 32))
//   load / store                 node <-> n bytes in device memory
//   hash_bytes(p, len)           core_hash over a byte string in device memory
//
// ADRS is passed as its eight 32-bit values (hash_address.c).  Word 7
// (key_and_mask) is set inside thash_f / thash_h as the reference does.
//
// Message words that are constant (domain-separation pads, padding, zero
// words) are written as literals so that the fully unrolled compressions fold
// them; for n = 32 (SHA-256) and n = 64 (SHA-512) the one-block prefix
// toByte(pad) || KEY of every PRF is absorbed once per key (midstate).
#pragma once
#include "dev_hash.h"

namespace xm {

enum : int { PAD_F = 0, PAD_H = 1, PAD_HASH = 2, PAD_PRF = 3, PAD_PRF_KEYGEN = 4 };

// ===========================================================================
// SHA-256, n = 24 (pad 4, output truncated, hash.c:36-39) and n = 32
// ===========================================================================
template <int N>
struct Sha256Fam {
    static constexpr int kN = N;
    static constexpr bool kChainPrf = (N == 32);  // chain_prf_init / chain_prf_step available
    static constexpr bool kPairHmsg = true;       // H_msg by two waves (k_hmsg_pair)
    static constexpr int kChainWaves = 1;         // no occupancy request for k_chains
    static constexpr int NW = N / 4;         // node words (big-endian values)
    static constexpr int LEN = 8 * N / 4 + 3;  // wots_len for w = 16
    static constexpr int BLOCK = 64;
    using Word = uint32_t;
    struct Node { uint32_t w[NW]; };
    struct Key {
        uint32_t pub[NW], sks[NW], prfk[NW];
        uint32_t ms_pub[8], ms_sks[8], ms_prfk[8];  // n = 32: state after toByte(pad,32)||key
    };

    __device__ static void load(Node& x, const uint8_t* p) {
        const uint32_t* q = reinterpret_cast<const uint32_t*>(p);
#pragma unroll
        for (int i = 0; i < NW; ++i) x.w[i] = bswap32(q[i]);
    }
    __device__ static void store(const Node& x, uint8_t* p) {
        uint32_t* q = reinterpret_cast<uint32_t*>(p);
#pragma unroll
        for (int i = 0; i < NW; ++i) q[i] = bswap32(x.w[i]);
    }

    __device__ static void midstate(uint32_t out[8], uint32_t pad, const uint32_t k[NW]) {
        uint32_t w[16];
#pragma unroll
        for (int i = 0; i < 7; ++i) w[i] = 0;
        w[7] = pad;
#pragma unroll
        for (int i = 0; i < 8; ++i) w[8 + i] = k[i];
        sha256_iv(out);
        sha256_compress(out, w);
    }
    __device__ static void prepare(Key& k) {
        if constexpr (N == 32) {
            midstate(k.ms_pub, PAD_PRF, k.pub);
            midstate(k.ms_sks, PAD_PRF_KEYGEN, k.sks);
            midstate(k.ms_prfk, PAD_PRF, k.prfk);
        }
    }

    // PRF(KEY, in) with in = 8 words; ms/key select the key (hash.c:70-81).
    __device__ static void prf8(Node& out, const uint32_t ms[8], const uint32_t key[NW],
                                const uint32_t (&in)[8]) {
        uint32_t st[8];
        if constexpr (N == 32) {
#pragma unroll
            for (int i = 0; i < 8; ++i) st[i] = ms[i];
            sha256_tail<8>(st, in, 96);
        } else {
            uint32_t m[1 + NW + 8];
            m[0] = PAD_PRF;
#pragma unroll
            for (int i = 0; i < NW; ++i) m[1 + i] = key[i];
#pragma unroll
            for (int i = 0; i < 8; ++i) m[1 + NW + i] = in[i];
            sha256_iv(st);
            sha256_tail<1 + NW + 8>(st, m, 4 + N + 32);
        }
#pragma unroll
        for (int i = 0; i < NW; ++i) out.w[i] = st[i];
    }

    // ---- n = 32: PRF(PUB_SEED, ADRS) computations that share work ----------
    // The PRF message block after the midstate is ADRS || padding: words 0..7
    // are ADRS, 8..15 constants.  Two PRFs whose ADRS agree in words 0..K-1
    // share rounds 0..K-1; schedule words W16..W20 depend only on words 0..5
    // and the constants, W21 also on word 6 (the hash address of a chain step).

    // State after rounds 0..5 and W16..W20 for one WOTS chain (ADRS words
    // 0..5 = layer, tree, type, ots, chain are fixed along the chain).
    struct ChainPrf {
        uint32_t v6[8];   // working variables after round 5
        uint32_t w[16];   // W16..W20 in slots 0..4, W5 in slot 5 (for W21)
    };
    __device__ static void chain_prf_init(ChainPrf& c, const Key& k, const uint32_t (&a)[8]) {
        uint32_t w[16];
#pragma unroll
        for (int i = 0; i < 8; ++i) w[i] = a[i];
        w[6] = 0;
        w[7] = 0;
        w[8] = 0x80000000u;
#pragma unroll
        for (int i = 9; i < 15; ++i) w[i] = 0;
        w[15] = 96 * 8;
#pragma unroll
        for (int i = 0; i < 8; ++i) c.v6[i] = k.ms_pub[i];
        sha256_rounds<0, 6>(c.v6, w);
        // W16..W20 (independent of words 6 and 7)
        uint32_t x[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] = w[i];
#pragma unroll
        for (int t = 16; t < 21; ++t) x[t & 15] = sha256_sched(x, t);
#pragma unroll
        for (int i = 0; i < 5; ++i) c.w[i] = x[i];
        c.w[5] = a[5];
    }

    // key = PRF(ADRS with hash = i, km = 0), mask = PRF(km = 1) for one chain step.
    __device__ static void chain_prf_step(Node& key, Node& mask, const Key& k, const ChainPrf& c,
                                          uint32_t i) {
        uint32_t w[16];
        w[6] = i;
        w[7] = 0;
        w[8] = 0x80000000u;
#pragma unroll
        for (int j = 9; j < 15; ++j) w[j] = 0;
        w[15] = 96 * 8;
        // W21 = sigma1(W19) + W14 + sigma0(W6) + W5, shared by key and mask
        const uint32_t w19 = c.w[3];
        const uint32_t w21 = xor3(rotr32(w19, 17), rotr32(w19, 19), w19 >> 10) + 0u +
                             xor3(rotr32(i, 7), rotr32(i, 18), i >> 3) + c.w[5];
        w[0] = c.w[0]; w[1] = c.w[1]; w[2] = c.w[2]; w[3] = c.w[3]; w[4] = c.w[4];
        w[5] = w21;
        uint32_t v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = c.v6[j];
        sha256_rounds<6, 7>(v, w);  // round 6 (W6 = i) is shared as well
        uint32_t v1[8], w1[16];
#pragma unroll
        for (int j = 0; j < 8; ++j) v1[j] = v[j];
#pragma unroll
        for (int j = 0; j < 16; ++j) w1[j] = w[j];
        w1[7] = 1;
        sha256_rounds<7, 64, 22>(v, w);
        sha256_rounds<7, 64, 22>(v1, w1);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            key.w[j] = k.ms_pub[j] + v[j];
            mask.w[j] = k.ms_pub[j] + v1[j];
        }
    }

    // F(key, x ^ mask) given key and mask (the outer hash of thash_f).
    __device__ static void f_outer(Node& out, const Node& key, const Node& mask, const Node& x) {
        constexpr int P = (N == 32) ? 8 : 1;
        uint32_t m[P + 2 * NW];
#pragma unroll
        for (int i = 0; i < P; ++i) m[i] = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i) { m[P + i] = key.w[i]; m[P + NW + i] = x.w[i] ^ mask.w[i]; }
        uint32_t st[8];
        sha256_iv(st);
        sha256_tail<P + 2 * NW>(st, m, 4 * P + 2 * N);
#pragma unroll
        for (int i = 0; i < NW; ++i) out.w[i] = st[i];
    }

    // M PRFs of one ADRS with key_and_mask = 0 .. M-1 (thash_f: key, mask;
    // thash_h: key, mask0, mask1): ADRS words 0..6 are common, so rounds
    // 0..6 and W16..W21 are computed once.
    template <int M>
    __device__ static void prf_km(Node* outs, const Key& k, const uint32_t (&a)[8]) {
        uint32_t w[16];
#pragma unroll
        for (int i = 0; i < 7; ++i) w[i] = a[i];
        w[7] = 0;
        w[8] = 0x80000000u;
#pragma unroll
        for (int i = 9; i < 15; ++i) w[i] = 0;
        w[15] = 96 * 8;
        uint32_t v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = k.ms_pub[i];
        sha256_rounds<0, 7>(v, w);
        uint32_t x[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] = w[i];
#pragma unroll
        for (int t = 16; t < 22; ++t) x[t & 15] = sha256_sched(x, t);
#pragma unroll
        for (int i = 0; i < 6; ++i) w[i] = x[i];  // W16..W21 into slots 0..5
#pragma unroll
        for (int km = 0; km < M; ++km) {
            uint32_t vv[8], ww[16];
#pragma unroll
            for (int j = 0; j < 8; ++j) vv[j] = v[j];
#pragma unroll
            for (int j = 0; j < 16; ++j) ww[j] = w[j];
            ww[7] = (uint32_t)km;
            sha256_rounds<7, 64, 22>(vv, ww);
#pragma unroll
            for (int j = 0; j < 8; ++j) outs[km].w[j] = k.ms_pub[j] + vv[j];
        }
    }

    __device__ static void prf_keygen(Node& out, const Key& k, const uint32_t (&a)[8]) {
        uint32_t st[8];
        if constexpr (N == 32) {
            uint32_t m[16];
#pragma unroll
            for (int i = 0; i < 8; ++i) { m[i] = k.pub[i]; m[8 + i] = a[i]; }
#pragma unroll
            for (int i = 0; i < 8; ++i) st[i] = k.ms_sks[i];
            sha256_tail<16>(st, m, 128);
        } else {
            uint32_t m[1 + 2 * NW + 8];
            m[0] = PAD_PRF_KEYGEN;
#pragma unroll
            for (int i = 0; i < NW; ++i) { m[1 + i] = k.sks[i]; m[1 + NW + i] = k.pub[i]; }
#pragma unroll
            for (int i = 0; i < 8; ++i) m[1 + 2 * NW + i] = a[i];
            sha256_iv(st);
            sha256_tail<1 + 2 * NW + 8>(st, m, 4 + 2 * N + 32);
        }
#pragma unroll
        for (int i = 0; i < NW; ++i) out.w[i] = st[i];
    }

    __device__ static void thash_f(Node& out, const Key& k, uint32_t (&a)[8], const Node& x) {
        Node key, mask;
        if constexpr (N == 32) {
            Node km[2];
            prf_km<2>(km, k, a);
            key = km[0];
            mask = km[1];
            a[7] = 1;
        } else {
            a[7] = 0;
            prf8(key, k.ms_pub, k.pub, a);
            a[7] = 1;
            prf8(mask, k.ms_pub, k.pub, a);
        }
        constexpr int P = (N == 32) ? 8 : 1;
        uint32_t m[P + 2 * NW];
#pragma unroll
        for (int i = 0; i < P; ++i) m[i] = 0;  // toByte(0, pad)
#pragma unroll
        for (int i = 0; i < NW; ++i) { m[P + i] = key.w[i]; m[P + NW + i] = x.w[i] ^ mask.w[i]; }
        uint32_t st[8];
        sha256_iv(st);
        sha256_tail<P + 2 * NW>(st, m, 4 * P + 2 * N);
#pragma unroll
        for (int i = 0; i < NW; ++i) out.w[i] = st[i];
    }

    __device__ static void thash_h(Node& out, const Key& k, uint32_t (&a)[8], const Node& l,
                                   const Node& r) {
        Node key, m0, m1;
        if constexpr (N == 32) {
            Node km[3];
            prf_km<3>(km, k, a);
            key = km[0];
            m0 = km[1];
            m1 = km[2];
            a[7] = 2;
        } else {
            a[7] = 0;
            prf8(key, k.ms_pub, k.pub, a);
            a[7] = 1;
            prf8(m0, k.ms_pub, k.pub, a);
            a[7] = 2;
            prf8(m1, k.ms_pub, k.pub, a);
        }
        constexpr int P = (N == 32) ? 8 : 1;
        uint32_t m[P + 3 * NW];
#pragma unroll
        for (int i = 0; i < P - 1; ++i) m[i] = 0;
        m[P - 1] = PAD_H;
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            m[P + i] = key.w[i];
            m[P + NW + i] = l.w[i] ^ m0.w[i];
            m[P + 2 * NW + i] = r.w[i] ^ m1.w[i];
        }
        uint32_t st[8];
        sha256_iv(st);
        sha256_tail<P + 3 * NW>(st, m, 4 * P + 3 * N);
#pragma unroll
        for (int i = 0; i < NW; ++i) out.w[i] = st[i];
    }

    // ---- split tweakable hashes ---------------------------------------
    // Everything in thash_f / thash_h that depends only on the address -- the
    // PRF key and masks and, for n = 32, the first block of the outer hash
    // (toByte(0|1, 32) || key) -- is computed by pre_f / pre_h; fin_f / fin_h
    // then finish with the chain value or the two children: one compression
    // (F) or two (H) for n = 32.  fin(pre(a), x) == thash(a, x) bit for bit.
    static constexpr int PREW = (N == 32) ? 8 : NW;  // midstate, or the key
    struct FPre { uint32_t s[PREW]; Node mask; };
    struct HPre { uint32_t s[PREW]; Node m0, m1; };

    __device__ static void pre_f(FPre& p, const Key& k, uint32_t (&a)[8]) {
        Node key;
        if constexpr (N == 32) {
            Node km[2];
            prf_km<2>(km, k, a);
            key = km[0];
            p.mask = km[1];
            uint32_t w[16];
#pragma unroll
            for (int i = 0; i < 8; ++i) { w[i] = 0; w[8 + i] = key.w[i]; }  // toByte(0, 32) || key
            sha256_iv(p.s);
            sha256_compress(p.s, w);
        } else {
            a[7] = 0;
            prf8(key, k.ms_pub, k.pub, a);
            a[7] = 1;
            prf8(p.mask, k.ms_pub, k.pub, a);
#pragma unroll
            for (int i = 0; i < NW; ++i) p.s[i] = key.w[i];
        }
    }
    __device__ static void fin_f(Node& out, const FPre& p, const Node& x) {
        uint32_t st[8];
        if constexpr (N == 32) {
            uint32_t m[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) { st[i] = p.s[i]; m[i] = x.w[i] ^ p.mask.w[i]; }
            sha256_tail<8>(st, m, 96);
        } else {
            uint32_t m[1 + 2 * NW];
            m[0] = PAD_F;
#pragma unroll
            for (int i = 0; i < NW; ++i) { m[1 + i] = p.s[i]; m[1 + NW + i] = x.w[i] ^ p.mask.w[i]; }
            sha256_iv(st);
            sha256_tail<1 + 2 * NW>(st, m, 4 + 2 * N);
        }
#pragma unroll
        for (int i = 0; i < NW; ++i) out.w[i] = st[i];
    }
    __device__ static void pre_h(HPre& p, const Key& k, uint32_t (&a)[8]) {
        Node key;
        if constexpr (N == 32) {
            Node km[3];
            prf_km<3>(km, k, a);
            key = km[0];
            p.m0 = km[1];
            p.m1 = km[2];
            uint32_t w[16];
#pragma unroll
            for (int i = 0; i < 7; ++i) w[i] = 0;
            w[7] = PAD_H;  // toByte(1, 32) || key
#pragma unroll
            for (int i = 0; i < 8; ++i) w[8 + i] = key.w[i];
            sha256_iv(p.s);
            sha256_compress(p.s, w);
        } else {
            a[7] = 0;
            prf8(key, k.ms_pub, k.pub, a);
            a[7] = 1;
            prf8(p.m0, k.ms_pub, k.pub, a);
            a[7] = 2;
            prf8(p.m1, k.ms_pub, k.pub, a);
#pragma unroll
            for (int i = 0; i < NW; ++i) p.s[i] = key.w[i];
        }
    }
    __device__ static void fin_h(Node& out, const HPre& p, const Node& l, const Node& r) {
        uint32_t st[8];
        if constexpr (N == 32) {
            uint32_t m[16];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                st[i] = p.s[i];
                m[i] = l.w[i] ^ p.m0.w[i];
                m[8 + i] = r.w[i] ^ p.m1.w[i];
            }
            sha256_tail<16>(st, m, 128);
        } else {
            uint32_t m[1 + 3 * NW];
            m[0] = PAD_H;
#pragma unroll
            for (int i = 0; i < NW; ++i) {
                m[1 + i] = p.s[i];
                m[1 + NW + i] = l.w[i] ^ p.m0.w[i];
                m[1 + 2 * NW + i] = r.w[i] ^ p.m1.w[i];
            }
            sha256_iv(st);
            sha256_tail<1 + 3 * NW>(st, m, 4 + 3 * N);
        }
#pragma unroll
        for (int i = 0; i < NW; ++i) out.w[i] = st[i];
    }

    __device__ static void prf_idx(Node& out, const Key& k, uint64_t idx) {
        uint32_t in[8] = {0, 0, 0, 0, 0, 0, (uint32_t)(idx >> 32), (uint32_t)idx};
        prf8(out, k.ms_prfk, k.prfk, in);
    }

    // SHA-256 of len bytes at p (p 16-byte aligned, readable kMsgSlack bytes
    // past the end), output truncated to n bytes.
    __device__ static void hash_bytes(Node& out, const uint8_t* p, uint32_t len) {
        uint32_t st[8];
        sha256_iv(st);
        const uint32_t nb = (len + 9 + 63) / 64;
        // blocks made only of message bytes: plain loads, the next block's
        // words in flight while this one compresses
        // (p is 16-byte aligned: four 16-byte loads per block)
        const uint32_t nfull = len / 64;
        uint32_t b = 0;
        if (nfull) {
            const uint4* q4 = reinterpret_cast<const uint4*>(p);
            uint4 nx[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) nx[j] = q4[j];
            for (; b < nfull; ++b) {
                uint32_t w[16];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    w[4 * j] = bswap32(nx[j].x);
                    w[4 * j + 1] = bswap32(nx[j].y);
                    w[4 * j + 2] = bswap32(nx[j].z);
                    w[4 * j + 3] = bswap32(nx[j].w);
                }
                const uint4* qn = q4 + 4 * (b + 1 < nfull ? b + 1 : b);
#pragma unroll
                for (int j = 0; j < 4; ++j) nx[j] = qn[j];
                sha256_compress(st, w);
            }
        }
        // the last one or two blocks: whole-block loads (the bytes past the
        // message are readable, kMsgSlack), then the padding in registers
        for (; b < nb; ++b) {
            const uint4* qb = reinterpret_cast<const uint4*>(p + 64 * (size_t)b);
            uint4 t[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) t[j] = qb[j];
            uint32_t w[16];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                w[4 * j] = t[j].x;
                w[4 * j + 1] = t[j].y;
                w[4 * j + 2] = t[j].z;
                w[4 * j + 3] = t[j].w;
            }
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int rem = (int)len - (int)(64 * b + 4 * j);  // message bytes from this word on
                const uint32_t v = bswap32(w[j]);
                if (rem >= 4) w[j] = v;
                else if (rem > 0) w[j] = (v & (0xFFFFFFFFu << (32 - 8 * rem))) | (0x80u << (24 - 8 * rem));
                else if (rem == 0) w[j] = 0x80000000u;
                else w[j] = 0;
            }
            if (b == nb - 1) {
                w[14] = 0;
                w[15] = len * 8u;
            }
            sha256_compress(st, w);
        }
#pragma unroll
        for (int i = 0; i < NW; ++i) out.w[i] = st[i];
    }
};

// ===========================================================================
// SHA-512, n = 64 (pad 64), vendored self_SHA512 in the reference
// (hash.c:54-56).  Node words are big-endian 64-bit values.
// ===========================================================================
struct Sha512Fam {
    static constexpr int kN = 64;
    static constexpr bool kChainPrf = false;
    static constexpr bool kPairHmsg = false;
    static constexpr int kChainWaves = 4;  // k_chains within 128 VGPRs
    static constexpr int NW = 8;
    static constexpr int LEN = 8 * 64 / 4 + 3;  // 131
    using Word = uint64_t;
    struct Node { uint64_t w[8]; };
    struct Key {
        uint64_t pub[8], sks[8], prfk[8];
        uint64_t ms_pub[8], ms_sks[8], ms_prfk[8];
    };

    __device__ static uint64_t bswap64(uint64_t x) { return __builtin_bswap64(x); }
    __device__ static void load(Node& x, const uint8_t* p) {
        const uint64_t* q = reinterpret_cast<const uint64_t*>(p);
#pragma unroll
        for (int i = 0; i < 8; ++i) x.w[i] = bswap64(q[i]);
    }
    __device__ static void store(const Node& x, uint8_t* p) {
        uint64_t* q = reinterpret_cast<uint64_t*>(p);
#pragma unroll
        for (int i = 0; i < 8; ++i) q[i] = bswap64(x.w[i]);
    }
    __device__ static void adrs64(uint64_t (&o)[4], const uint32_t (&a)[8]) {
#pragma unroll
        for (int i = 0; i < 4; ++i) o[i] = ((uint64_t)a[2 * i] << 32) | a[2 * i + 1];
    }
    __device__ static void midstate(uint64_t out[8], uint64_t pad, const uint64_t k[8]) {
        uint64_t w[16];
#pragma unroll
        for (int i = 0; i < 7; ++i) w[i] = 0;
        w[7] = pad;
#pragma unroll
        for (int i = 0; i < 8; ++i) w[8 + i] = k[i];
        sha512_iv(out);
        sha512_compress(out, w);
    }
    __device__ static void prepare(Key& k) {
        midstate(k.ms_pub, PAD_PRF, k.pub);
        midstate(k.ms_sks, PAD_PRF_KEYGEN, k.sks);
        midstate(k.ms_prfk, PAD_PRF, k.prfk);
    }
    __device__ static void prf4(Node& out, const uint64_t ms[8], const uint64_t (&in)[4]) {
        uint64_t st[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) st[i] = ms[i];
        sha512_tail<4>(st, in, 128 + 32);
#pragma unroll
        for (int i = 0; i < 8; ++i) out.w[i] = st[i];
    }
    __device__ static void prf_keygen(Node& out, const Key& k, const uint32_t (&a)[8]) {
        uint64_t m[12], ad[4];
        adrs64(ad, a);
#pragma unroll
        for (int i = 0; i < 8; ++i) m[i] = k.pub[i];
#pragma unroll
        for (int i = 0; i < 4; ++i) m[8 + i] = ad[i];
        uint64_t st[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) st[i] = k.ms_sks[i];
        sha512_tail<12>(st, m, 128 + 96);
#pragma unroll
        for (int i = 0; i < 8; ++i) out.w[i] = st[i];
    }
    // The outer hash absorbs toByte(pad) || key first and the masked input
    // after it, so each PRF result is consumed as soon as it exists: at most
    // one 512-bit PRF output is live next to the running state.
    __device__ static void absorb_block(uint64_t st[8], const uint64_t (&lo)[8], const uint64_t (&hi)[8]) {
        uint64_t w[16];
#pragma unroll
        for (int i = 0; i < 8; ++i) { w[i] = lo[i]; w[8 + i] = hi[i]; }
        sha512_compress(st, w);
    }
    __device__ static void thash_f(Node& out, const Key& k, uint32_t (&a)[8], const Node& x) {
        Node t;
        uint64_t ad[4];
        uint64_t st[8];
        sha512_iv(st);
        a[7] = 0;
        adrs64(ad, a);
        prf4(t, k.ms_pub, ad);  // key
        {
            const uint64_t z[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            absorb_block(st, z, t.w);
        }
        a[7] = 1;
        adrs64(ad, a);
        prf4(t, k.ms_pub, ad);  // mask
        {
            uint64_t xm[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) xm[i] = x.w[i] ^ t.w[i];
            const uint64_t pad[8] = {0x8000000000000000ULL, 0, 0, 0, 0, 0, 0, 192 * 8};
            absorb_block(st, xm, pad);
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) out.w[i] = st[i];
    }
    __device__ __forceinline__ static void thash_h(Node& out, const Key& k, uint32_t (&a)[8], const Node& l,
                                   const Node& r) {
        Node t;
        uint64_t ad[4];
        uint64_t st[8];
        sha512_iv(st);
        a[7] = 0;
        adrs64(ad, a);
        prf4(t, k.ms_pub, ad);  // key
        {
            const uint64_t p[8] = {0, 0, 0, 0, 0, 0, 0, (uint64_t)PAD_H};
            absorb_block(st, p, t.w);
        }
        uint64_t lm[8];
        a[7] = 1;
        adrs64(ad, a);
        prf4(t, k.ms_pub, ad);  // mask 0
#pragma unroll
        for (int i = 0; i < 8; ++i) lm[i] = l.w[i] ^ t.w[i];
        a[7] = 2;
        adrs64(ad, a);
        prf4(t, k.ms_pub, ad);  // mask 1
        {
            uint64_t rm[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) rm[i] = r.w[i] ^ t.w[i];
            absorb_block(st, lm, rm);
        }
        {
            const uint64_t p0[8] = {0x8000000000000000ULL, 0, 0, 0, 0, 0, 0, 0};
            const uint64_t p1[8] = {0, 0, 0, 0, 0, 0, 0, 256 * 8};
            absorb_block(st, p0, p1);
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) out.w[i] = st[i];
    }
    // ---- split tweakable hashes (see Sha256Fam): pre_* leaves the state
    // after toByte(0|1, 64) || key and the masks; fin_f is one compression,
    // fin_h two.
    struct FPre { uint64_t s[8]; Node mask; };
    struct HPre { uint64_t s[8]; Node m0, m1; };
    __device__ static void pre_f(FPre& p, const Key& k, uint32_t (&a)[8]) {
        Node key;
        uint64_t ad[4];
        a[7] = 0;
        adrs64(ad, a);
        prf4(key, k.ms_pub, ad);
        a[7] = 1;
        adrs64(ad, a);
        prf4(p.mask, k.ms_pub, ad);
        const uint64_t z[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        sha512_iv(p.s);
        absorb_block(p.s, z, key.w);
    }
    __device__ static void fin_f(Node& out, const FPre& p, const Node& x) {
        uint64_t st[8], xm[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) { st[i] = p.s[i]; xm[i] = x.w[i] ^ p.mask.w[i]; }
        const uint64_t pad[8] = {0x8000000000000000ULL, 0, 0, 0, 0, 0, 0, 192 * 8};
        absorb_block(st, xm, pad);
#pragma unroll
        for (int i = 0; i < 8; ++i) out.w[i] = st[i];
    }
    __device__ static void pre_h(HPre& p, const Key& k, uint32_t (&a)[8]) {
        Node key;
        uint64_t ad[4];
        a[7] = 0;
        adrs64(ad, a);
        prf4(key, k.ms_pub, ad);
        a[7] = 1;
        adrs64(ad, a);
        prf4(p.m0, k.ms_pub, ad);
        a[7] = 2;
        adrs64(ad, a);
        prf4(p.m1, k.ms_pub, ad);
        const uint64_t pw[8] = {0, 0, 0, 0, 0, 0, 0, (uint64_t)PAD_H};
        sha512_iv(p.s);
        absorb_block(p.s, pw, key.w);
    }
    __device__ static void fin_h(Node& out, const HPre& p, const Node& l, const Node& r) {
        uint64_t st[8], lm[8], rm[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            st[i] = p.s[i];
            lm[i] = l.w[i] ^ p.m0.w[i];
            rm[i] = r.w[i] ^ p.m1.w[i];
        }
        absorb_block(st, lm, rm);
        const uint64_t p0[8] = {0x8000000000000000ULL, 0, 0, 0, 0, 0, 0, 0};
        const uint64_t p1[8] = {0, 0, 0, 0, 0, 0, 0, 256 * 8};
        absorb_block(st, p0, p1);
#pragma unroll
        for (int i = 0; i < 8; ++i) out.w[i] = st[i];
    }
    __device__ static void prf_idx(Node& out, const Key& k, uint64_t idx) {
        uint64_t in[4] = {0, 0, 0, idx};
        prf4(out, k.ms_prfk, in);
    }
    __device__ static void hash_bytes(Node& out, const uint8_t* p, uint32_t len) {
        uint64_t st[8];
        sha512_iv(st);
        const uint32_t nb = (len + 17 + 127) / 128;
        // blocks made only of message bytes: plain loads, the next block's
        // words in flight while this one compresses
        // (p is 16-byte aligned: eight 16-byte loads per block)
        const uint32_t nfull = len / 128;
        uint32_t b = 0;
        if (nfull) {
            const uint4* q4 = reinterpret_cast<const uint4*>(p);
            uint4 nx[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) nx[j] = q4[j];
            for (; b < nfull; ++b) {
                uint64_t w[16];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    w[2 * j] = ((uint64_t)bswap32(nx[j].x) << 32) | bswap32(nx[j].y);
                    w[2 * j + 1] = ((uint64_t)bswap32(nx[j].z) << 32) | bswap32(nx[j].w);
                }
                const uint4* qn = q4 + 8 * (b + 1 < nfull ? b + 1 : b);
#pragma unroll
                for (int j = 0; j < 8; ++j) nx[j] = qn[j];
                sha512_compress(st, w);
            }
        }
        // the last one or two blocks: whole-block loads (the bytes past the
        // message are readable, kMsgSlack), then the padding in registers
        for (; b < nb; ++b) {
            const uint4* qb = reinterpret_cast<const uint4*>(p + 128 * (size_t)b);
            uint32_t raw[32];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const uint4 t = qb[j];
                raw[4 * j] = t.x;
                raw[4 * j + 1] = t.y;
                raw[4 * j + 2] = t.z;
                raw[4 * j + 3] = t.w;
            }
            uint64_t w[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                uint64_t v = 0;
#pragma unroll
                for (int h = 0; h < 2; ++h) {  // two 32-bit halves, big-endian
                    const int rem = (int)len - (int)(128 * b + 8 * j + 4 * h);
                    const uint32_t x = bswap32(raw[2 * j + h]);
                    uint32_t u;
                    if (rem >= 4) u = x;
                    else if (rem > 0) u = (x & (0xFFFFFFFFu << (32 - 8 * rem))) | (0x80u << (24 - 8 * rem));
                    else if (rem == 0) u = 0x80000000u;
                    else u = 0;
                    v |= (uint64_t)u << (h ? 0 : 32);
                }
                w[j] = v;
            }
            if (b == nb - 1) {
                w[14] = 0;
                w[15] = (uint64_t)len * 8u;
            }
            sha512_compress(st, w);
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) out.w[i] = st[i];
    }
};

// ===========================================================================
// SHAKE128 (rate 168) / SHAKE256 (rate 136), n = 24 (pad 4), 32, 64
// (hash.c:41-59).  Messages and nodes are kept as little-endian 32-bit words
// of the byte string; pairs of words form the 64-bit Keccak lanes.
// ===========================================================================
template <int RATE, int N>
struct ShakeFam {
    static constexpr int kN = N;
    static constexpr bool kChainPrf = false;
    static constexpr bool kPairHmsg = false;
    static constexpr int kChainWaves = 1;
    static constexpr int NW = N / 4;
    static constexpr int LEN = 8 * N / 4 + 3;
    static constexpr int PADW = (N == 24) ? 1 : N / 4;  // padding_len in words
    using Word = uint32_t;
    struct Node { uint32_t w[NW]; };
    struct Key { uint32_t pub[NW], sks[NW], prfk[NW]; };

    __device__ static void load(Node& x, const uint8_t* p) {
        const uint32_t* q = reinterpret_cast<const uint32_t*>(p);
#pragma unroll
        for (int i = 0; i < NW; ++i) x.w[i] = q[i];
    }
    __device__ static void store(const Node& x, uint8_t* p) {
        uint32_t* q = reinterpret_cast<uint32_t*>(p);
#pragma unroll
        for (int i = 0; i < NW; ++i) q[i] = x.w[i];
    }
    __device__ static void prepare(Key&) {}

    // SHAKE over a message of M words (4*M bytes, compile-time length).
    template <int M>
    __device__ static void shake_words(Node& out, const uint32_t (&m)[M]) {
        constexpr int LEN_B = 4 * M;
        constexpr int NBLK = LEN_B / RATE + 1;
        uint64_t s[25];
#pragma unroll
        for (int i = 0; i < 25; ++i) s[i] = 0;
#pragma unroll
        for (int b = 0; b < NBLK; ++b) {
#pragma unroll
            for (int j = 0; j < RATE / 8; ++j) {
                uint32_t lohi[2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int byte = b * RATE + 8 * j + 4 * h;
                    uint32_t v = (byte / 4 < M) ? m[byte / 4] : 0u;
                    // domain separation 0x1F right after the message, 0x80 in
                    // the last byte of the last block (pad10*1)
                    if (b == NBLK - 1) {
                        if (byte <= LEN_B && LEN_B < byte + 4) v ^= 0x1Fu << (8 * (LEN_B - byte));
                        if (byte <= NBLK * RATE - 1 && NBLK * RATE - 1 < byte + 4)
                            v ^= 0x80u << (8 * (NBLK * RATE - 1 - byte));
                    }
                    lohi[h] = v;
                }
                s[j] ^= ((uint64_t)lohi[1] << 32) | lohi[0];
            }
            keccak_f1600(s);
        }
#pragma unroll
        for (int i = 0; i < NW; ++i) out.w[i] = (i & 1) ? (uint32_t)(s[i / 2] >> 32) : (uint32_t)s[i / 2];
    }

    __device__ static void pad_words(uint32_t* m, uint32_t v) {
#pragma unroll
        for (int i = 0; i < PADW - 1; ++i) m[i] = 0;
        m[PADW - 1] = v << 24;  // toByte(v, pad): last byte holds v
    }
    __device__ static void adrs_le(uint32_t* m, const uint32_t (&a)[8]) {
#pragma unroll
        for (int i = 0; i < 8; ++i) m[i] = bswap32(a[i]);
    }
    __device__ static void prf8(Node& out, const uint32_t key[NW], const uint32_t (&in_le)[8]) {
        uint32_t m[PADW + NW + 8];
        pad_words(m, PAD_PRF);
#pragma unroll
        for (int i = 0; i < NW; ++i) m[PADW + i] = key[i];
#pragma unroll
        for (int i = 0; i < 8; ++i) m[PADW + NW + i] = in_le[i];
        shake_words(out, m);
    }
    __device__ static void prf_keygen(Node& out, const Key& k, const uint32_t (&a)[8]) {
        uint32_t m[PADW + 2 * NW + 8];
        pad_words(m, PAD_PRF_KEYGEN);
#pragma unroll
        for (int i = 0; i < NW; ++i) { m[PADW + i] = k.sks[i]; m[PADW + NW + i] = k.pub[i]; }
        uint32_t al[8];
        adrs_le(al, a);
#pragma unroll
        for (int i = 0; i < 8; ++i) m[PADW + 2 * NW + i] = al[i];
        shake_words(out, m);
    }
    __device__ static void thash_f(Node& out, const Key& k, uint32_t (&a)[8], const Node& x) {
        Node key, mask;
        uint32_t al[8];
        a[7] = 0;
        adrs_le(al, a);
        prf8(key, k.pub, al);
        a[7] = 1;
        adrs_le(al, a);
        prf8(mask, k.pub, al);
        uint32_t m[PADW + 2 * NW];
        pad_words(m, PAD_F);
#pragma unroll
        for (int i = 0; i < NW; ++i) { m[PADW + i] = key.w[i]; m[PADW + NW + i] = x.w[i] ^ mask.w[i]; }
        shake_words(out, m);
    }
    __device__ static void thash_h(Node& out, const Key& k, uint32_t (&a)[8], const Node& l,
                                   const Node& r) {
        Node key, m0, m1;
        uint32_t al[8];
        a[7] = 0;
        adrs_le(al, a);
        prf8(key, k.pub, al);
        a[7] = 1;
        adrs_le(al, a);
        prf8(m0, k.pub, al);
        a[7] = 2;
        adrs_le(al, a);
        prf8(m1, k.pub, al);
        uint32_t m[PADW + 3 * NW];
        pad_words(m, PAD_H);
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            m[PADW + i] = key.w[i];
            m[PADW + NW + i] = l.w[i] ^ m0.w[i];
            m[PADW + 2 * NW + i] = r.w[i] ^ m1.w[i];
        }
        shake_words(out, m);
    }
    // ---- split tweakable hashes (see Sha256Fam): the sponge absorbs key and
    // masked input in one block, so pre_* keeps key and masks and fin_* is
    // the outer permutation.
    struct FPre { Node key, mask; };
    struct HPre { Node key, m0, m1; };
    __device__ static void pre_f(FPre& p, const Key& k, uint32_t (&a)[8]) {
        uint32_t al[8];
        a[7] = 0;
        adrs_le(al, a);
        prf8(p.key, k.pub, al);
        a[7] = 1;
        adrs_le(al, a);
        prf8(p.mask, k.pub, al);
    }
    __device__ static void fin_f(Node& out, const FPre& p, const Node& x) {
        uint32_t m[PADW + 2 * NW];
        pad_words(m, PAD_F);
#pragma unroll
        for (int i = 0; i < NW; ++i) { m[PADW + i] = p.key.w[i]; m[PADW + NW + i] = x.w[i] ^ p.mask.w[i]; }
        shake_words(out, m);
    }
    __device__ static void pre_h(HPre& p, const Key& k, uint32_t (&a)[8]) {
        uint32_t al[8];
        a[7] = 0;
        adrs_le(al, a);
        prf8(p.key, k.pub, al);
        a[7] = 1;
        adrs_le(al, a);
        prf8(p.m0, k.pub, al);
        a[7] = 2;
        adrs_le(al, a);
        prf8(p.m1, k.pub, al);
    }
    __device__ static void fin_h(Node& out, const HPre& p, const Node& l, const Node& r) {
        uint32_t m[PADW + 3 * NW];
        pad_words(m, PAD_H);
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            m[PADW + i] = p.key.w[i];
            m[PADW + NW + i] = l.w[i] ^ p.m0.w[i];
            m[PADW + 2 * NW + i] = r.w[i] ^ p.m1.w[i];
        }
        shake_words(out, m);
    }
    __device__ static void prf_idx(Node& out, const Key& k, uint64_t idx) {
        uint32_t in[8] = {0, 0, 0, 0, 0, 0, bswap32((uint32_t)(idx >> 32)), bswap32((uint32_t)idx)};
        prf8(out, k.prfk, in);
    }
    __device__ static void hash_bytes(Node& out, const uint8_t* p, uint32_t len) {
        uint64_t s[25];
#pragma unroll
        for (int i = 0; i < 25; ++i) s[i] = 0;
        const uint32_t nb = len / RATE + 1;
        // whole-block 8-byte loads (p is 16-byte aligned and RATE a multiple
        // of 8; the bytes past the message are readable, kMsgSlack), the next
        // block's words in flight while this one is permuted; the padding of
        // the last block is applied in registers
        const uint2* q = reinterpret_cast<const uint2*>(p);
        uint2 nx[RATE / 8];
#pragma unroll
        for (int j = 0; j < RATE / 8; ++j) nx[j] = q[j];
        for (uint32_t b = 0; b < nb; ++b) {
            uint2 cur[RATE / 8];
#pragma unroll
            for (int j = 0; j < RATE / 8; ++j) cur[j] = nx[j];
            const uint2* qn = q + (RATE / 8) * (b + 1 < nb ? b + 1 : b);
#pragma unroll
            for (int j = 0; j < RATE / 8; ++j) nx[j] = qn[j];
            const bool lastb = b + 1 == nb;
#pragma unroll
            for (int j = 0; j < RATE / 8; ++j) {
                uint32_t lohi[2] = {cur[j].x, cur[j].y};
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int rem = (int)len - (int)(b * RATE + 8 * j + 4 * h);
                    uint32_t v = lohi[h];
                    if (rem < 4) v = rem > 0 ? v & (0xFFFFFFFFu >> (32 - 8 * rem)) : 0;
                    if (rem >= 0 && rem < 4) v ^= 0x1Fu << (8 * rem);
                    if (lastb && 8 * j + 4 * h == RATE - 4) v ^= 0x80000000u;
                    lohi[h] = v;
                }
                s[j] ^= ((uint64_t)lohi[1] << 32) | lohi[0];
            }
            keccak_f1600(s);
        }
#pragma unroll
        for (int i = 0; i < NW; ++i) out.w[i] = (i & 1) ? (uint32_t)(s[i / 2] >> 32) : (uint32_t)s[i / 2];
    }
};

}  // namespace xm
