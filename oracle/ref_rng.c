/*
 * ref_rng.c -- seeded stand-in for the reference's randombytes()
 * (cuxmss/randombytes.c:10-41 reads /dev/urandom).  Linked only into the
 * oracle/_ref builds of the reference sources, so that their keypairs are
 * reproducible.  Same SplitMix64 byte stream as oracle/xo_hash.c
 * (xo_rng_seed / xo_randombytes).  Test infrastructure only.
 */
#include <stdint.h>

static uint64_t st;
static unsigned char buf[8];
static unsigned int left;

void ref_rng_seed(uint64_t seed)
{
    st = seed;
    left = 0;
}

void randombytes(unsigned char *x, unsigned long long xlen)
{
    while (xlen > 0) {
        if (left == 0) {
            uint64_t z = (st += 0x9E3779B97F4A7C15ULL);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
            z ^= z >> 31;
            for (int i = 0; i < 8; i++) {
                buf[i] = (unsigned char)(z >> (8 * i));
            }
            left = 8;
        }
        *x++ = buf[8 - left];
        left--;
        xlen--;
    }
}
