// rng.cpp -- randombytes() for keypair generation.
//
// Default: /dev/urandom, as cuxmss/randombytes.c:10-41.  After
// xmss_mi355x_rng_seed(s) it is a reproducible stream instead: successive
// SplitMix64 outputs, each emitted as 8 little-endian bytes, consumed in order
// across calls (the harness stream used to compare keypairs with the
// reference byte for byte).
#include <fcntl.h>
#include <stdint.h>
#include <unistd.h>

#include <mutex>

#include "../../include/randombytes.h"
#include "../../include/xmss_mi355x.h"

namespace {
std::mutex g_mu;
bool g_seeded = false;
uint64_t g_state = 0;
unsigned char g_buf[8];
unsigned g_left = 0;
int g_fd = -1;

uint64_t splitmix64() {
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
}  // namespace

extern "C" {

void xmss_mi355x_rng_seed(uint64_t seed) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_seeded = true;
    g_state = seed;
    g_left = 0;
}

void xmss_mi355x_rng_urandom(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_seeded = false;
}

void randombytes(unsigned char* x, unsigned long long xlen) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_seeded) {
        while (xlen > 0) {
            if (g_left == 0) {
                const uint64_t v = splitmix64();
                for (int i = 0; i < 8; ++i) g_buf[i] = (unsigned char)(v >> (8 * i));
                g_left = 8;
            }
            *x++ = g_buf[8 - g_left];
            --g_left;
            --xlen;
        }
        return;
    }
    while (g_fd == -1) {
        g_fd = open("/dev/urandom", O_RDONLY);
        if (g_fd == -1) sleep(1);
    }
    while (xlen > 0) {
        const size_t want = xlen < 1048576 ? (size_t)xlen : 1048576;
        const ssize_t got = read(g_fd, x, want);
        if (got < 1) {
            sleep(1);
            continue;
        }
        x += got;
        xlen -= (unsigned long long)got;
    }
}

}  // extern "C"
