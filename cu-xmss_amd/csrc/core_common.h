// core_common.h -- helpers shared by the plain and BDS cores.
#pragma once
#include <stdint.h>
#include <string.h>

#include <mutex>
#include <new>

#include "../../include/params.h"
#include "engine.h"

namespace xm {

// utils.c:6-30
inline void ull_to_bytes(unsigned char* out, unsigned outlen, unsigned long long in) {
    for (int i = (int)outlen - 1; i >= 0; --i) {
        out[i] = (unsigned char)(in & 0xff);
        in >>= 8;
    }
}
inline unsigned long long bytes_to_ull(const unsigned char* in, unsigned inlen) {
    unsigned long long r = 0;
    for (unsigned i = 0; i < inlen; ++i) r |= ((unsigned long long)in[i]) << (8 * (inlen - 1 - i));
    return r;
}

// Index check of both cores (xmss_core.c:222-231, xmss_core_fast.c:742-751,
// :970-979): from idx = 2^h - 1 on the key is wiped (index bytes 0xFF, the
// rest zero); returns -2 when no one-time key is left, 0 otherwise (the last
// index still signs, with the wiped key).
inline int index_check(const xmss_params* p, unsigned char* sk, unsigned long long idx) {
    if (idx >= ((1ULL << p->full_height) - 1)) {
        memset(sk, 0xFF, p->index_bytes);
        memset(sk + p->index_bytes, 0, (size_t)(p->sk_bytes - p->index_bytes));
        if (idx > ((1ULL << p->full_height) - 1)) return -2;
        if (p->full_height == 64 && idx == ((1ULL << p->full_height) - 1)) return -2;
    }
    return 0;
}

// All device work of the library is serialised on the engine of the current
// device (one stream); callers on other threads wait here.
std::mutex& engine_mutex();

// Return code when no GPU is usable (the library has no CPU path) or a device
// call failed.
constexpr int XMSS_ENODEV = -3;

// Longest message the device path hashes: an H_msg descriptor holds the
// prefixed message length in 32 bits (toByte(2, pad) || R || root ||
// toByte(idx, n) || M, plus a block of padding).  Longer messages are
// rejected with -1 before any byte is moved and with the sk unchanged,
// where the reference would hash them (cuxmss/hash.c:107-120).
constexpr unsigned long long kMaxMsgBytes = 0xFFFFF000ull;

// Runs f (an entry point's body) and turns a device failure into -3: the
// engine of the current device is brought back to a consistent state and
// the caller's process goes on.
template <class Fn>
int guarded(Fn&& f) {
    try {
        return f();
    } catch (const DeviceError&) {
    } catch (const std::bad_alloc&) {
    }
    try {
        std::lock_guard<std::mutex> lk(engine_mutex());
        if (Engine* e = Engine::get()) e->recover();
    } catch (...) {
    }
    return XMSS_ENODEV;
}

// Return code of the allocation-free entry point for a request larger than
// its static device scratch.
constexpr int XMSS_ETOOBIG = -4;

[[noreturn]] void fatal(const char* what);

// Shared verification (xmss_commons.c:136-215); no_alloc: the
// allocation-free form (xmss_mi355x_sign_open_static).
int core_sign_open(const xmss_params* p, unsigned char* m, unsigned long long* mlen, const unsigned char* sm,
                   unsigned long long smlen, const unsigned char* pk, bool no_alloc);

}  // namespace xm
