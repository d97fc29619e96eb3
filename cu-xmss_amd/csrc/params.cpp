// params.cpp -- parameter sets and OIDs (drop-in for cuxmss/params.c).
//
// The reference spells its tables out as if/else and switch ladders; here
// they are data: 21 XMSS sets and seven XMSS^MT families of eight
// (height, layers) shapes each.  Names, OIDs and derived sizes are identical
// (checked name by name against the reference build in tests/test_params.py).
#include <stdint.h>
#include <string.h>

#include <string>

#include "../../include/params.h"
#include "../../include/xmss_core.h"

namespace {

struct XmssSet {
    const char* name;
    uint32_t oid;
    unsigned func, n, h;
};

// cuxmss/params.c:9-67 (names) and :256-372 (func, n, height)
const XmssSet kXmss[] = {
    {"XMSS-SHA2_10_256", 0x01, XMSS_SHA2, 32, 10},        {"XMSS-SHA2_16_256", 0x02, XMSS_SHA2, 32, 16},
    {"XMSS-SHA2_20_256", 0x03, XMSS_SHA2, 32, 20},        {"XMSS-SHA2_10_512", 0x04, XMSS_SHA2, 64, 10},
    {"XMSS-SHA2_16_512", 0x05, XMSS_SHA2, 64, 16},        {"XMSS-SHA2_20_512", 0x06, XMSS_SHA2, 64, 20},
    {"XMSS-SHAKE_10_256", 0x07, XMSS_SHAKE128, 32, 10},   {"XMSS-SHAKE_16_256", 0x08, XMSS_SHAKE128, 32, 16},
    {"XMSS-SHAKE_20_256", 0x09, XMSS_SHAKE128, 32, 20},   {"XMSS-SHAKE_10_512", 0x0a, XMSS_SHAKE256, 64, 10},
    {"XMSS-SHAKE_16_512", 0x0b, XMSS_SHAKE256, 64, 16},   {"XMSS-SHAKE_20_512", 0x0c, XMSS_SHAKE256, 64, 20},
    {"XMSS-SHA2_10_192", 0x0d, XMSS_SHA2, 24, 10},        {"XMSS-SHA2_16_192", 0x0e, XMSS_SHA2, 24, 16},
    {"XMSS-SHA2_20_192", 0x0f, XMSS_SHA2, 24, 20},        {"XMSS-SHAKE256_10_256", 0x10, XMSS_SHAKE256, 32, 10},
    {"XMSS-SHAKE256_16_256", 0x11, XMSS_SHAKE256, 32, 16}, {"XMSS-SHAKE256_20_256", 0x12, XMSS_SHAKE256, 32, 20},
    {"XMSS-SHAKE256_10_192", 0x13, XMSS_SHAKE256, 24, 10}, {"XMSS-SHAKE256_16_192", 0x14, XMSS_SHAKE256, 24, 16},
    {"XMSS-SHAKE256_20_192", 0x15, XMSS_SHAKE256, 24, 20},
};

// cuxmss/params.c:69-252 and :374-698: OIDs come in runs of eight shapes.
struct MtFamily {
    const char* prefix;
    const char* suffix;
    unsigned func, n;
    uint32_t first_oid;
};
const MtFamily kMtFamilies[] = {
    {"XMSSMT-SHA2_", "_256", XMSS_SHA2, 32, 0x01},     {"XMSSMT-SHA2_", "_512", XMSS_SHA2, 64, 0x09},
    {"XMSSMT-SHAKE_", "_256", XMSS_SHAKE128, 32, 0x11}, {"XMSSMT-SHAKE_", "_512", XMSS_SHAKE256, 64, 0x19},
    {"XMSSMT-SHA2_", "_192", XMSS_SHA2, 24, 0x21},     {"XMSSMT-SHAKE256_", "_256", XMSS_SHAKE256, 32, 0x29},
    {"XMSSMT-SHAKE256_", "_192", XMSS_SHAKE256, 24, 0x31},
};
struct MtShape {
    const char* hd;
    unsigned h, d;
};
const MtShape kMtShapes[8] = {{"20/2", 20, 2}, {"20/4", 20, 4}, {"40/2", 40, 2}, {"40/4", 40, 4},
                              {"40/8", 40, 8}, {"60/3", 60, 3}, {"60/6", 60, 6}, {"60/12", 60, 12}};

int fill(xmss_params* p, unsigned func, unsigned n, unsigned h, unsigned d) {
    p->func = func;
    p->n = n;
    p->padding_len = (n == 24) ? 4 : n;
    p->full_height = h;
    p->d = d;
    p->wots_w = 16;
    p->bds_k = 0;  // params.c:372,695
    return xmss_xmssmt_initialize_params(p);
}

}  // namespace

extern "C" {

int xmss_str_to_oid(uint32_t* oid, const char* s) {
    for (const XmssSet& e : kXmss) {
        if (!strcmp(s, e.name)) {
            *oid = e.oid;
            return 0;
        }
    }
    return -1;
}

int xmssmt_str_to_oid(uint32_t* oid, const char* s) {
    for (const MtFamily& f : kMtFamilies) {
        for (unsigned k = 0; k < 8; ++k) {
            const std::string name = std::string(f.prefix) + kMtShapes[k].hd + f.suffix;
            if (name == s) {
                *oid = f.first_oid + k;
                return 0;
            }
        }
    }
    return -1;
}

int xmss_parse_oid(xmss_params* params, const uint32_t oid) {
    for (const XmssSet& e : kXmss) {
        if (e.oid == oid) return fill(params, e.func, e.n, e.h, 1);
    }
    return -1;
}

int xmssmt_parse_oid(xmss_params* params, const uint32_t oid) {
    for (const MtFamily& f : kMtFamilies) {
        if (oid >= f.first_oid && oid < f.first_oid + 8) {
            const MtShape& s = kMtShapes[oid - f.first_oid];
            return fill(params, f.func, f.n, s.h, s.d);
        }
    }
    return -1;
}

// cuxmss/params.c:710-753
int xmss_xmssmt_initialize_params(xmss_params* p) {
    p->tree_height = p->full_height / p->d;
    switch (p->wots_w) {
        case 4: p->wots_log_w = 2; p->wots_len2 = 5; break;
        case 16: p->wots_log_w = 4; p->wots_len2 = 3; break;
        case 256: p->wots_log_w = 8; p->wots_len2 = 2; break;
        default: return -1;
    }
    p->wots_len1 = 8 * p->n / p->wots_log_w;
    p->wots_len = p->wots_len1 + p->wots_len2;
    p->wots_sig_bytes = p->wots_len * p->n;
    p->index_bytes = (p->d == 1) ? 4 : (p->full_height + 7) / 8;
    p->sig_bytes = p->index_bytes + p->n + p->d * p->wots_sig_bytes + p->full_height * p->n;
    p->pk_bytes = 2 * p->n;
    p->sk_bytes = xmss_xmssmt_core_sk_bytes(p);
    return 0;
}

}  // extern "C"
