// VALU issue-rate probe for the roofline of the hash kernels.
//
// gfx950 does not issue every 32-bit VALU instruction at one wave64 per four
// cycles per SIMD: VOP2/VOP1 ops (v_add_u32, v_xor_b32, v_lshrrev_b32) and
// v_bitop3_b32 sustain up to 1.6-1.9x the rate of three-operand VOP3 ops
// (v_alignbit_b32, v_add3_u32, v_perm_b32, v_bfi_b32).  The attainable peak of
// a hash kernel therefore depends on its instruction mix, and this probe
// issues the mix of the SHA-256 WOTS chain step: per 12 instructions 5
// v_alignbit_b32, 2 v_add3_u32, 3 v_bitop3_b32, 1 v_add_u32, 1 v_lshrrev_b32
// (the chain step is 41.5% / 17.0% / 25.6% / 9.2% / 6.4% of these).  Eight
// independent accumulators per lane keep every wave issuing; the count is
// fixed by the inline assembly, so
//     lane-ops = blocks x 256 x iters x 96
// and lane-ops / kernel time is the rate the chip sustains for that mix.
// Kinds 1-5 issue one instruction class each (v_alignbit_b32; v_add3_u32;
// v_bitop3_b32; VOP2 v_add_u32 / v_xor_b32; v_perm_b32 / v_bfi_b32) to
// measure the per-class rates.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "engine.h"
#include "launch.h"

namespace xm {

#define XM_MIX12(x, y, z)                                   \
    asm volatile(                                           \
        "v_alignbit_b32 %0, %0, %0, 7\n\t"                  \
        "v_bitop3_b32 %0, %0, %1, %2 bitop3:0x96\n\t"       \
        "v_alignbit_b32 %0, %1, %0, 11\n\t"                 \
        "v_add3_u32 %0, %0, %1, %2\n\t"                     \
        "v_alignbit_b32 %0, %0, %0, 25\n\t"                 \
        "v_bitop3_b32 %0, %0, %1, %2 bitop3:0xe8\n\t"       \
        "v_lshrrev_b32 %0, 3, %0\n\t"                       \
        "v_alignbit_b32 %0, %2, %0, 13\n\t"                 \
        "v_add_u32 %0, %0, %1\n\t"                          \
        "v_bitop3_b32 %0, %0, %1, %2 bitop3:0xca\n\t"       \
        "v_alignbit_b32 %0, %0, %0, 22\n\t"                 \
        "v_add3_u32 %0, %0, %1, %2"                          \
        : "+v"(x)                                           \
        : "v"(y), "v"(z))

// One instruction class, three dependent instructions per accumulator.
#define XM_CLASS3(x, y, z, ins)                              \
    asm volatile(ins "\n\t" ins "\n\t" ins : "+v"(x) : "v"(y), "v"(z))

template <int KIND>
__device__ __forceinline__ void probe_step(uint32_t& x, uint32_t y, uint32_t z) {
    if constexpr (KIND == 0) {
        XM_MIX12(x, y, z);
    } else if constexpr (KIND == 1) {
        XM_CLASS3(x, y, z, "v_alignbit_b32 %0, %1, %0, 7");
        XM_CLASS3(x, y, z, "v_alignbit_b32 %0, %2, %0, 9");
        XM_CLASS3(x, y, z, "v_alignbit_b32 %0, %0, %0, 11");
        XM_CLASS3(x, y, z, "v_alignbit_b32 %0, %1, %0, 25");
    } else if constexpr (KIND == 2) {
        XM_CLASS3(x, y, z, "v_add3_u32 %0, %0, %1, %2");
        XM_CLASS3(x, y, z, "v_add3_u32 %0, %0, %2, %1");
        XM_CLASS3(x, y, z, "v_add3_u32 %0, %0, %1, %2");
        XM_CLASS3(x, y, z, "v_add3_u32 %0, %0, %2, %1");
    } else if constexpr (KIND == 3) {
        XM_CLASS3(x, y, z, "v_bitop3_b32 %0, %0, %1, %2 bitop3:0x96");
        XM_CLASS3(x, y, z, "v_bitop3_b32 %0, %0, %1, %2 bitop3:0xe8");
        XM_CLASS3(x, y, z, "v_bitop3_b32 %0, %0, %1, %2 bitop3:0xca");
        XM_CLASS3(x, y, z, "v_bitop3_b32 %0, %0, %1, %2 bitop3:0xd2");
    } else if constexpr (KIND == 4) {
        XM_CLASS3(x, y, z, "v_add_u32 %0, %0, %1");
        XM_CLASS3(x, y, z, "v_xor_b32 %0, %0, %2");
        XM_CLASS3(x, y, z, "v_add_u32 %0, %0, %2");
        XM_CLASS3(x, y, z, "v_xor_b32 %0, %0, %1");
    } else {
        XM_CLASS3(x, y, z, "v_perm_b32 %0, %0, %1, %2");
        XM_CLASS3(x, y, z, "v_bfi_b32 %0, %0, %1, %2");
        XM_CLASS3(x, y, z, "v_perm_b32 %0, %0, %2, %1");
        XM_CLASS3(x, y, z, "v_bfi_b32 %0, %0, %2, %1");
    }
}

template <int KIND>
__global__ __launch_bounds__(256) void k_valu_probe(uint32_t iters, uint32_t* __restrict__ out) {
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    uint32_t a0 = t, a1 = t + 1, a2 = t + 2, a3 = t + 3, a4 = t + 4, a5 = t + 5, a6 = t + 6, a7 = t + 7;
    const uint32_t y = t * 0x9e3779b9u, z = t ^ 0x5bd1e995u;
    for (uint32_t i = 0; i < iters; ++i) {
        probe_step<KIND>(a0, y, z);
        probe_step<KIND>(a1, y, z);
        probe_step<KIND>(a2, y, z);
        probe_step<KIND>(a3, y, z);
        probe_step<KIND>(a4, y, z);
        probe_step<KIND>(a5, y, z);
        probe_step<KIND>(a6, y, z);
        probe_step<KIND>(a7, y, z);
    }
    out[t] = a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7;
}

#undef XM_CLASS3
#undef XM_MIX12

// Runs probe `kind` once (after one warm-up launch) on its own stream of the
// current device; returns lane-ops/s and the kernel time in ms.  Every kind
// issues 96 instructions per lane and iteration.
int valu_probe(int kind, uint32_t blocks, uint32_t iters, double* lane_ops_per_s, double* ms_out) {
    using K = void (*)(uint32_t, uint32_t*);
    static const K kernels[VALU_PROBE_KINDS] = {k_valu_probe<0>, k_valu_probe<1>, k_valu_probe<2>,
                                                k_valu_probe<3>, k_valu_probe<4>, k_valu_probe<5>};
    if (kind < 0 || kind >= VALU_PROBE_KINDS || blocks == 0 || iters == 0) return -1;
    uint32_t* out = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCK(hipMalloc(&out, (size_t)blocks * 256 * sizeof(uint32_t)));
    HIPCK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    HIPCK(hipEventCreate(&e0));
    HIPCK(hipEventCreate(&e1));
    hipLaunchKernelGGL(kernels[kind], dim3(blocks), dim3(256), 0, s, iters / 8 + 1, out);
    HIPCK(hipGetLastError());
    HIPCK(hipEventRecord(e0, s));
    hipLaunchKernelGGL(kernels[kind], dim3(blocks), dim3(256), 0, s, iters, out);
    HIPCK(hipGetLastError());
    HIPCK(hipEventRecord(e1, s));
    HIPCK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCK(hipEventElapsedTime(&ms, e0, e1));
    HIPCK(hipEventDestroy(e0));
    HIPCK(hipEventDestroy(e1));
    HIPCK(hipStreamDestroy(s));
    HIPCK(hipFree(out));
    *ms_out = ms;
    *lane_ops_per_s = (double)blocks * 256.0 * iters * 96.0 / (ms / 1000.0);
    return 0;
}

}  // namespace xm
