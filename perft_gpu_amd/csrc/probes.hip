// probes.hip -- microbenchmarks that measure the ceilings the engine is
// reported against, on the box where it runs:
//   kind 0  int32 VALU rate: 8 independent add/xor chains per lane, all
//           waves resident; result in int32 lane-ops/s
//   kind 1  streaming copy of `bytes` (16-byte vector loads and stores);
//           result in bytes/s of read + write traffic (also the known-size
//           reference for rocprofv3 FETCH_SIZE / WRITE_SIZE scaling)
//   kind 2  random 16-byte read-modify-write over a `bytes`-sized table
//           (the transposition table's access pattern); result in
//           probes/s
#include <hip/hip_runtime.h>

#include <algorithm>

#include "engine.h"

namespace pf {

__global__ __launch_bounds__(256) void k_valu(u32 iters, u32 seed, u32* sink) {
    u32 a0 = seed ^ threadIdx.x, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3;
    u32 a4 = a0 + 4, a5 = a0 + 5, a6 = a0 + 6, a7 = a0 + 7;
    for (u32 i = 0; i < iters; i++) {
#pragma unroll
        for (int k = 0; k < 8; k++) {  // 8 ops, 4 independent at a time
            a0 += a1; a2 += a3; a4 += a5; a6 += a7;
            a1 ^= a2; a3 ^= a4; a5 ^= a6; a7 ^= a0;
        }
    }
    const u32 r = a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7;
    if (r == 0x12345678u) sink[0] = r;  // keeps the chains live
}

// Per-opcode issue rate: 8 independent chains of one instruction kind.
#define PF_OP8(OPC)                                                                      \
    asm volatile(OPC " %0, %0, %8\n\t" OPC " %1, %1, %8\n\t" OPC " %2, %2, %8\n\t"    \
                 OPC " %3, %3, %8\n\t" OPC " %4, %4, %8\n\t" OPC " %5, %5, %8\n\t"    \
                 OPC " %6, %6, %8\n\t" OPC " %7, %7, %8"                                \
                 : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7) \
                 : "v"(k))
// three-source forms: d = op(d, k, k2)
#define PF_OP8_3(OPC, MOD)                                                                           \
    asm volatile(OPC " %0, %0, %8, %9" MOD "\n\t" OPC " %1, %1, %8, %9" MOD "\n\t"                   \
                 OPC " %2, %2, %8, %9" MOD "\n\t" OPC " %3, %3, %8, %9" MOD "\n\t"                   \
                 OPC " %4, %4, %8, %9" MOD "\n\t" OPC " %5, %5, %8, %9" MOD "\n\t"                   \
                 OPC " %6, %6, %8, %9" MOD "\n\t" OPC " %7, %7, %8, %9" MOD                          \
                 : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7)       \
                 : "v"(k), "v"(k2))
// select with a scalar-register mask: d = m ? k : d
#define PF_OP8_S(OPC)                                                                                  \
    asm volatile(OPC " %0, %0, %8, %9\n\t" OPC " %1, %1, %8, %9\n\t" OPC " %2, %2, %8, %9\n\t"              \
                 OPC " %3, %3, %8, %9\n\t" OPC " %4, %4, %8, %9\n\t" OPC " %5, %5, %8, %9\n\t"              \
                 OPC " %6, %6, %8, %9\n\t" OPC " %7, %7, %8, %9"                                             \
                 : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7)       \
                 : "v"(k), "s"(sm))
template <int OP>
__global__ __launch_bounds__(256) void k_opc(u32 iters, u32* sink) {
    u32 a0 = threadIdx.x, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3, a4 = a0 + 4, a5 = a0 + 5, a6 = a0 + 6, a7 = a0 + 7;
    const u32 k = blockIdx.x | 3, k2 = blockIdx.x * 5u + 1u;
    const unsigned long long sm = 0x5555555555555555ULL ^ blockIdx.x;
    for (u32 i = 0; i < iters; i++) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            if constexpr (OP == 0) PF_OP8("v_add_u32");
            if constexpr (OP == 1) PF_OP8("v_xor_b32");
            if constexpr (OP == 2) PF_OP8("v_and_b32");
            if constexpr (OP == 3) PF_OP8("v_bcnt_u32_b32");
            if constexpr (OP == 4) PF_OP8("v_lshlrev_b32");
            if constexpr (OP == 5) PF_OP8("v_mul_lo_u32");
            if constexpr (OP == 7) PF_OP8_3("v_bitop3_b32", " bitop3:0xf8");
            if constexpr (OP == 8) PF_OP8_3("v_and_or_b32", "");
            if constexpr (OP == 9) PF_OP8_3("v_or3_b32", "");
            if constexpr (OP == 10) PF_OP8_S("v_cndmask_b32_e64");
            if constexpr (OP == 11) PF_OP8_3("v_perm_b32", "");
            if constexpr (OP == 12) PF_OP8_3("v_add3_u32", "");
            if constexpr (OP == 13) PF_OP8("v_min_u32");
            if constexpr (OP == 14) PF_OP8("v_ashrrev_i32");
            if constexpr (OP == 15) PF_OP8_3("v_bfe_i32", "");
        }
    }
    if ((a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7) == 0x12345678u) sink[0] = 1;
}
#undef PF_OP8
#undef PF_OP8_3
#undef PF_OP8_S

__global__ __launch_bounds__(256) void k_opc64(u32 iters, u32* sink) {
    u64 a0 = threadIdx.x, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3;
    const u32 k = (blockIdx.x & 7) | 1;
    for (u32 i = 0; i < iters; i++) {
#pragma unroll
        for (int j = 0; j < 16; j++) {
            asm volatile("v_lshlrev_b64 %0, %4, %0\n\tv_lshlrev_b64 %1, %4, %1\n\t"
                         "v_lshlrev_b64 %2, %4, %2\n\tv_lshlrev_b64 %3, %4, %3"
                         : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3)
                         : "v"(k));
        }
    }
    if ((a0 ^ a1 ^ a2 ^ a3) == 0x12345678u) sink[0] = 1;
}

__global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, u64 n) {
    for (u64 i = (u64)blockIdx.x * 256 + threadIdx.x; i < n; i += (u64)gridDim.x * 256) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void k_random16(uint4* table, u64 mask, u32 per_thread, u64 seed) {
    u64 x = seed ^ ((u64)(blockIdx.x * 256 + threadIdx.x) * 0x9E3779B97F4A7C15ULL);
    for (u32 i = 0; i < per_thread; i++) {
        x ^= x << 13; x ^= x >> 7; x ^= x << 17;
        uint4* e = table + (x & mask);
        uint4 v = *e;
        v.x += 1;
        *e = v;
    }
}

}  // namespace pf

pf_status pf::Engine::probe(int kind, u64 bytes, u64 iters, double* out) {
    if (!out) return PF_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return PF_EDEVICE;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return PF_EDEVICE;
    float ms = 0;
    pf_status st = PF_OK;
    if (kind == 0) {
        u32* sink = nullptr;
        if (hipMalloc(&sink, 4) != hipSuccess) return PF_ENOMEM;
        const u32 it = iters ? (u32)iters : 4096;
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_valu, 256, 0) != hipSuccess) (void)hipGetLastError();
        const int grid = std::max(per_cu, 1) * n_cu * 4;
        hipLaunchKernelGGL(k_valu, dim3(grid), dim3(256), 0, stream, 16u, 1u, sink);  // warm
        hipEventRecord(a, stream);
        hipLaunchKernelGGL(k_valu, dim3(grid), dim3(256), 0, stream, it, 7u, sink);
        hipEventRecord(b, stream);
        if (hipEventSynchronize(b) != hipSuccess) st = PF_EDEVICE;
        hipEventElapsedTime(&ms, a, b);
        // 8 unrolled steps x 8 ops per iteration
        *out = (double)grid * 256.0 * it * 64.0 / (ms * 1e-3);
        hipFree(sink);
    } else if (kind == 3) {
        // iters selects the opcode: 0 add, 1 xor, 2 and, 3 bcnt, 4 lshl32, 5 mul_lo, 6 lshl64,
        // 7 bitop3 (d | k & k2), 8 and_or, 9 or3, 10 cndmask (scalar mask), 11 perm, 12 add3, 13 min_u32,
        // 14 ashrrev_i32, 15 bfe_i32
        u32* sink = nullptr;
        if (hipMalloc(&sink, 4) != hipSuccess) return PF_ENOMEM;
        const u32 it = 2048;
        const void* f = nullptr;
        switch ((int)iters) {
        case 0: f = (const void*)k_opc<0>; break;
        case 1: f = (const void*)k_opc<1>; break;
        case 2: f = (const void*)k_opc<2>; break;
        case 3: f = (const void*)k_opc<3>; break;
        case 4: f = (const void*)k_opc<4>; break;
        case 5: f = (const void*)k_opc<5>; break;
        case 7: f = (const void*)k_opc<7>; break;
        case 8: f = (const void*)k_opc<8>; break;
        case 9: f = (const void*)k_opc<9>; break;
        case 10: f = (const void*)k_opc<10>; break;
        case 11: f = (const void*)k_opc<11>; break;
        case 12: f = (const void*)k_opc<12>; break;
        case 13: f = (const void*)k_opc<13>; break;
        case 14: f = (const void*)k_opc<14>; break;
        case 15: f = (const void*)k_opc<15>; break;
        default: f = (const void*)k_opc64; break;
        }
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, 256, 0) != hipSuccess) (void)hipGetLastError();
        const int grid = std::max(per_cu, 1) * n_cu * 4;
        u32 warm = 8, itv = it;
        void* a1[] = {(void*)&warm, (void*)&sink};
        void* a2[] = {(void*)&itv, (void*)&sink};
        hipLaunchKernel(f, dim3(grid), dim3(256), a1, 0, stream);
        hipEventRecord(a, stream);
        hipLaunchKernel(f, dim3(grid), dim3(256), a2, 0, stream);
        hipEventRecord(b, stream);
        if (hipEventSynchronize(b) != hipSuccess) st = PF_EDEVICE;
        hipEventElapsedTime(&ms, a, b);
        // wave64 instructions per second per SIMD-lane-equivalent: 64 per loop iteration
        *out = (double)grid * 256.0 * it * 64.0 / (ms * 1e-3);
        hipFree(sink);
    } else if (kind == 1 || kind == 2) {
        if (!bytes) bytes = 1ULL << 30;
        bytes &= ~15ULL;
        char* buf = nullptr;
        const u64 need = kind == 1 ? 2 * bytes : bytes;
        if (hipMalloc(&buf, need) != hipSuccess) return PF_ENOMEM;
        hipMemsetAsync(buf, 0, need, stream);
        if (kind == 1) {
            const u64 n = bytes / 16;
            const int grid = n_cu * 32;
            hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, stream, (const uint4*)buf, (uint4*)(buf + bytes), n);
            hipEventRecord(a, stream);
            const int reps = iters ? (int)iters : 5;
            for (int r = 0; r < reps; r++)
                hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, stream, (const uint4*)buf, (uint4*)(buf + bytes), n);
            hipEventRecord(b, stream);
            if (hipEventSynchronize(b) != hipSuccess) st = PF_EDEVICE;
            hipEventElapsedTime(&ms, a, b);
            *out = 2.0 * (double)bytes * reps / (ms * 1e-3);
        } else {
            u64 entries = 1;
            while (entries * 2 * 16 <= bytes) entries *= 2;
            const u32 per = iters ? (u32)iters : 64;
            const int grid = n_cu * 32;
            hipEventRecord(a, stream);
            hipLaunchKernelGGL(k_random16, dim3(grid), dim3(256), 0, stream, (uint4*)buf, entries - 1, per, 12345ULL);
            hipEventRecord(b, stream);
            if (hipEventSynchronize(b) != hipSuccess) st = PF_EDEVICE;
            hipEventElapsedTime(&ms, a, b);
            *out = (double)grid * 256.0 * per / (ms * 1e-3);
        }
        hipFree(buf);
    } else {
        st = PF_EINVAL;
    }
    hipEventDestroy(a);
    hipEventDestroy(b);
    return st;
}
