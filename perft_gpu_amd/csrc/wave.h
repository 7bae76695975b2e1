// wave.h -- wave64 scans and sums shared by the level kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pf {

// Inclusive prefix sum over the 64 lanes of a wave, in DPP moves: four row
// shifts (1, 2, 4, 8 lanes, a lane with no source adding 0) give each 16-lane
// row its own scan, then two row broadcasts carry the totals forward (rows 1
// and 3 add lane 15 of the row before, rows 2 and 3 add lane 31).  Every lane
// of the wave must be active.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return v;
}

// Sum over the 64 lanes, every lane getting it.
__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long v) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

}  // namespace pf
