// sweep_common.hpp -- device helpers shared by the SpMM sweep kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fh {

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  With remap set, hardware block b runs logical block xcd_block(), so
// each XCD works through one contiguous range of rows and the columns that
// neighbouring rows gather stay in that XCD's L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t remap) {
    const uint32_t b = blockIdx.x;
    if (!remap) return b;
    const uint32_t G = gridDim.x, x = b & 7u, l = b >> 3, q = G >> 3, rem = G & 7u;
    return x < rem ? x * (q + 1u) + l : rem * (q + 1u) + (x - rem) * q + l;
}

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

template <int NBC>
__device__ __forceinline__ void gather(const float* __restrict__ p, float w, float (&acc)[NBC]) {
    if constexpr (NBC == 1) {
        acc[0] += w * p[0];
    } else {
#pragma unroll
        for (int q = 0; q < NBC / 4; ++q) {
            const float4 v = *reinterpret_cast<const float4*>(p + 4 * q);
            acc[4 * q + 0] += w * v.x;
            acc[4 * q + 1] += w * v.y;
            acc[4 * q + 2] += w * v.z;
            acc[4 * q + 3] += w * v.w;
        }
    }
}

template <int NBC, int TEAM>
__device__ __forceinline__ void team_reduce(float (&acc)[NBC]) {
#pragma unroll
    for (int j = 0; j < NBC; ++j) {
#pragma unroll
        for (int m = TEAM / 2; m > 0; m >>= 1) acc[j] += __shfl_xor(acc[j], m, TEAM);
    }
}


// max-reduce n floats of a 256-thread block into lane 0 of wave 0
template <int N>
__device__ __forceinline__ void block_max(float (&v)[N], float (&out)[N]) {
#pragma unroll
    for (int q = 0; q < N; ++q)
#pragma unroll
        for (int m = 32; m > 0; m >>= 1) v[q] = fmaxf(v[q], __shfl_xor(v[q], m, 64));
    __shared__ float sh[4][N];
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < N; ++q) sh[threadIdx.x >> 6][q] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < N; ++q) out[q] = fmaxf(fmaxf(sh[0][q], sh[1][q]), fmaxf(sh[2][q], sh[3][q]));
}

}  // namespace fh
