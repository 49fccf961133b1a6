// sweep_common.hpp -- device helpers shared by the SpMM sweep kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fh {

// A control-block word (iteration count, active bands, done flags) as a
// wave-uniform value.  The word is written only by kernels launched before
// the reading one, so every lane reads the same value; readfirstlane moves
// it into a scalar register, and the pointers selected by it (and the
// buffer descriptors built from them) stay scalar.  Without it a buffer
// load through a descriptor in vector registers becomes a loop over the
// distinct descriptors of the wave (readfirstlane, compare, exec mask) at
// every gather.
__device__ __forceinline__ uint32_t ctrl_word(const uint32_t* p) {
    return __builtin_amdgcn_readfirstlane(*reinterpret_cast<const volatile uint32_t*>(p));
}

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

// round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f32_to_bf16(float x) {
    const uint32_t u = __float_as_uint(x);
    return uint16_t((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// Packed CSR entry: column in the low `bits` bits, ray count above.  The
// weight count / den of the row equals the FP32 weight array's correctly
// rounded quotient: with rcp = RN(1 / den), q = RN(count rcp) and the exact
// residual r = count - q den (one FMA), RN(q + r rcp) is the correctly
// rounded quotient (Markstein's correction step; checked exhaustively for
// 0 <= count <= den <= 65535, the range packed graphs are built for).
// den 0 passes rcp 0 and gives weight 0.
struct RowDen {
    float den, rcp;
};
__device__ __forceinline__ RowDen row_den(uint32_t d) {
    const float f = float(d);
    return {f, d ? __frcp_rn(f) : 0.0f};
}
__device__ __forceinline__ uint32_t pk_col(uint32_t p, uint32_t bits) { return p & ((1u << bits) - 1u); }
__device__ __forceinline__ float pk_weight(uint32_t p, uint32_t bits, RowDen d) {
    const float c = float(p >> bits);
    const float q = __fmul_rn(c, d.rcp);
    const float r = fmaf(-q, d.den, c);
    return fmaf(r, d.rcp, q);
}

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

// Single-band gathers of the two rows [b0, e0) and [b1, e1) by the lanes of
// a team: lane handles entries lane + k TEAM of both rows, U of each row per
// step with all column / weight loads issued, then all 2 U gathers, then the
// FMAs, so that a row of ~U TEAM entries costs one round of memory latency.
// Per-lane partial sums; the caller reduces them over the team.
template <int TEAM, int U, bool W16>
__device__ __forceinline__ void team_rows_dot2(const uint32_t* __restrict__ col, const float* __restrict__ w32,
                                               const uint16_t* __restrict__ w16, const float* __restrict__ x,
                                               uint32_t b0, uint32_t e0, uint32_t b1, uint32_t e1, int lane,
                                               float& g0, float& g1) {
    const uint32_t len = max(e0 - b0, e1 - b1);
    for (uint32_t k = uint32_t(lane); k < len; k += uint32_t(TEAM * U)) {
        uint32_t n0[U], n1[U];
        float w0[U], w1[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i0 = b0 + k + uint32_t(u * TEAM), i1 = b1 + k + uint32_t(u * TEAM);
            const bool ok0 = i0 < e0, ok1 = i1 < e1;
            n0[u] = ok0 ? col[i0] : 0u;
            n1[u] = ok1 ? col[i1] : 0u;
            w0[u] = ok0 ? (W16 ? bf16_to_f32(w16[i0]) : w32[i0]) : 0.0f;
            w1[u] = ok1 ? (W16 ? bf16_to_f32(w16[i1]) : w32[i1]) : 0.0f;
        }
        float v0[U], v1[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            v0[u] = x[n0[u]];
            v1[u] = x[n1[u]];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            g0 += w0[u] * v0[u];
            g1 += w1[u] * v1[u];
        }
    }
}

// team_rows_dot2 over packed entries (pk_col / pk_weight with the rows'
// denominators d0, d1)
template <int TEAM, int U>
__device__ __forceinline__ void team_rows_dot2_pk(const uint32_t* __restrict__ pk, uint32_t bits, RowDen d0, RowDen d1,
                                                  const float* __restrict__ x, uint32_t b0, uint32_t e0, uint32_t b1,
                                                  uint32_t e1, int lane, float& g0, float& g1) {
    const uint32_t len = max(e0 - b0, e1 - b1);
    for (uint32_t k = uint32_t(lane); k < len; k += uint32_t(TEAM * U)) {
        uint32_t p0[U], p1[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i0 = b0 + k + uint32_t(u * TEAM), i1 = b1 + k + uint32_t(u * TEAM);
            p0[u] = i0 < e0 ? pk[i0] : 0u;
            p1[u] = i1 < e1 ? pk[i1] : 0u;
        }
        float v0[U], v1[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            v0[u] = x[pk_col(p0[u], bits)];
            v1[u] = x[pk_col(p1[u], bits)];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            g0 += pk_weight(p0[u], bits, d0) * v0[u];
            g1 += pk_weight(p1[u], bits, d1) * v1[u];
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

// max-reduce n floats of a 1024-thread block; every thread gets the result
template <int N>
__device__ __forceinline__ void block_max_1024(float (&v)[N], float (&out)[N]) {
#pragma unroll
    for (int q = 0; q < N; ++q)
#pragma unroll
        for (int m = 32; m > 0; m >>= 1) v[q] = fmaxf(v[q], __shfl_xor(v[q], m, 64));
    __shared__ float sh[16][N];
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < N; ++q) sh[threadIdx.x >> 6][q] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < N; ++q) {
        float r = sh[0][q];
        for (int w = 1; w < 16; ++w) r = fmaxf(r, sh[w][q]);
        out[q] = r;
    }
}

}  // namespace fh
