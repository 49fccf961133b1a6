// spmm.hip -- multi-band two-sided Jacobi radiosity sweep as a CSR SpMM.
//
// One sweep of the reference's shader/facetrt/radiosity.comp:63-106 for all
// bands at once.  Per side s (partner side o = s ^ 1) and band b:
//   G_s[b] = sum_e w_e * B[col_e][b]                   (CSR gather, FP32 acc)
//   c_s[b] = e_s + r_s (fsky_s S_b + D_s) + t_s (fsky_o S_b + D_o)  (per solve)
//   B'_s[b] = B_s[b] + omega * (c_s[b] + r_s G_s[b] + t_s G_o[b] - B_s[b])
// with w_e = count_e / den_s and fsky_s = sky_s / den_s (1 when den_s = 0).
//
// Layout in HBM: B, c, r, t are [side][ldb] with all bands of a side
// contiguous (ldb = nb rounded up to 4), so a CSR entry's gather reads one
// contiguous run of bands; the two sides of a facet are adjacent rows.
// A team of TEAM lanes owns one facet: lanes stride over the row's entries,
// each accumulating NBC bands in registers, then the team reduces with
// cross-lane shuffles and lane j finishes band j.  Bands beyond NBC are
// processed in chunks by the same team (the row is re-read from cache).
//
// Modes:
//   SWEEP    FP32 (W16=false) or BF16 (W16=true) weights; residual reduced.
//   REFRESH  defect correction: G32 and G16 in one pass; writes the FP32
//            Jacobi update (the judge sweep) and c2 = c + r(G32-G16)_s +
//            t(G32-G16)_o, the corrected source the BF16 sweeps then use.
//   INCIDENT H_s[b] = max(0, fsky_s S_b + G_s[b] + D_s[b]).
//
// Convergence is decided on the device: every block max-reduces |dB| and
// |dB|/|B| into the control block; the last block to finish advances the
// iteration count and raises `done`, after which queued sweeps exit at once.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <string>

#include "solver.hpp"
#include "sweep_common.hpp"

namespace fh {

enum { MODE_SWEEP = 0, MODE_REFRESH = 1, MODE_INCIDENT = 2 };

struct SweepArgs {
    const uint32_t* rowptr;
    const uint32_t* col;
    const float* w32;
    const uint16_t* w16;
    const float* Bin;
    float* Bout;
    const float* c;   // source term (c2 during BF16 sweeps)
    const float* r;
    const float* t;
    const float* cBase;  // REFRESH: uncorrected source
    float* c2;           // REFRESH: corrected source out
    const float* fsky;
    const float* D;
    const float* skyv;   // [ldb] sky radiosity per band (device)
    float* H;
    uint32_t F;
    uint32_t ldb;
    uint32_t nb;
    uint32_t nchunks;
    float relax;
    uint32_t judge;  // 1: this sweep may end the solve on tolRel
    uint32_t standalone;  // 1: timing launch, ignores the control block
    float* buf0;  // when set: input / output are buf[iter & 1] / buf[~iter & 1],
    float* buf1;  // iter = sweeps executed so far in this solve (device side)
    SolveCtrl* ctrl;
    float2* partial;  // per-block (max |dB|, max |dB|/|B|)
    uint32_t xcd;     // 1: XCD-contiguous block order (xcd_block)
};

template <int NBC, int TEAM, int MODE, bool W16>
__global__ __launch_bounds__(256) void k_sweep(SweepArgs a) {
    SolveCtrl* ctrl = a.ctrl;
    if (MODE != MODE_INCIDENT && !a.standalone) {
        if (*reinterpret_cast<volatile uint32_t*>(&ctrl->done)) return;
        if (MODE == MODE_SWEEP && W16 && *reinterpret_cast<volatile uint32_t*>(&ctrl->innerDone)) return;
    }
    const float* __restrict__ Bin = a.Bin;
    float* __restrict__ Bout = a.Bout;
    if (MODE != MODE_INCIDENT && !a.standalone && a.buf0) {
        const uint32_t it = *reinterpret_cast<volatile uint32_t*>(&ctrl->iter);
        Bin = (it & 1u) ? a.buf1 : a.buf0;
        Bout = (it & 1u) ? a.buf0 : a.buf1;
    }
    const int lane = threadIdx.x % TEAM;
    const uint32_t f = (xcd_block(a.xcd) * 256u + threadIdx.x) / TEAM;
    float myAbs = 0.0f, myRel = 0.0f;
    if (f < a.F) {
        const uint32_t s0 = 2u * f;
        const uint32_t e0 = a.rowptr[s0], e1 = a.rowptr[s0 + 1], e2 = a.rowptr[s0 + 2];
        for (uint32_t ch = 0; ch < a.nchunks; ++ch) {
            const uint32_t b0 = ch * NBC;
            float g0[NBC], g1[NBC];
            float q0[NBC], q1[NBC];  // BF16 sums in REFRESH mode
#pragma unroll
            for (int j = 0; j < NBC; ++j) g0[j] = g1[j] = q0[j] = q1[j] = 0.0f;
            for (uint32_t e = e0 + lane; e < e1; e += TEAM) {
                const uint32_t n = a.col[e];
                const float* p = Bin + size_t(n) * a.ldb + b0;
                if (MODE == MODE_REFRESH) {
                    gather<NBC>(p, a.w32[e], g0);
                    gather<NBC>(p, bf16_to_f32(a.w16[e]), q0);
                } else {
                    gather<NBC>(p, W16 ? bf16_to_f32(a.w16[e]) : a.w32[e], g0);
                }
            }
            for (uint32_t e = e1 + lane; e < e2; e += TEAM) {
                const uint32_t n = a.col[e];
                const float* p = Bin + size_t(n) * a.ldb + b0;
                if (MODE == MODE_REFRESH) {
                    gather<NBC>(p, a.w32[e], g1);
                    gather<NBC>(p, bf16_to_f32(a.w16[e]), q1);
                } else {
                    gather<NBC>(p, W16 ? bf16_to_f32(a.w16[e]) : a.w32[e], g1);
                }
            }
            team_reduce<NBC, TEAM>(g0);
            team_reduce<NBC, TEAM>(g1);
            if (MODE == MODE_REFRESH) {
                team_reduce<NBC, TEAM>(q0);
                team_reduce<NBC, TEAM>(q1);
            }
#pragma unroll
            for (int j = 0; j < NBC; ++j) {
                if ((j % TEAM) != lane) continue;
                const uint32_t b = b0 + j;
                if (b >= a.ldb) continue;
                const size_t i0 = size_t(s0) * a.ldb + b, i1 = i0 + a.ldb;
                if (MODE == MODE_INCIDENT) {
                    const float sk = a.skyv[b];
                    const float h0 = a.fsky[s0] * sk + g0[j] + a.D[i0];
                    const float h1 = a.fsky[s0 + 1] * sk + g1[j] + a.D[i1];
                    a.H[i0] = fmaxf(h0, 0.0f);
                    a.H[i1] = fmaxf(h1, 0.0f);
                    continue;
                }
                const float r0 = a.r[i0], t0 = a.t[i0], r1 = a.r[i1], t1 = a.t[i1];
                const float* csrc = MODE == MODE_REFRESH ? a.cBase : a.c;
                const float c0 = csrc[i0], c1 = csrc[i1];
                const float n0 = c0 + r0 * g0[j] + t0 * g1[j];
                const float n1 = c1 + r1 * g1[j] + t1 * g0[j];
                if (MODE == MODE_REFRESH) {
                    a.c2[i0] = c0 + r0 * (g0[j] - q0[j]) + t0 * (g1[j] - q1[j]);
                    a.c2[i1] = c1 + r1 * (g1[j] - q1[j]) + t1 * (g0[j] - q0[j]);
                }
                const float o0 = Bin[i0], o1 = Bin[i1];
                const float x0 = o0 + a.relax * (n0 - o0);
                const float x1 = o1 + a.relax * (n1 - o1);
                Bout[i0] = x0;
                Bout[i1] = x1;
                const float d0 = fabsf(x0 - o0), d1 = fabsf(x1 - o1);
                myAbs = fmaxf(myAbs, fmaxf(d0, d1));
                myRel = fmaxf(myRel, fmaxf(d0 / fmaxf(fabsf(x0), 1.0e-30f), d1 / fmaxf(fabsf(x1), 1.0e-30f)));
            }
        }
    }
    if (MODE == MODE_INCIDENT) return;
    // block max-reduction of the residuals into this block's partial slot;
    // k_finalize reduces the partials (no atomics or fences on the hot path)
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) {
        myAbs = fmaxf(myAbs, __shfl_xor(myAbs, m, 64));
        myRel = fmaxf(myRel, __shfl_xor(myRel, m, 64));
    }
    __shared__ float sAbs[4], sRel[4];
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sAbs[wave] = myAbs;
        sRel[wave] = myRel;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const float bA = fmaxf(fmaxf(sAbs[0], sAbs[1]), fmaxf(sAbs[2], sAbs[3]));
        const float bR = fmaxf(fmaxf(sRel[0], sRel[1]), fmaxf(sRel[2], sRel[3]));
        a.partial[blockIdx.x] = make_float2(bA, bR);
    }
}

// Band-parallel variant for multi-band B (ldb >= 4): a team of BT lanes owns
// one facet and each lane owns KB bands (b = chunk + lane + k BT), so the
// gather of one CSR entry reads BT * KB contiguous floats of B[col] with
// coalesced loads.  The team loads BT entries (col, w) at a time, one per
// lane, and broadcasts each by cross-lane shuffle.  Bands beyond BT * KB are
// processed in chunks.
template <int BT, int KB, int MODE, bool W16>
__global__ __launch_bounds__(256) void k_sweep_bp(SweepArgs a) {
    SolveCtrl* ctrl = a.ctrl;
    if (MODE != MODE_INCIDENT && !a.standalone) {
        if (*reinterpret_cast<volatile uint32_t*>(&ctrl->done)) return;
        if (MODE == MODE_SWEEP && W16 && *reinterpret_cast<volatile uint32_t*>(&ctrl->innerDone)) return;
    }
    const float* __restrict__ Bin = a.Bin;
    float* __restrict__ Bout = a.Bout;
    if (MODE != MODE_INCIDENT && !a.standalone && a.buf0) {
        const uint32_t it = *reinterpret_cast<volatile uint32_t*>(&ctrl->iter);
        Bin = (it & 1u) ? a.buf1 : a.buf0;
        Bout = (it & 1u) ? a.buf0 : a.buf1;
    }
    const int lane = threadIdx.x % BT;
    const uint32_t f = (xcd_block(a.xcd) * 256u + threadIdx.x) / BT;
    const uint32_t ldb = a.ldb;
    float myAbs = 0.0f, myRel = 0.0f;
    if (f < a.F) {
        const uint32_t s0 = 2u * f;
        const uint32_t rb[3] = {a.rowptr[s0], a.rowptr[s0 + 1], a.rowptr[s0 + 2]};
        for (uint32_t cb = 0; cb < ldb; cb += BT * KB) {
            float g[2][KB], q[2][KB];
#pragma unroll
            for (int k = 0; k < KB; ++k) g[0][k] = g[1][k] = q[0][k] = q[1][k] = 0.0f;
            const float* __restrict__ Bl = Bin + cb + lane;
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const uint32_t end = rb[side + 1];
                // U entries per step: all column / weight loads, then all
                // U * KB gathers, then the FMAs (entries past the row end use
                // column 0 with weight 0)
                constexpr int U = 8;
                for (uint32_t base = rb[side]; base < end; base += U) {
                    uint32_t n[U];
                    float w[U], wq[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const uint32_t e = base + uint32_t(u);
                        const bool ok = e < end;
                        n[u] = ok ? a.col[e] : 0u;
                        if (MODE == MODE_REFRESH) {
                            w[u] = ok ? a.w32[e] : 0.0f;
                            wq[u] = ok ? bf16_to_f32(a.w16[e]) : 0.0f;
                        } else {
                            w[u] = ok ? (W16 ? bf16_to_f32(a.w16[e]) : a.w32[e]) : 0.0f;
                        }
                    }
                    float v[U][KB];
#pragma unroll
                    for (int u = 0; u < U; ++u)
#pragma unroll
                        for (int k = 0; k < KB; ++k)
                            v[u][k] = cb + uint32_t(lane + k * BT) < ldb ? Bl[size_t(n[u]) * ldb + k * BT] : 0.0f;
#pragma unroll
                    for (int u = 0; u < U; ++u)
#pragma unroll
                        for (int k = 0; k < KB; ++k) {
                            g[side][k] += w[u] * v[u][k];
                            if (MODE == MODE_REFRESH) q[side][k] += wq[u] * v[u][k];
                        }
                }
            }
#pragma unroll
            for (int k = 0; k < KB; ++k) {
                const uint32_t b = cb + uint32_t(lane) + uint32_t(k * BT);
                if (b >= ldb) continue;
                const size_t i0 = size_t(s0) * ldb + b, i1 = i0 + ldb;
                if (MODE == MODE_INCIDENT) {
                    const float sk = a.skyv[b];
                    a.H[i0] = fmaxf(a.fsky[s0] * sk + g[0][k] + a.D[i0], 0.0f);
                    a.H[i1] = fmaxf(a.fsky[s0 + 1] * sk + g[1][k] + a.D[i1], 0.0f);
                    continue;
                }
                const float r0 = a.r[i0], t0 = a.t[i0], r1 = a.r[i1], t1 = a.t[i1];
                const float* csrc = MODE == MODE_REFRESH ? a.cBase : a.c;
                const float c0 = csrc[i0], c1 = csrc[i1];
                const float n0 = c0 + r0 * g[0][k] + t0 * g[1][k];
                const float n1 = c1 + r1 * g[1][k] + t1 * g[0][k];
                if (MODE == MODE_REFRESH) {
                    a.c2[i0] = c0 + r0 * (g[0][k] - q[0][k]) + t0 * (g[1][k] - q[1][k]);
                    a.c2[i1] = c1 + r1 * (g[1][k] - q[1][k]) + t1 * (g[0][k] - q[0][k]);
                }
                const float o0 = Bin[i0], o1 = Bin[i1];
                const float x0 = o0 + a.relax * (n0 - o0);
                const float x1 = o1 + a.relax * (n1 - o1);
                Bout[i0] = x0;
                Bout[i1] = x1;
                const float d0 = fabsf(x0 - o0), d1 = fabsf(x1 - o1);
                myAbs = fmaxf(myAbs, fmaxf(d0, d1));
                myRel = fmaxf(myRel, fmaxf(d0 / fmaxf(fabsf(x0), 1.0e-30f), d1 / fmaxf(fabsf(x1), 1.0e-30f)));
            }
        }
    }
    if (MODE == MODE_INCIDENT) return;
    float v[2] = {myAbs, myRel}, out[2];
    block_max<2>(v, out);
    if (threadIdx.x == 0) a.partial[blockIdx.x] = make_float2(out[0], out[1]);
}

// Reduces the per-block residuals of the sweep just run and advances the
// control block: iteration count, last residuals, done flag.
__global__ __launch_bounds__(1024) void k_finalize(SolveCtrl* ctrl, const float2* __restrict__ partial, uint32_t n,
                                                   uint32_t judge, uint32_t refresh, uint32_t inner) {
    if (*reinterpret_cast<volatile uint32_t*>(&ctrl->done)) return;
    if (inner && *reinterpret_cast<volatile uint32_t*>(&ctrl->innerDone)) return;
    float mA = 0.0f, mR = 0.0f;
    for (uint32_t i = threadIdx.x; i < n; i += 1024u) {
        const float2 p = partial[i];
        mA = fmaxf(mA, p.x);
        mR = fmaxf(mR, p.y);
    }
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) {
        mA = fmaxf(mA, __shfl_xor(mA, m, 64));
        mR = fmaxf(mR, __shfl_xor(mR, m, 64));
    }
    __shared__ float sA[16], sR[16];
    if ((threadIdx.x & 63) == 0) {
        sA[threadIdx.x >> 6] = mA;
        sR[threadIdx.x >> 6] = mR;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float absd = 0.0f, reld = 0.0f;
        for (int w = 0; w < 16; ++w) {
            absd = fmaxf(absd, sA[w]);
            reld = fmaxf(reld, sR[w]);
        }
        const uint32_t it = ctrl->iter + 1u;
        ctrl->iter = it;
        ctrl->lastAbs = absd;
        ctrl->lastRel = reld;
        if (refresh) {
            ctrl->refreshes += 1u;
            ctrl->innerDone = 0u;
        }
        if (inner && ctrl->tolRel >= 0.0f && reld <= ctrl->tolRel) ctrl->innerDone = 1u;
        bool stop = it >= ctrl->maxIter;
        if (judge && ctrl->tolRel >= 0.0f && reld <= ctrl->tolRel) stop = true;
        if (ctrl->tolAbs >= 0.0f && (it % ctrl->checkEvery) == 0u && absd <= ctrl->tolAbs) stop = true;
        if (stop) ctrl->done = 1u;
    }
}

// c = e + r (fsky_s S + D_s) + t (fsky_o S + D_o): folds the per-solve
// constant terms of radiosity.comp:85-95 into one source array.
__global__ void k_fold_source(const float* __restrict__ e, const float* __restrict__ r, const float* __restrict__ t,
                              const float* __restrict__ D, const float* __restrict__ fsky,
                              const float* __restrict__ skyv, uint32_t S, uint32_t ldb, float* __restrict__ c) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(S) * ldb) return;
    const uint32_t s = uint32_t(i / ldb), b = uint32_t(i - size_t(s) * ldb);
    const size_t io = size_t(s ^ 1u) * ldb + b;
    const float sk = skyv[b];
    const float hs = fsky[s] * sk + D[i];
    const float ho = fsky[s ^ 1u] * sk + D[io];
    c[i] = e[i] + r[i] * hs + t[i] * ho;
}

// [S][nb] -> [S][ldb] (zero padding); and back
__global__ void k_pad(const float* __restrict__ src, uint32_t S, uint32_t nb, uint32_t ldb, float* __restrict__ dst) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(S) * ldb) return;
    const uint32_t s = uint32_t(i / ldb), b = uint32_t(i - size_t(s) * ldb);
    dst[i] = b < nb ? src[size_t(s) * nb + b] : 0.0f;
}

__global__ void k_unpad(const float* __restrict__ src, uint32_t S, uint32_t nb, uint32_t ldb, float* __restrict__ dst) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(S) * nb) return;
    const uint32_t s = uint32_t(i / nb), b = uint32_t(i - size_t(s) * nb);
    dst[i] = src[size_t(s) * ldb + b];
}

// AoS SurfaceOptics -> r, t, e, D (nb = 1)
__global__ void k_split_optics(const float4* __restrict__ o, uint32_t S, float* __restrict__ r, float* __restrict__ t,
                               float* __restrict__ e, float* __restrict__ D) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const float4 v = o[s];
    r[s] = v.x;
    t[s] = v.y;
    e[s] = v.z;
    D[s] = v.w;
}

int sweep_team(const fh_context* c);

uint32_t band_stride(uint32_t nb) { return nb <= 1u ? 1u : (nb + 3u) & ~3u; }

static int pick_nbc(uint32_t ldb) {
    if (ldb == 1u) return 1;
    if (ldb == 4u) return 4;
    if (ldb == 8u) return 8;
    if (ldb % 16u == 0u) return 16;
    if (ldb % 8u == 0u) return 8;
    return 4;
}

int sweep_team(const fh_context* c) {
    // FH_SPMM_TEAM=4|16|64 overrides the choice (tuning knob)
    static const int forced = [] {
        const char* e = std::getenv("FH_SPMM_TEAM");
        return e ? std::atoi(e) : 0;
    }();
    if (forced == 4 || forced == 8 || forced == 16 || forced == 32 || forced == 64) return forced;
    const double avg = c->S ? double(c->nnz) / double(c->S) : 0.0;
    if (avg < 12.0) return 4;
    if (avg < 24.0) return 8;
    if (avg < 96.0) return 16;
    return 32;
}

template <int NBC, int TEAM>
static void launch_mode(int mode, bool w16, const SweepArgs& a, uint32_t grid, hipStream_t st) {
    if (mode == MODE_SWEEP && !w16)
        hipLaunchKernelGGL((k_sweep<NBC, TEAM, MODE_SWEEP, false>), dim3(grid), dim3(256), 0, st, a);
    else if (mode == MODE_SWEEP)
        hipLaunchKernelGGL((k_sweep<NBC, TEAM, MODE_SWEEP, true>), dim3(grid), dim3(256), 0, st, a);
    else if (mode == MODE_REFRESH)
        hipLaunchKernelGGL((k_sweep<NBC, TEAM, MODE_REFRESH, false>), dim3(grid), dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL((k_sweep<NBC, TEAM, MODE_INCIDENT, false>), dim3(grid), dim3(256), 0, st, a);
}

template <int NBC>
static void launch_team(int team, int mode, bool w16, const SweepArgs& a, hipStream_t st) {
    const uint64_t threads = uint64_t(a.F) * team;
    const uint32_t grid = uint32_t((threads + 255u) / 256u);
    const uint32_t g = grid ? grid : 1u;
    if (team == 4) launch_mode<NBC, 4>(mode, w16, a, g, st);
    else if (team == 8) launch_mode<NBC, 8>(mode, w16, a, g, st);
    else if (team == 16) launch_mode<NBC, 16>(mode, w16, a, g, st);
    else if (team == 32) launch_mode<NBC, 32>(mode, w16, a, g, st);
    else launch_mode<NBC, 64>(mode, w16, a, g, st);
}

template <int BT, int KB>
static void launch_bp_mode(int mode, bool w16, const SweepArgs& a, uint32_t grid, hipStream_t st) {
    if (mode == MODE_SWEEP && !w16)
        hipLaunchKernelGGL((k_sweep_bp<BT, KB, MODE_SWEEP, false>), dim3(grid), dim3(256), 0, st, a);
    else if (mode == MODE_SWEEP)
        hipLaunchKernelGGL((k_sweep_bp<BT, KB, MODE_SWEEP, true>), dim3(grid), dim3(256), 0, st, a);
    else if (mode == MODE_REFRESH)
        hipLaunchKernelGGL((k_sweep_bp<BT, KB, MODE_REFRESH, false>), dim3(grid), dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL((k_sweep_bp<BT, KB, MODE_INCIDENT, false>), dim3(grid), dim3(256), 0, st, a);
}

// lanes per facet (BT) and bands per lane (KB) of the band-parallel kernel
static void pick_bp(uint32_t ldb, int& bt, int& kb) {
    if (ldb % 16u == 0u) {
        bt = 16;
        kb = ldb >= 64u ? 4 : int(ldb / 16u);
    } else if (ldb % 8u == 0u) {
        bt = 8;
        kb = 1;
    } else {
        bt = 4;
        kb = ldb == 12u ? 3 : 1;
    }
}

// FH_SPMM_KERNEL=entry|band forces one of the two kernels (tuning knob)
static bool band_parallel(uint32_t ldb) {
    static const int forced = [] {
        const char* e = std::getenv("FH_SPMM_KERNEL");
        return e ? (std::string(e) == "entry" ? 1 : (std::string(e) == "band" ? 2 : 0)) : 0;
    }();
    if (forced == 1) return false;
    if (forced == 2) return ldb >= 4u;
    // measured on gfx950: entry-parallel wins up to 8 bands, band-parallel
    // from 16 (fewer cache-line requests per gather instruction)
    return ldb >= 16u;
}

// FH_XCD_REMAP=0|1 forces the block order (tuning knob)
static uint32_t xcd_remap() {
    static const uint32_t v = [] {
        const char* e = std::getenv("FH_XCD_REMAP");
        return e ? uint32_t(std::atoi(e) != 0) : 0u;
    }();
    return v;
}

void launch_sweep(fh_context* c, int mode, bool w16, SweepArgs a) {
    a.xcd = xcd_remap();
    if (band_parallel(a.ldb)) {
        int bt = 16, kb = 1;
        pick_bp(a.ldb, bt, kb);
        a.F = c->F;
        a.rowptr = c->rowptr.p;
        a.col = c->col.p;
        a.w32 = c->val32.p;
        a.w16 = c->val16.p;
        a.fsky = c->fsky.p;
        a.ctrl = c->ctrl.p;
        const uint32_t grid = uint32_t((uint64_t(c->F) * uint64_t(bt) + 255u) / 256u);
        c->partial.ensure(grid ? grid : 1u);
        a.partial = c->partial.p;
        const uint32_t g = grid ? grid : 1u;
        if (bt == 16 && kb == 1) launch_bp_mode<16, 1>(mode, w16, a, g, c->stream);
        else if (bt == 16 && kb == 2) launch_bp_mode<16, 2>(mode, w16, a, g, c->stream);
        else if (bt == 16 && kb == 3) launch_bp_mode<16, 3>(mode, w16, a, g, c->stream);
        else if (bt == 16) launch_bp_mode<16, 4>(mode, w16, a, g, c->stream);
        else if (bt == 8) launch_bp_mode<8, 1>(mode, w16, a, g, c->stream);
        else if (kb == 3) launch_bp_mode<4, 3>(mode, w16, a, g, c->stream);
        else launch_bp_mode<4, 1>(mode, w16, a, g, c->stream);
        if (mode != MODE_INCIDENT && !a.standalone)
            hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, c->stream, c->ctrl.p, c->partial.p, g, a.judge,
                               uint32_t(mode == MODE_REFRESH), uint32_t(mode == MODE_SWEEP && w16));
        FH_HIP(hipGetLastError());
        return;
    }
    const int nbc = pick_nbc(a.ldb);
    const int team = sweep_team(c);
    a.nchunks = (a.ldb + uint32_t(nbc) - 1u) / uint32_t(nbc);
    a.F = c->F;
    a.rowptr = c->rowptr.p;
    a.col = c->col.p;
    a.w32 = c->val32.p;
    a.w16 = c->val16.p;
    a.fsky = c->fsky.p;
    a.ctrl = c->ctrl.p;
    const uint64_t threads = uint64_t(c->F) * uint64_t(team);
    const uint32_t grid = uint32_t((threads + 255u) / 256u);
    c->partial.ensure(grid ? grid : 1u);
    a.partial = c->partial.p;
    switch (nbc) {
        case 1: launch_team<1>(team, mode, w16, a, c->stream); break;
        case 4: launch_team<4>(team, mode, w16, a, c->stream); break;
        case 8: launch_team<8>(team, mode, w16, a, c->stream); break;
        default: launch_team<16>(team, mode, w16, a, c->stream); break;
    }
    if (mode != MODE_INCIDENT && !a.standalone)
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, c->stream, c->ctrl.p, c->partial.p, grid ? grid : 1u,
                           a.judge, uint32_t(mode == MODE_REFRESH), uint32_t(mode == MODE_SWEEP && w16));
    FH_HIP(hipGetLastError());
}

void launch_fold(fh_context* c, const float* e, const float* r, const float* t, const float* D, uint32_t ldb,
                 float* out) {
    const size_t n = size_t(c->S) * ldb;
    hipLaunchKernelGGL(k_fold_source, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, c->stream, e, r, t, D, c->fsky.p,
                       c->skyv.p, c->S, ldb, out);
    FH_HIP(hipGetLastError());
}

void launch_pad(fh_context* c, const float* src, uint32_t nb, uint32_t ldb, float* dst) {
    const size_t n = size_t(c->S) * ldb;
    hipLaunchKernelGGL(k_pad, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, c->stream, src, c->S, nb, ldb, dst);
    FH_HIP(hipGetLastError());
}

void launch_unpad(fh_context* c, const float* src, uint32_t nb, uint32_t ldb, float* dst) {
    const size_t n = size_t(c->S) * nb;
    hipLaunchKernelGGL(k_unpad, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, c->stream, src, c->S, nb, ldb, dst);
    FH_HIP(hipGetLastError());
}

void launch_split_optics(fh_context* c, const float4* o, float* r, float* t, float* e, float* D) {
    hipLaunchKernelGGL(k_split_optics, dim3((c->S + 255u) / 256u), dim3(256), 0, c->stream, o, c->S, r, t, e, D);
    FH_HIP(hipGetLastError());
}


__global__ void k_ctrl_reset(SolveCtrl* ctrl, uint32_t maxIter, float tolRel, float tolAbs, uint32_t checkEvery) {
    if (threadIdx.x == 0) {
        SolveCtrl z{};
        z.maxIter = maxIter;
        z.tolRel = tolRel;
        z.tolAbs = tolAbs;
        z.checkEvery = checkEvery ? checkEvery : 1u;
        *ctrl = z;
    }
}

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_f32(float* p, size_t n, float v) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void launch_ctrl_reset(fh_context* c, uint32_t maxIter, float tolRel, float tolAbs, uint32_t checkEvery) {
    hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(64), 0, c->stream, c->ctrl.p, maxIter, tolRel, tolAbs, checkEvery);
    FH_HIP(hipGetLastError());
}

void launch_fill_u32(fh_context* c, uint32_t* p, size_t n, uint32_t v) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_u32, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, c->stream, p, n, v);
    FH_HIP(hipGetLastError());
}

void launch_fill_f32(fh_context* c, float* p, size_t n, float v) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_f32, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, c->stream, p, n, v);
    FH_HIP(hipGetLastError());
}

static void read_ctrl(fh_context* c) {
    FH_HIP(hipMemcpyAsync(c->hctrl, c->ctrl.p, sizeof(SolveCtrl), hipMemcpyDeviceToHost, c->stream));
    FH_HIP(hipStreamSynchronize(c->stream));
}

void set_sky(fh_context* c, const float* sky, uint32_t nb, uint32_t ldb) {
    std::vector<float> pad(ldb, 0.0f);
    for (uint32_t b = 0; b < nb; ++b) pad[b] = sky[b];
    c->skyv.ensure(ldb);
    FH_HIP(hipMemcpyAsync(c->skyv.p, pad.data(), sizeof(float) * ldb, hipMemcpyHostToDevice, c->stream));
    FH_HIP(hipStreamSynchronize(c->stream));
}

int run_jacobi(fh_context* c, const JacobiOptions& o, const float* cc, const float* rr, const float* tt,
               float* bufs[2], fh_solve_stats* stats) {
    launch_ctrl_reset(c, o.maxIter ? o.maxIter : 1u, o.tolRel, o.tolAbs, o.checkEvery);
    SweepArgs a{};
    a.r = rr;
    a.t = tt;
    a.ldb = o.ldb;
    a.relax = o.relax;
    uint64_t launched = 0;
    uint32_t batch = 4;
    const bool bf16 = o.precision == FH_PRECISION_BF16_DC;
    if (bf16) c->c2.ensure(size_t(c->S) * o.ldb);
    while (true) {
        a.buf0 = bufs[0];
        a.buf1 = bufs[1];
        for (uint32_t k = 0; k < batch; ++k) {
            if (!bf16) {
                a.c = cc;
                a.judge = 1u;
                launch_sweep(c, MODE_SWEEP, false, a);
            } else {
                // groups of BF16 sweeps (stopping early once their relative
                // change is within tolRel) closed by an FP32 refresh that
                // judges convergence and recomputes the defect correction;
                // the first group runs on the uncorrected source
                const uint32_t period = o.innerBf16 + 1u;
                if (launched % period == period - 1u) {
                    a.c = cc;
                    a.cBase = cc;
                    a.c2 = c->c2.p;
                    a.judge = 1u;
                    launch_sweep(c, MODE_REFRESH, false, a);
                } else {
                    a.c = launched < period ? cc : c->c2.p;
                    a.judge = 0u;
                    launch_sweep(c, MODE_SWEEP, true, a);
                }
            }
            ++launched;
        }
        read_ctrl(c);
        if (c->hctrl->done) break;
        if (batch < 32u) batch *= 2u;
    }
    if (stats) {
        stats->iterations = c->hctrl->iter;
        stats->maxDelta = c->hctrl->lastAbs;
        stats->maxRelDelta = c->hctrl->lastRel;
        stats->refreshes = c->hctrl->refreshes;
    }
    return int(c->hctrl->iter & 1u);
}

void run_incident(fh_context* c, uint32_t ldb, const float* B, const float* D, float* H) {
    SweepArgs a{};
    a.Bin = B;
    a.D = D;
    a.H = H;
    a.skyv = c->skyv.p;
    a.ldb = ldb;
    launch_sweep(c, MODE_INCIDENT, false, a);
}

void launch_single_sweep(fh_context* c, uint32_t ldb, const float* cc, const float* rr, const float* tt,
                         const float* Bin, float* Bout, bool w16) {
    SweepArgs a{};
    a.c = cc;
    a.r = rr;
    a.t = tt;
    a.Bin = Bin;
    a.Bout = Bout;
    a.ldb = ldb;
    a.relax = 1.0f;
    a.judge = 0u;
    a.standalone = 1u;
    launch_sweep(c, MODE_SWEEP, w16, a);
}
}  // namespace fh
