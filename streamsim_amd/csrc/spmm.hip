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

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

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
    // band retirement (band-parallel kernels): bands whose relative change
    // in a judge sweep is within tolRel stop iterating; ctrl->nAct bounds
    // the bands still iterated, partAct[block] = 1 + highest band above tol
    uint32_t retire;
    uint32_t* partAct;
    // BF16 copies of B (row stride ldh, a multiple of 64 bands = 128 bytes):
    // BF16 sweeps gather Bin16; sweeps of a BF16 solve also write Bout16.
    // Ping-pong between h0 / h1 by ctrl->iter like buf0 / buf1.
    uint16_t* h0;
    uint16_t* h1;
    uint32_t ldh;
    uint32_t bt;      // band-parallel kernels: lanes per facet team
    uint32_t fuseH;   // judge sweeps also write H = max(0, fsky S + G + D) (band-parallel)
    const uint32_t* sched;  // facet processed by logical team i (null: facet i)
    const float* omega;  // Chebyshev weights per band for this sweep (band-parallel FP32), or null
    const float* rho2;   // squared spectral radius estimate per band (Chebyshev)
    // increment sweeps (band-parallel BF16 solves): 1 = base sweep, 2 =
    // increment sweep, 0 = neither; gacc = accumulated gathers M Q, carry =
    // B - Q per side-band ([S][ldb] FP32), Q the state the gathers have seen
    uint32_t inc;
    float* gacc;
    float* carry;
    // band-parallel sweeps of energy-balance shortwave solves: r, t from the
    // facet's material tables and c, D from per-side direct factors instead
    // of the [S][ldb] arrays (null fmat: the arrays)
    const uint32_t* fmat;  // material per facet
    const float* rtab;     // [materials][ldb] reflectance per internal band
    const float* ttab;     // [materials][ldb] transmittance
    const float* dirf;     // [S] clamp(sunlit, 0, 1) max(0, cos(normal, sun))
    const float* dniv;     // [ldb] direct normal irradiance per internal band
    // packed CSR entries (col | cnt << pkBits) read instead of col / w32 by
    // the entry-parallel FP32-weight gathers, weight cnt / den[row]; null:
    // col / w32.  (The band-parallel kernels keep col / w32: every lane of a
    // team would recompute the same weights.)
    const uint32_t* pk;
    const uint32_t* den;
    uint32_t pkBits;
    // pair blocks (band-parallel FP32 and increment sweeps): team i updates
    // the two facets of units[i] together, with the CSR entries between them
    // solved exactly (unitW); null: one facet per team
    const uint2* units;
    const float4* unitW;
    uint32_t nUnits;
    uint32_t jacobiOnly;  // 1: plain Jacobi sweep (single-sweep entry points)
    uint32_t sides;       // sides of the scene (rows of the gathered arrays)
    // per-block partials given by the caller (capacity partialCap); null:
    // the context's partial buffer, grown as needed
    float2* partialExt;
    uint32_t partialCap;
    // FP8 increment sweeps (k_sweep_q8_single / _pair): the gathered increments as rows of
    // ldq bytes (16-byte chunks of 12 bands and their exponent), ping-pong
    // between q0 / q1 by ctrl->iter like buf0 / buf1
    uint8_t* q0;
    uint8_t* q1;
    uint32_t ldq;
    uint32_t q8;  // 1: increments in FP8 instead of BF16
    uint32_t partOff;     // band-parallel: first partial of this launch (a sweep may be two launches)
    // 1 (increment sweep, BF16): the first sweep of a solve continuing the
    // increment state of the solve before: no gathers (the accumulated
    // gathers hold M (B - carry) already), the increments that solve wrote
    // last (Bin16, never gathered) added to the carries
    uint32_t fold;
};

constexpr uint32_t kNoFacet = 0xffffffffu;  // no partner facet

template <int NBC, int TEAM, int MODE, bool W16>
__global__ __launch_bounds__(256) void k_sweep(SweepArgs a) {
    SolveCtrl* ctrl = a.ctrl;
    if (MODE != MODE_INCIDENT && !a.standalone) {
        if (ctrl_word(&ctrl->done)) return;
        if (MODE == MODE_SWEEP && W16 && ctrl_word(&ctrl->innerDone)) return;
    }
    const float* __restrict__ Bin = a.Bin;
    float* __restrict__ Bout = a.Bout;
    if (MODE != MODE_INCIDENT && !a.standalone && a.buf0) {
        const uint32_t it = ctrl_word(&ctrl->iter);
        Bin = (it & 1u) ? a.buf1 : a.buf0;
        Bout = (it & 1u) ? a.buf0 : a.buf1;
    }
    const int lane = threadIdx.x % TEAM;
    const uint32_t fl = (xcd_block(a.xcd) * 256u + threadIdx.x) / TEAM;
    float myAbs = 0.0f, myRel = 0.0f;
    if (fl < a.F) {
        const uint32_t f = a.sched ? a.sched[fl] : fl;
        const uint32_t s0 = 2u * f;
        const uint32_t e0 = a.rowptr[s0], e1 = a.rowptr[s0 + 1], e2 = a.rowptr[s0 + 2];
        for (uint32_t ch = 0; ch < a.nchunks; ++ch) {
            const uint32_t b0 = ch * NBC;
            float g0[NBC], g1[NBC];
            float q0[NBC], q1[NBC];  // BF16 sums in REFRESH mode
#pragma unroll
            for (int j = 0; j < NBC; ++j) g0[j] = g1[j] = q0[j] = q1[j] = 0.0f;
            if constexpr (NBC == 1) {
                // single band: both rows with U entries per lane in flight
                constexpr int U = TEAM >= 32 ? 4 : 8;
                if (MODE == MODE_REFRESH) {
                    team_rows_dot2<TEAM, U, false>(a.col, a.w32, a.w16, Bin, e0, e1, e1, e2, lane, g0[0], g1[0]);
                    team_rows_dot2<TEAM, U, true>(a.col, a.w32, a.w16, Bin, e0, e1, e1, e2, lane, q0[0], q1[0]);
                } else if (!W16 && a.pk) {
                    team_rows_dot2_pk<TEAM, U>(a.pk, a.pkBits, row_den(a.den[s0]), row_den(a.den[s0 + 1]), Bin, e0,
                                               e1, e1, e2, lane, g0[0], g1[0]);
                } else {
                    team_rows_dot2<TEAM, U, W16>(a.col, a.w32, a.w16, Bin, e0, e1, e1, e2, lane, g0[0], g1[0]);
                }
            } else {
            const bool packed = MODE != MODE_REFRESH && !W16 && a.pk;
            const RowDen d0 = row_den(packed ? a.den[s0] : 0u), d1 = row_den(packed ? a.den[s0 + 1] : 0u);
            for (uint32_t e = e0 + lane; e < e1; e += TEAM) {
                const uint32_t q = packed ? a.pk[e] : 0u;
                const uint32_t n = packed ? pk_col(q, a.pkBits) : a.col[e];
                const float* p = Bin + size_t(n) * a.ldb + b0;
                if (MODE == MODE_REFRESH) {
                    gather<NBC>(p, a.w32[e], g0);
                    gather<NBC>(p, bf16_to_f32(a.w16[e]), q0);
                } else {
                    gather<NBC>(p, packed ? pk_weight(q, a.pkBits, d0) : W16 ? bf16_to_f32(a.w16[e]) : a.w32[e], g0);
                }
            }
            for (uint32_t e = e1 + lane; e < e2; e += TEAM) {
                const uint32_t q = packed ? a.pk[e] : 0u;
                const uint32_t n = packed ? pk_col(q, a.pkBits) : a.col[e];
                const float* p = Bin + size_t(n) * a.ldb + b0;
                if (MODE == MODE_REFRESH) {
                    gather<NBC>(p, a.w32[e], g1);
                    gather<NBC>(p, bf16_to_f32(a.w16[e]), q1);
                } else {
                    gather<NBC>(p, packed ? pk_weight(q, a.pkBits, d1) : W16 ? bf16_to_f32(a.w16[e]) : a.w32[e], g1);
                }
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
                if (MODE == MODE_SWEEP && a.fuseH && a.judge) {  // incident of this sweep's gather
                    const float sk = a.skyv[b];
                    const float h0 = a.fsky[s0] * sk + g0[j] + a.D[i0];
                    const float h1 = a.fsky[s0 + 1] * sk + g1[j] + a.D[i1];
                    a.H[i0] = fmaxf(h0, 0.0f);
                    a.H[i1] = fmaxf(h1, 0.0f);
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
        a.partial[a.partOff + blockIdx.x] = make_float2(bA, bR);
    }
}

// Band-parallel sweeps (ldb >= 16).  A team of bt lanes owns one facet and
// every lane owns one 16-byte vector of bands: 4 FP32 bands (k_sweep_v) or
// 8 BF16 bands of the BF16 copy (k_sweep_hv), so the gather of one CSR entry
// is a single load instruction per team covering the row's contiguous bands
// (2 lines of 128 bytes for 48 FP32 bands, 1 line for up to 64 BF16 bands).
// Every lane of a team loads the same U entries (col, w) (one address per
// team), then issues the U gathers before the FMAs.  Teams do not straddle
// lanes' work: a block holds 256 / bt teams; bands beyond one team pass are
// processed in chunks.
//
// With band retirement only bands [0, nAct) are gathered and updated; bands
// [nAct, nActPrev), retired by the previous judge sweep, are copied from Bin
// to Bout once so that both ping-pong buffers hold their final value.
constexpr int kU = 8;  // CSR entries in flight per lane (one aligned group)
#ifndef FH_PAIR_GROUPS
#define FH_PAIR_GROUPS 2
#endif
constexpr int kPairGroups = FH_PAIR_GROUPS;  // aligned groups in flight per lane in the pair gathers

// The 8-aligned group of CSR entries [base, base + 8): columns and weights
// with 16-byte vector loads (one request per team and vector; the CSR arrays
// are padded by kCsrPad zero entries), entries outside [start, end) get
// weight 0 and ok[u] = false.
template <bool W16>
__device__ __forceinline__ void load_group(const SweepArgs& a, uint32_t base, uint32_t start, uint32_t end,
                                           uint32_t (&n)[kU], float (&w)[kU], bool (&ok)[kU]) {
    const uint4 cA = *reinterpret_cast<const uint4*>(a.col + base);
    const uint4 cB = *reinterpret_cast<const uint4*>(a.col + base + 4);
    n[0] = cA.x; n[1] = cA.y; n[2] = cA.z; n[3] = cA.w;
    n[4] = cB.x; n[5] = cB.y; n[6] = cB.z; n[7] = cB.w;
    float h[kU];
    if (W16) {
        const uint4 q = *reinterpret_cast<const uint4*>(a.w16 + base);
        const uint32_t qq[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            h[2 * j] = __uint_as_float(qq[j] << 16);
            h[2 * j + 1] = __uint_as_float(qq[j] & 0xffff0000u);
        }
    }
    float f[kU];
    if (!W16) {
        const float4 wA = *reinterpret_cast<const float4*>(a.w32 + base);
        const float4 wB = *reinterpret_cast<const float4*>(a.w32 + base + 4);
        f[0] = wA.x; f[1] = wA.y; f[2] = wA.z; f[3] = wA.w;
        f[4] = wB.x; f[5] = wB.y; f[6] = wB.z; f[7] = wB.w;
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
        const uint32_t e = base + uint32_t(u);
        ok[u] = e >= start && e < end;
        w[u] = ok[u] ? (W16 ? h[u] : f[u]) : 0.0f;
    }
}

struct BandCtx {
    uint32_t nAct, nPrev;
    float tolRel;
    const float* Bin;
    float* Bout;
    const uint16_t* Bin16;
    uint16_t* Bout16;
    const uint8_t* Bin8;
    uint8_t* Bout8;
};

// control-block state of a band-parallel sweep; false: the sweep is skipped
template <int MODE>
__device__ __forceinline__ bool band_ctx(const SweepArgs& a, bool inner, BandCtx& x) {
    SolveCtrl* ctrl = a.ctrl;
    x.nAct = x.nPrev = a.ldb;
    x.tolRel = -1.0f;
    x.Bin = a.Bin;
    x.Bout = a.Bout;
    x.Bin16 = a.h0;  // standalone: h0 -> h1
    x.Bout16 = a.h1;
    x.Bin8 = a.q0;
    x.Bout8 = a.q1;
    if (MODE == MODE_INCIDENT || a.standalone) return true;
    if (ctrl_word(&ctrl->done)) return false;
    if (inner && ctrl_word(&ctrl->innerDone)) return false;
    if (a.retire) {
        x.nAct = ctrl_word(&ctrl->nAct);
        x.nPrev = ctrl_word(&ctrl->nActPrev);
        x.tolRel = ctrl->tolRel;
    }
    if (a.buf0) {
        const uint32_t it = ctrl_word(&ctrl->iter);
        x.Bin = (it & 1u) ? a.buf1 : a.buf0;
        x.Bout = (it & 1u) ? a.buf0 : a.buf1;
        x.Bin16 = a.h0 ? ((it & 1u) ? a.h1 : a.h0) : nullptr;
        x.Bout16 = a.h0 ? ((it & 1u) ? a.h0 : a.h1) : nullptr;
        x.Bin8 = a.q0 ? ((it & 1u) ? a.q1 : a.q0) : nullptr;
        x.Bout8 = a.q0 ? ((it & 1u) ? a.q0 : a.q1) : nullptr;
    }
    return true;
}

__device__ __forceinline__ float4 ld4(const float* p) { return *reinterpret_cast<const float4*>(p); }
__device__ __forceinline__ void st4(float* p, const float (&v)[4]) {
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ void arr4(const float4 v, float (&o)[4]) {
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
    o[3] = v.w;
}

// r, t, c and D of the side pair (s0, s0 + 1) at the bands b0 .. b0 + 3:
// from the [S][ldb] arrays, or (fmat set) from the facet's material row and
// the direct factors, D = dir dni and c = r (fsky_s S + D_s) + t (fsky_o S + D_o)
// (both sides of a facet share its material).  D only when wantD.
__device__ __forceinline__ void load_rtc(const SweepArgs& a, uint32_t s0, uint32_t b0, size_t i0, size_t i1,
                                         bool wantD, float (&r0)[4], float (&r1)[4], float (&t0)[4], float (&t1)[4],
                                         float (&c0)[4], float (&c1)[4], float (&D0)[4], float (&D1)[4]) {
    if (a.fmat) {
        const size_t m = size_t(a.fmat[s0 >> 1]) * a.ldb + b0;
        float sk[4], dn[4];
        arr4(ld4(a.rtab + m), r0);
        arr4(ld4(a.ttab + m), t0);
        arr4(ld4(a.skyv + b0), sk);
        arr4(ld4(a.dniv + b0), dn);
        const float f0 = a.fsky[s0], f1 = a.fsky[s0 + 1], d0 = a.dirf[s0], d1 = a.dirf[s0 + 1];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            r1[j] = r0[j];
            t1[j] = t0[j];
            D0[j] = __fmul_rn(d0, dn[j]);
            D1[j] = __fmul_rn(d1, dn[j]);
            const float h0 = fmaf(f0, sk[j], D0[j]), h1 = fmaf(f1, sk[j], D1[j]);
            c0[j] = fmaf(t0[j], h1, __fmul_rn(r0[j], h0));
            c1[j] = fmaf(t0[j], h0, __fmul_rn(r0[j], h1));
        }
        return;
    }
    arr4(ld4(a.r + i0), r0);
    arr4(ld4(a.r + i1), r1);
    arr4(ld4(a.t + i0), t0);
    arr4(ld4(a.t + i1), t1);
    arr4(ld4(a.c + i0), c0);
    arr4(ld4(a.c + i1), c1);
    if (wantD) {
        arr4(ld4(a.D + i0), D0);
        arr4(ld4(a.D + i1), D1);
    }
}

// Per-facet state of a 4-band row update (bands b0 .. b0 + 3, b0 a multiple
// of 4, 16-byte row accesses) between the gathers and the stores.
struct Upd4 {
    size_t i0, i1;        // [S][ldb] offsets of the sides s0, s0 + 1
    float o0[4], o1[4];   // B of this sweep
    float n0[4], n1[4];   // Jacobi value c + r g_s + t g_o from the gathers
    float r0[4], r1[4], t0[4], t1[4];
    float k0[4], k1[4];   // increment sweeps: carry B - Q; else 0
    bool act[4];          // band below nAct
};

// First half of the update of the side pair (s0, s0 + 1) from the gathered
// sums g.  Bands in [nAct, nActPrev), retired by the last judge sweep, are
// copied once so both ping-pong buffers hold their final value, and bands
// from nActPrev on are left alone (false for both); otherwise the Jacobi
// value n and, on judge sweeps of fused solves, the incident of this
// sweep's gather.  INC: 0 FP32 sweep, 1 / 2 base / increment sweep.
//
// Increment sweeps: the gathers see Q, a copy of B that only ever changes
// by BF16 increments: sweep k gathers the increments dq_k (BF16) and
// accumulates G = M Q_k in FP32 (gacc); each side keeps carry = B - Q
// exactly.  The update from G is the FP32 one (Jacobi or Chebyshev), and the
// next increment is dq_k+1 = bf16(B_k+1 - Q_k) with carry_k+1 = (B_k+1 - Q_k)
// - dq_k+1, so what rounding drops in one sweep is sent in the next and
// Q -> B as B converges: the fixed point is the FP32 one.  The base sweep
// gathers Q_0 = bf16(B_0) itself.
template <int INC>
__device__ __forceinline__ bool upd_begin(const SweepArgs& a, const BandCtx& x, uint32_t s0, uint32_t b0,
                                          float (&g0)[4], float (&g1)[4], Upd4& u) {
    if (b0 >= x.nPrev) return false;
    const uint32_t ldb = a.ldb;
    u.i0 = size_t(s0) * ldb + b0;
    u.i1 = u.i0 + ldb;
    const size_t i0 = u.i0, i1 = u.i1;
    arr4(ld4(x.Bin + i0), u.o0);
    arr4(ld4(x.Bin + i1), u.o1);
    if (b0 >= x.nAct) {
        st4(x.Bout + i0, u.o0);
        st4(x.Bout + i1, u.o1);
        if (INC && x.Bout16) {  // the band's last increment, not gathered, follows into both buffers
            *reinterpret_cast<uint2*>(x.Bout16 + size_t(s0) * a.ldh + b0) =
                *reinterpret_cast<const uint2*>(x.Bin16 + size_t(s0) * a.ldh + b0);
            *reinterpret_cast<uint2*>(x.Bout16 + size_t(s0 + 1) * a.ldh + b0) =
                *reinterpret_cast<const uint2*>(x.Bin16 + size_t(s0 + 1) * a.ldh + b0);
        }
        return false;
    }
    if (INC == 1) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            u.k0[j] = u.o0[j] - bf16_to_f32(f32_to_bf16(u.o0[j]));
            u.k1[j] = u.o1[j] - bf16_to_f32(f32_to_bf16(u.o1[j]));
        }
    } else if (INC == 2) {
        float p0[4], p1[4];
        arr4(ld4(a.gacc + i0), p0);
        arr4(ld4(a.gacc + i1), p1);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            g0[j] += p0[j];
            g1[j] += p1[j];
        }
        arr4(ld4(a.carry + i0), u.k0);
        arr4(ld4(a.carry + i1), u.k1);
        if (a.fold) {
            const uint2 p0 = *reinterpret_cast<const uint2*>(x.Bin16 + size_t(s0) * a.ldh + b0);
            const uint2 p1 = *reinterpret_cast<const uint2*>(x.Bin16 + size_t(s0 + 1) * a.ldh + b0);
            const uint32_t w0[2] = {p0.x, p0.y}, w1[2] = {p1.x, p1.y};
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                u.k0[2 * j] += __uint_as_float(w0[j] << 16);
                u.k0[2 * j + 1] += __uint_as_float(w0[j] & 0xffff0000u);
                u.k1[2 * j] += __uint_as_float(w1[j] << 16);
                u.k1[2 * j + 1] += __uint_as_float(w1[j] & 0xffff0000u);
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) u.k0[j] = u.k1[j] = 0.0f;
    }
    if (INC) {
        st4(a.gacc + i0, g0);
        st4(a.gacc + i1, g1);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) u.act[j] = b0 + uint32_t(j) < x.nAct;
    const bool fuse = a.fuseH && a.judge;
    float c0[4], c1[4], D0[4], D1[4];
    load_rtc(a, s0, b0, i0, i1, fuse, u.r0, u.r1, u.t0, u.t1, c0, c1, D0, D1);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        u.n0[j] = c0[j] + u.r0[j] * g0[j] + u.t0[j] * g1[j];
        u.n1[j] = c1[j] + u.r1[j] * g1[j] + u.t1[j] * g0[j];
    }
    if (fuse) {
        float sk[4], H0[4], H1[4];
        arr4(ld4(a.skyv + b0), sk);
        const float f0 = a.fsky[s0], f1 = a.fsky[s0 + 1];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            H0[j] = fmaxf(f0 * sk[j] + g0[j] + D0[j], 0.0f);
            H1[j] = fmaxf(f1 * sk[j] + g1[j] + D1[j], 0.0f);
        }
        if (u.act[3]) {
            st4(a.H + i0, H0);
            st4(a.H + i1, H1);
        } else {  // retired bands keep the incident of their last judge sweep
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (u.act[j]) {
                    a.H[i0 + j] = H0[j];
                    a.H[i1 + j] = H1[j];
                }
        }
    }
    return true;
}

// Second half: the new iterate from m, the sweep's value before
// acceleration (m = n with one facet per team) — Chebyshev y = p +
// omega_b (m - p) with p the iterate before B, which Bout still holds
// (ping-pong), or y = B + relax (m - B); the residual maxima of the Jacobi
// change |n - B| (relaxed: of relax (n - B)), the reference's criterion;
// the highest band above tolRel; the stores (and the next increment).
// KEEP_V (increment sweeps): the next increment before rounding, v = (B'
// - B) + carry, is left in u.k0 / u.k1 for the caller to round and store
// (FP8 increments round a chunk of bands spread over three lanes together).
template <int INC, bool KEEP_V = false>
__device__ __forceinline__ void upd_end(const SweepArgs& a, const BandCtx& x, uint32_t s0, uint32_t b0, Upd4& u,
                                        const float (&m0)[4], const float (&m1)[4], float& myAbs, float& myRel,
                                        float& myAct) {
    const size_t i0 = u.i0, i1 = u.i1;
    float y0[4], y1[4], d0[4], d1[4], e0[4], e1[4];
    if (INC || a.omega) {
        float w[4] = {1.0f, 1.0f, 1.0f, 1.0f};
        if (a.omega) arr4(ld4(a.omega + b0), w);
        const bool anyW = w[0] != 1.0f || w[1] != 1.0f || w[2] != 1.0f || w[3] != 1.0f;
        float p0[4] = {0.0f, 0.0f, 0.0f, 0.0f}, p1[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (anyW) {
            arr4(ld4(x.Bout + i0), p0);
            arr4(ld4(x.Bout + i1), p1);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            y0[j] = w[j] == 1.0f ? m0[j] : p0[j] + w[j] * (m0[j] - p0[j]);
            y1[j] = w[j] == 1.0f ? m1[j] : p1[j] + w[j] * (m1[j] - p1[j]);
            d0[j] = fabsf(u.n0[j] - u.o0[j]);
            d1[j] = fabsf(u.n1[j] - u.o1[j]);
            e0[j] = d0[j] / fmaxf(fabsf(u.n0[j]), 1.0e-30f);
            e1[j] = d1[j] / fmaxf(fabsf(u.n1[j]), 1.0e-30f);
        }
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float j0 = u.o0[j] + a.relax * (u.n0[j] - u.o0[j]);  // relaxed Jacobi iterate
            const float j1 = u.o1[j] + a.relax * (u.n1[j] - u.o1[j]);
            y0[j] = u.o0[j] + a.relax * (m0[j] - u.o0[j]);
            y1[j] = u.o1[j] + a.relax * (m1[j] - u.o1[j]);
            d0[j] = fabsf(j0 - u.o0[j]);
            d1[j] = fabsf(j1 - u.o1[j]);
            e0[j] = d0[j] / fmaxf(fabsf(j0), 1.0e-30f);
            e1[j] = d1[j] / fmaxf(fabsf(j1), 1.0e-30f);
        }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (!u.act[j]) {  // beyond nAct inside this vector: unchanged
            y0[j] = u.o0[j];
            y1[j] = u.o1[j];
            continue;
        }
        myAbs = fmaxf(myAbs, fmaxf(d0[j], d1[j]));
        myRel = fmaxf(myRel, fmaxf(e0[j], e1[j]));
        if (fmaxf(e0[j], e1[j]) > x.tolRel) myAct = fmaxf(myAct, float(b0 + uint32_t(j) + 1u));
    }
    st4(x.Bout + i0, y0);
    st4(x.Bout + i1, y1);
    if (INC && KEEP_V) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            u.k0[j] = (y0[j] - u.o0[j]) + u.k0[j];
            u.k1[j] = (y1[j] - u.o1[j]) + u.k1[j];
        }
    } else if (INC) {
        // next increment and the part of it that BF16 does not carry
        uint32_t q[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float v0 = (y0[j] - u.o0[j]) + u.k0[j], v1 = (y1[j] - u.o1[j]) + u.k1[j];
            const uint16_t h0 = f32_to_bf16(v0), h1 = f32_to_bf16(v1);
            u.k0[j] = v0 - bf16_to_f32(h0);
            u.k1[j] = v1 - bf16_to_f32(h1);
            q[j] = uint32_t(h0) | (uint32_t(h1) << 16);
        }
        st4(a.carry + i0, u.k0);
        st4(a.carry + i1, u.k1);
        *reinterpret_cast<uint2*>(x.Bout16 + size_t(s0) * a.ldh + b0) =
            make_uint2((q[0] & 0xffffu) | (q[1] << 16), (q[2] & 0xffffu) | (q[3] << 16));
        *reinterpret_cast<uint2*>(x.Bout16 + size_t(s0 + 1) * a.ldh + b0) =
            make_uint2((q[0] >> 16) | (q[1] & 0xffff0000u), (q[2] >> 16) | (q[3] & 0xffff0000u));
    } else if (x.Bout16) {
        const uint32_t h0 = uint32_t(f32_to_bf16(y0[0])) | (uint32_t(f32_to_bf16(y0[1])) << 16);
        const uint32_t h1 = uint32_t(f32_to_bf16(y0[2])) | (uint32_t(f32_to_bf16(y0[3])) << 16);
        const uint32_t h2 = uint32_t(f32_to_bf16(y1[0])) | (uint32_t(f32_to_bf16(y1[1])) << 16);
        const uint32_t h3 = uint32_t(f32_to_bf16(y1[2])) | (uint32_t(f32_to_bf16(y1[3])) << 16);
        *reinterpret_cast<uint2*>(x.Bout16 + size_t(s0) * a.ldh + b0) = make_uint2(h0, h1);
        *reinterpret_cast<uint2*>(x.Bout16 + size_t(s0 + 1) * a.ldh + b0) = make_uint2(h2, h3);
    }
}

// Update of the side pair (s0, s0 + 1) at the 4 bands b0 .. b0 + 3 from the
// gathered sums (one facet per team): MODE_SWEEP the Jacobi (or Chebyshev)
// update, MODE_INCIDENT the incident H = max(0, fsky S + G + D).
template <int MODE>
__device__ __forceinline__ void band_update4(const SweepArgs& a, const BandCtx& x, uint32_t s0, uint32_t b0,
                                             float (&g0)[4], float (&g1)[4], float& myAbs, float& myRel,
                                             float& myAct) {
    if (b0 >= x.nPrev) return;
    if (MODE == MODE_INCIDENT) {
        const size_t i0 = size_t(s0) * a.ldb + b0, i1 = i0 + a.ldb;
        float sk[4], D0[4], D1[4], H0[4], H1[4];
        arr4(ld4(a.skyv + b0), sk);
        arr4(ld4(a.D + i0), D0);
        arr4(ld4(a.D + i1), D1);
        const float f0 = a.fsky[s0], f1 = a.fsky[s0 + 1];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            H0[j] = fmaxf(f0 * sk[j] + g0[j] + D0[j], 0.0f);
            H1[j] = fmaxf(f1 * sk[j] + g1[j] + D1[j], 0.0f);
        }
        st4(a.H + i0, H0);
        st4(a.H + i1, H1);
        return;
    }
    Upd4 u;
    if (upd_begin<0>(a, x, s0, b0, g0, g1, u)) upd_end<0>(a, x, s0, b0, u, u.n0, u.n1, myAbs, myRel, myAct);
}

// Row update of an increment sweep (one facet per team); BASE: base sweep.
template <bool BASE>
__device__ __forceinline__ void band_update_inc4(const SweepArgs& a, const BandCtx& x, uint32_t s0, uint32_t b0,
                                                 float (&g0)[4], float (&g1)[4], float& myAbs, float& myRel,
                                                 float& myAct) {
    Upd4 u;
    if (upd_begin<BASE ? 1 : 2>(a, x, s0, b0, g0, g1, u))
        upd_end<BASE ? 1 : 2>(a, x, s0, b0, u, u.n0, u.n1, myAbs, myRel, myAct);
}

// Pair blocks.  With the CSR entries between the two facets (f, g) of a
// unit taken at the new values, the changes D = B' - B satisfy
// Df = jf + Af Dg and Dg = jg + Ag Df: j = n - B (plus A carry of the
// partner in increment sweeps, whose gathers saw Q = B - carry), and A
// (2 x 2 per band) maps the partner's side changes into this facet's
// Jacobi value through the unit weights W of its rows: A00 = r0 W00 +
// t0 W10, A01 = r0 W01 + t0 W11, A10 = r1 W10 + t1 W00, A11 = r1 W11 +
// t1 W01.  So (I - Af Ag) Df = jf + Af jg, then Dg; a near-singular block
// (det <= 1e-6) keeps the Jacobi values.  The new value is written as the
// Jacobi value plus the pair correction, B'f = nf + Af (Dg + carry_g),
// and a correction within 2^-21 of nf (a few ulps) is dropped: near the
// fixed point the sweep is then exactly the FP32 Jacobi map, whose fixed
// point it reaches (B' = B + D would let partners flip each other's last
// bit indefinitely).  To keep the register footprint of a team at two
// facets small, a facet's state between the gathers and the solve is only
// its Jacobi change, A and carry; B and the carry are read again for the
// stores.
struct Pair4 {
    float j0[4], j1[4];                      // Jacobi change n - B
    float a00[4], a01[4], a10[4], a11[4];    // A
    float k0[4], k1[4];                      // carry (increment sweeps)
};

template <int INC>
__device__ __forceinline__ bool pair_begin(const SweepArgs& a, const BandCtx& x, uint32_t s0, uint32_t b0,
                                           float (&g0)[4], float (&g1)[4], const float4 W, Pair4& p) {
    Upd4 u;
    if (!upd_begin<INC>(a, x, s0, b0, g0, g1, u)) return false;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        p.j0[j] = u.n0[j] - u.o0[j];
        p.j1[j] = u.n1[j] - u.o1[j];
        p.a00[j] = u.r0[j] * W.x + u.t0[j] * W.z;
        p.a01[j] = u.r0[j] * W.y + u.t0[j] * W.w;
        p.a10[j] = u.r1[j] * W.z + u.t1[j] * W.x;
        p.a11[j] = u.r1[j] * W.w + u.t1[j] * W.y;
        p.k0[j] = u.k0[j];
        p.k1[j] = u.k1[j];
    }
    return true;
}

// pair corrections c = A (D + carry) of the partner, per facet side
__device__ __forceinline__ void pair_solve4(const Pair4& f, const Pair4& g, float (&cf0)[4], float (&cf1)[4],
                                            float (&cg0)[4], float (&cg1)[4]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float jf0 = f.j0[j] + f.a00[j] * g.k0[j] + f.a01[j] * g.k1[j];
        const float jf1 = f.j1[j] + f.a10[j] * g.k0[j] + f.a11[j] * g.k1[j];
        const float jg0 = g.j0[j] + g.a00[j] * f.k0[j] + g.a01[j] * f.k1[j];
        const float jg1 = g.j1[j] + g.a10[j] * f.k0[j] + g.a11[j] * f.k1[j];
        const float p00 = f.a00[j] * g.a00[j] + f.a01[j] * g.a10[j];
        const float p01 = f.a00[j] * g.a01[j] + f.a01[j] * g.a11[j];
        const float p10 = f.a10[j] * g.a00[j] + f.a11[j] * g.a10[j];
        const float p11 = f.a10[j] * g.a01[j] + f.a11[j] * g.a11[j];
        const float det = (1.0f - p00) * (1.0f - p11) - p01 * p10;
        cf0[j] = cf1[j] = cg0[j] = cg1[j] = 0.0f;
        if (det > 1.0e-6f) {
            const float h0 = jf0 + f.a00[j] * jg0 + f.a01[j] * jg1;
            const float h1 = jf1 + f.a10[j] * jg0 + f.a11[j] * jg1;
            const float inv = 1.0f / det;
            const float df0 = ((1.0f - p11) * h0 + p01 * h1) * inv;
            const float df1 = (p10 * h0 + (1.0f - p00) * h1) * inv;
            const float dg0 = jg0 + g.a00[j] * df0 + g.a01[j] * df1;
            const float dg1 = jg1 + g.a10[j] * df0 + g.a11[j] * df1;
            const float ug0 = dg0 + g.k0[j], ug1 = dg1 + g.k1[j], uf0 = df0 + f.k0[j], uf1 = df1 + f.k1[j];
            cf0[j] = f.a00[j] * ug0 + f.a01[j] * ug1;
            cf1[j] = f.a10[j] * ug0 + f.a11[j] * ug1;
            cg0[j] = g.a00[j] * uf0 + g.a01[j] * uf1;
            cg1[j] = g.a10[j] * uf0 + g.a11[j] * uf1;
        }
    }
}

// stores of one facet of a pair: B and the carry read again, n = B + j
// (exact where B and n are within a factor 2), m = n + c beyond 2^-21 |n|
// (KEEP_V: the next increment before rounding into v, see upd_end)
template <int INC, bool KEEP_V = false>
__device__ __forceinline__ void pair_end(const SweepArgs& a, const BandCtx& x, uint32_t s0, uint32_t b0,
                                         const float (&j0)[4], const float (&j1)[4], const float (&c0)[4],
                                         const float (&c1)[4], float& myAbs, float& myRel, float& myAct,
                                         float (&v)[8]) {
    Upd4 u;
    u.i0 = size_t(s0) * a.ldb + b0;
    u.i1 = u.i0 + a.ldb;
    arr4(ld4(x.Bin + u.i0), u.o0);
    arr4(ld4(x.Bin + u.i1), u.o1);
    if (INC == 1) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            u.k0[j] = u.o0[j] - bf16_to_f32(f32_to_bf16(u.o0[j]));
            u.k1[j] = u.o1[j] - bf16_to_f32(f32_to_bf16(u.o1[j]));
        }
    } else if (INC == 2) {
        arr4(ld4(a.carry + u.i0), u.k0);
        arr4(ld4(a.carry + u.i1), u.k1);
    }
    constexpr float kDrop = 4.76837158e-07f;  // 2^-21
    float m0[4], m1[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        u.act[j] = b0 + uint32_t(j) < x.nAct;
        u.n0[j] = u.o0[j] + j0[j];
        u.n1[j] = u.o1[j] + j1[j];
        m0[j] = fabsf(c0[j]) > kDrop * fabsf(u.n0[j]) ? u.n0[j] + c0[j] : u.n0[j];
        m1[j] = fabsf(c1[j]) > kDrop * fabsf(u.n1[j]) ? u.n1[j] + c1[j] : u.n1[j];
    }
    upd_end<INC, KEEP_V>(a, x, s0, b0, u, m0, m1, myAbs, myRel, myAct);
    if (KEEP_V) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            v[j] = u.k0[j];
            v[4 + j] = u.k1[j];
        }
    }
}

// Update of a unit's two facets (side pairs s0 and s1) at 4 bands; INC as
// upd_begin.  KEEP_V: vf / vg receive each facet's next increments before
// rounding (sides 0 and 1, 4 bands each); returns whether the bands were
// updated (false: beyond the active bands).
template <int INC, bool KEEP_V = false>
__device__ __forceinline__ bool pair_update4(const SweepArgs& a, const BandCtx& x, uint32_t s0, uint32_t s1,
                                             uint32_t b0, const float4 Wf, const float4 Wg, float (&gf0)[4],
                                             float (&gf1)[4], float (&gg0)[4], float (&gg1)[4], float& myAbs,
                                             float& myRel, float& myAct, float (&vf)[8], float (&vg)[8]) {
    Pair4 pf, pg;
    const bool okf = pair_begin<INC>(a, x, s0, b0, gf0, gf1, Wf, pf);
    const bool okg = pair_begin<INC>(a, x, s1, b0, gg0, gg1, Wg, pg);  // same bands: same outcome as f
    if (!okf || !okg) return false;
    float cf0[4], cf1[4], cg0[4], cg1[4];
    pair_solve4(pf, pg, cf0, cf1, cg0, cg1);
    pair_end<INC, KEEP_V>(a, x, s0, b0, pf.j0, pf.j1, cf0, cf1, myAbs, myRel, myAct, vf);
    pair_end<INC, KEEP_V>(a, x, s1, b0, pg.j0, pg.j1, cg0, cg1, myAbs, myRel, myAct, vg);
    return true;
}
template <int INC>
__device__ __forceinline__ void pair_update4(const SweepArgs& a, const BandCtx& x, uint32_t s0, uint32_t s1,
                                             uint32_t b0, const float4 Wf, const float4 Wg, float (&gf0)[4],
                                             float (&gf1)[4], float (&gg0)[4], float (&gg1)[4], float& myAbs,
                                             float& myRel, float& myAct) {
    float vf[8], vg[8];
    (void)pair_update4<INC, false>(a, x, s0, s1, b0, Wf, Wg, gf0, gf1, gg0, gg1, myAbs, myRel, myAct, vf, vg);
}

__device__ __forceinline__ void band_partials(const SweepArgs& a, float myAbs, float myRel, float myAct) {
    float v[3] = {myAbs, myRel, myAct}, out[3];
    block_max<3>(v, out);
    if (threadIdx.x == 0) {
        a.partial[a.partOff + blockIdx.x] = make_float2(out[0], out[1]);
        if (a.partAct) a.partAct[a.partOff + blockIdx.x] = uint32_t(out[2]);
    }
}

// Sums of the two rows of a facet (entries [rb[0], rb[1]) and [rb[1],
// rb[2])) at the 4 FP32 bands of this lane, Bl = Bin + b0.
template <bool W16>
__device__ __forceinline__ void gather_v(const SweepArgs& a, const float* __restrict__ Bl, uint32_t ldb,
                                         const uint32_t (&rb)[3], float (&g)[2][4]) {
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const uint32_t start = rb[side], end = rb[side + 1];
        for (uint32_t base = start & ~uint32_t(kU - 1); base < end; base += kU) {
            uint32_t n[kU];
            float w[kU];
            bool ok[kU];
            load_group<W16>(a, base, start, end, n, w, ok);
            float4 v[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u)
                v[u] = ok[u] ? *reinterpret_cast<const float4*>(Bl + size_t(n[u]) * ldb)
                             : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const float vv[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
                for (int j = 0; j < 4; ++j) g[side][j] += w[u] * vv[j];
            }
        }
    }
}

// Buffer resource over `bytes` bytes at `base` (raw buffer, no swizzle):
// loads at byte offsets past `bytes` return zero without touching memory.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, int(bytes), 0x00020000);
}
constexpr uint32_t kNoRow = 0xfffffff0u;  // byte offset past any gathered array: loads zero

// FP8 increments (k_sweep_q8_single / _pair): a row of ldq bytes holds 16-byte chunks of 12
// bands, bytes 0..11 the bands' values in FP8 (E4M3) and byte 12 the
// chunk's power-of-two exponent e (int8): value = fp8 * 2^e.  Each chunk
// is rounded with its own exponent, chosen so that its largest magnitude
// becomes [64, 128) (inside E4M3's range whatever its bias): values more
// than 2^15 below it round to zero, which the carry keeps for later.
constexpr uint32_t kQ8Bands = 12;

__device__ __forceinline__ float pow2f(int e) { return __uint_as_float(uint32_t(e + 127) << 23); }  // e in [-126, 127]

// the 4 FP8 values of a 32-bit word
__device__ __forceinline__ void fp8x4(int w, float (&o)[4]) {
    const auto lo = __builtin_amdgcn_cvt_pk_f32_fp8(w, false);
    const auto hi = __builtin_amdgcn_cvt_pk_f32_fp8(w, true);
    o[0] = lo[0];
    o[1] = lo[1];
    o[2] = hi[0];
    o[3] = hi[1];
}

// 4 values (already scaled) rounded to FP8, packed in a 32-bit word
__device__ __forceinline__ int to_fp8x4(float a, float b, float c, float d) {
    const int w = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
    return __builtin_amdgcn_cvt_pk_fp8_f32(c, d, w, true);
}

__device__ __forceinline__ int chunk_exponent(float m) {
    if (!(m > 0.0f)) return 0;
    const int k = max(int((__float_as_uint(m) >> 23) & 0xffu) - 127, -120);  // floor(log2 m), tiny m floored
    return k - 6;
}

// Gathers of both rows of a facet with the CSR loads software-pipelined.
// The aligned groups of 8 entries of row 0 ([rb[0], rb[1])) and then of
// row 1 ([rb[1], rb[2])) form one sequence; the columns and weights of the
// next group are loaded while this group's gathers are in flight, so a
// group costs one round of memory latency instead of two, and the first
// group of row 1 overlaps the last one of row 0.  One accumulator: it is
// handed to side 0 where row 1 begins.  Each gather is a 16-byte buffer
// load at byte offset col * rowBytes + laneBytes of `rows`; entries outside
// the row load at kNoRow, which returns zero without a memory access.  The
// 16 bytes are 4 FP32 bands (KIND 0), 8 BF16 bands (KIND 1) or an FP8 chunk
// (KIND 2) of which the lane takes word `wsel` (4 bands) and the exponent.
enum { GATHER_F32 = 0, GATHER_BF16 = 1, GATHER_FP8 = 2, GATHER_FP8_CHUNK = 3 };

// (GATHER_FP8_CHUNK: the lane takes the whole FP8 chunk, 12 bands)
template <int KIND>
struct RowSums {
    static constexpr int NB = KIND == GATHER_BF16 ? 8 : (KIND == GATHER_FP8_CHUNK ? 12 : 4);
    float s0[NB], s1[NB];
};

template <int KIND>
__device__ __forceinline__ RowSums<KIND> gather_rows_pipe(const SweepArgs& a, const uint32_t (&rb)[3],
                                                          __amdgpu_buffer_rsrc_t rows, uint32_t rowBytes,
                                                          uint32_t laneBytes, uint32_t wsel = 0u) {
    constexpr int NB = RowSums<KIND>::NB;
    const uint32_t lo0 = rb[0], lo1 = rb[1], hi1 = rb[2];
    const uint32_t base0 = lo0 & ~uint32_t(kU - 1), base1 = lo1 & ~uint32_t(kU - 1);
    const uint32_t n0 = lo1 > lo0 ? (lo1 - base0 + uint32_t(kU - 1)) / uint32_t(kU) : 0u;
    const uint32_t n1 = hi1 > lo1 ? (hi1 - base1 + uint32_t(kU - 1)) / uint32_t(kU) : 0u;
    const uint32_t n = n0 + n1;
    float acc[NB], r0[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[j] = r0[j] = 0.0f;
    uint32_t base = n0 ? base0 : base1;
    uint4 cA = make_uint4(0u, 0u, 0u, 0u), cB = cA;
    float4 wA = make_float4(0.0f, 0.0f, 0.0f, 0.0f), wB = wA;
    if (n) {
        cA = *reinterpret_cast<const uint4*>(a.col + base);
        cB = *reinterpret_cast<const uint4*>(a.col + base + 4);
        wA = *reinterpret_cast<const float4*>(a.w32 + base);
        wB = *reinterpret_cast<const float4*>(a.w32 + base + 4);
    }
    for (uint32_t k = 0; k < n; ++k) {
        // row 1 begins: the sums so far are row 0's
        const bool hand = k == n0;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            r0[j] = hand ? acc[j] : r0[j];
            acc[j] = hand ? 0.0f : acc[j];
        }
        const uint32_t lo = k < n0 ? lo0 : lo1, hi = k < n0 ? lo1 : hi1;
        const uint32_t c[kU] = {cA.x, cA.y, cA.z, cA.w, cB.x, cB.y, cB.z, cB.w};
        const float w[kU] = {wA.x, wA.y, wA.z, wA.w, wB.x, wB.y, wB.z, wB.w};
        uint4 v[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const uint32_t e = base + uint32_t(u);
            const uint32_t off = e >= lo && e < hi ? __umul24(c[u], rowBytes) + laneBytes : kNoRow;
            v[u] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rows, off, 0, 0));
        }
        if (k + 1u < n) {  // the next group lies within the CSR (and its padding)
            base = k + 1u < n0 ? base0 + (k + 1u) * uint32_t(kU) : base1 + (k + 1u - n0) * uint32_t(kU);
            cA = *reinterpret_cast<const uint4*>(a.col + base);
            cB = *reinterpret_cast<const uint4*>(a.col + base + 4);
            wA = *reinterpret_cast<const float4*>(a.w32 + base);
            wB = *reinterpret_cast<const float4*>(a.w32 + base + 4);
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const uint32_t vv[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
            if constexpr (KIND == GATHER_BF16) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    acc[2 * j] += w[u] * __uint_as_float(vv[j] << 16);
                    acc[2 * j + 1] += w[u] * __uint_as_float(vv[j] & 0xffff0000u);
                }
            } else if constexpr (KIND == GATHER_FP8_CHUNK) {
                const float ws = w[u] * pow2f(int(int8_t(vv[3] & 0xffu)));
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    float q[4];
                    fp8x4(int(vv[t]), q);
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[4 * t + j] += ws * q[j];
                }
            } else if constexpr (KIND == GATHER_FP8) {
                // the chunk's exponent folded into the weight (exact: a power of two)
                const float ws = w[u] * pow2f(int(int8_t(vv[3] & 0xffu)));
                float q[4];
                fp8x4(int(wsel == 0u ? vv[0] : (wsel == 1u ? vv[1] : vv[2])), q);
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[j] += ws * q[j];
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[j] += w[u] * __uint_as_float(vv[j]);
            }
        }
    }
    RowSums<KIND> out;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        out.s0[j] = n1 ? r0[j] : acc[j];
        out.s1[j] = n1 ? acc[j] : 0.0f;
    }
    return out;
}

// FP32 gathers, 4 bands per lane: MODE_SWEEP with FP32 (W16 false) or BF16
// weights, or MODE_INCIDENT.  PAIR (MODE_SWEEP, FP32 weights): team i
// updates the two facets of units[i] with the pair block solve.
template <int MODE, bool W16, bool PAIR>
__global__ __launch_bounds__(256) void k_sweep_v(SweepArgs a) {
    BandCtx x;
    if (!band_ctx<MODE>(a, MODE == MODE_SWEEP && W16, x)) return;
    const uint32_t bt = a.bt, tpb = 256u / bt, ldb = a.ldb;
    const uint32_t team = threadIdx.x / bt, lane = threadIdx.x - team * bt;
    const uint32_t fl = xcd_block(a.xcd) * tpb + team;
    float myAbs = 0.0f, myRel = 0.0f, myAct = 0.0f;
    if (team < tpb && fl < (PAIR ? a.nUnits : a.F)) {
        uint32_t f, h = kNoFacet;
        if (PAIR) {
            const uint2 un = a.units[fl];
            f = un.x;
            h = un.y;
        } else {
            f = a.sched ? a.sched[fl] : fl;
        }
        const uint32_t s0 = 2u * f;
        const uint32_t rb[3] = {a.rowptr[s0], a.rowptr[s0 + 1], a.rowptr[s0 + 2]};
        if (!PAIR) {
            for (uint32_t cb = 0; cb < x.nPrev; cb += 4u * bt) {
                const uint32_t b0 = cb + 4u * lane;
                float g[2][4] = {};
                if (b0 < x.nAct) {
                    if constexpr (!W16) {
                        const RowSums<GATHER_F32> rs =
                            gather_rows_pipe<GATHER_F32>(a, rb, make_rsrc(x.Bin, a.sides * ldb * 4u), ldb * 4u, b0 * 4u);
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            g[0][j] = rs.s0[j];
                            g[1][j] = rs.s1[j];
                        }
                    } else {
                        gather_v<W16>(a, x.Bin + b0, ldb, rb, g);
                    }
                }
                band_update4<MODE>(a, x, s0, b0, g[0], g[1], myAbs, myRel, myAct);
            }
        } else {
            const uint32_t s1 = 2u * h;
            const uint32_t rc[3] = {a.rowptr[s1], a.rowptr[s1 + 1], a.rowptr[s1 + 2]};
            const float4 Wf = a.unitW[2u * fl], Wg = a.unitW[2u * fl + 1u];
            for (uint32_t cb = 0; cb < x.nPrev; cb += 4u * bt) {
                const uint32_t b0 = cb + 4u * lane;
                float gf[2][4] = {}, gg[2][4] = {};
                if (b0 < x.nAct) {
                    if constexpr (!W16) {
                        const __amdgpu_buffer_rsrc_t rows = make_rsrc(x.Bin, a.sides * ldb * 4u);
                        const RowSums<GATHER_F32> rf = gather_rows_pipe<GATHER_F32>(a, rb, rows, ldb * 4u, b0 * 4u);
                        const RowSums<GATHER_F32> rg = gather_rows_pipe<GATHER_F32>(a, rc, rows, ldb * 4u, b0 * 4u);
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            gf[0][j] = rf.s0[j];
                            gf[1][j] = rf.s1[j];
                            gg[0][j] = rg.s0[j];
                            gg[1][j] = rg.s1[j];
                        }
                    } else {
                        gather_v<W16>(a, x.Bin + b0, ldb, rb, gf);
                        gather_v<W16>(a, x.Bin + b0, ldb, rc, gg);
                    }
                }
                pair_update4<0>(a, x, s0, s1, b0, Wf, Wg, gf[0], gf[1], gg[0], gg[1], myAbs, myRel, myAct);
            }
        }
    }
    if (MODE == MODE_INCIDENT) return;
    band_partials(a, myAbs, myRel, myAct);
}

// Sums of the two rows of a facet at the 8 BF16 bands of this lane,
// Hl = Bin16 + b0 (row stride ldh), BF16 weights (W16) or FP32.
template <bool W16>
__device__ __forceinline__ void gather_h(const SweepArgs& a, const uint16_t* __restrict__ Hl, uint32_t ldh,
                                         const uint32_t (&rb)[3], float (&g)[2][8]) {
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const uint32_t start = rb[side], end = rb[side + 1];
        for (uint32_t base = start & ~uint32_t(kU - 1); base < end; base += kU) {
            uint32_t n[kU];
            float w[kU];
            bool ok[kU];
            load_group<W16>(a, base, start, end, n, w, ok);
            uint4 v[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u)
                v[u] = ok[u] ? *reinterpret_cast<const uint4*>(Hl + size_t(n[u]) * ldh) : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const uint32_t vv[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    g[side][2 * j] += w[u] * __uint_as_float(vv[j] << 16);
                    g[side][2 * j + 1] += w[u] * __uint_as_float(vv[j] & 0xffff0000u);
                }
            }
        }
    }
}

// Sums of the two rows of a facet at 4 BF16 bands of this lane (8-byte
// loads), FP32 weights: the pair blocks' increment sweeps, which hold two
// facets' state per lane and so run fewer waves per SIMD: each lane keeps
// 8 G entries in flight (G aligned groups; the CSR padding of kCsrPad = 16
// entries covers the reads past a row's end).
template <int G>
__device__ __forceinline__ void gather_h4(const SweepArgs& a, const uint16_t* __restrict__ Hl, uint32_t ldh,
                                          const uint32_t (&rb)[3], float (&g)[2][4]) {
    static_assert(G * kU <= int(kCsrPad) + 1, "reads past the CSR padding");
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const uint32_t start = rb[side], end = rb[side + 1];
        for (uint32_t base = start & ~uint32_t(kU - 1); base < end; base += uint32_t(G * kU)) {
            uint32_t n[G][kU];
            float w[G][kU];
            bool ok[G][kU];
#pragma unroll
            for (int q = 0; q < G; ++q) load_group<false>(a, base + uint32_t(q * kU), start, end, n[q], w[q], ok[q]);
            uint2 v[G][kU];
#pragma unroll
            for (int q = 0; q < G; ++q)
#pragma unroll
                for (int u = 0; u < kU; ++u)
                    v[q][u] = ok[q][u] ? *reinterpret_cast<const uint2*>(Hl + size_t(n[q][u]) * ldh) : make_uint2(0u, 0u);
#pragma unroll
            for (int q = 0; q < G; ++q)
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    const uint32_t vv[2] = {v[q][u].x, v[q][u].y};
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        g[side][2 * j] += w[q][u] * __uint_as_float(vv[j] << 16);
                        g[side][2 * j + 1] += w[q][u] * __uint_as_float(vv[j] & 0xffff0000u);
                    }
                }
        }
    }
}

// BF16 sweep: gathers the BF16 copy Bin16 (row stride ldh, a multiple of 64
// bands) 8 bands per lane with BF16 weights; accumulates, updates and writes
// B in FP32 and refreshes Bout16.  INC = 1 / 2: base / increment sweep of the
// increment scheme (upd_begin), FP32 weights, Bin16 / Bout16 hold the BF16
// increments; PAIR (INC 1 / 2): team i updates the two facets of units[i]
// with the pair block solve, 4 bands per lane.
template <int INC, bool PAIR>
__global__ __launch_bounds__(256) void k_sweep_hv(SweepArgs a) {
    BandCtx x;
    if (!band_ctx<MODE_SWEEP>(a, INC == 0, x)) return;
    const uint32_t bt = a.bt, tpb = 256u / bt, ldh = a.ldh;
    const uint32_t team = threadIdx.x / bt, lane = threadIdx.x - team * bt;
    const uint32_t fl = xcd_block(a.xcd) * tpb + team;
    float myAbs = 0.0f, myRel = 0.0f, myAct = 0.0f;
    if (team < tpb && fl < (PAIR ? a.nUnits : a.F)) {
        uint32_t f, h = kNoFacet;
        if (PAIR) {
            const uint2 un = a.units[fl];
            f = un.x;
            h = un.y;
        } else {
            f = a.sched ? a.sched[fl] : fl;
        }
        const uint32_t s0 = 2u * f;
        const uint32_t rb[3] = {a.rowptr[s0], a.rowptr[s0 + 1], a.rowptr[s0 + 2]};
        if (!PAIR) {
            for (uint32_t cb = 0; cb < x.nPrev; cb += 8u * bt) {
                const uint32_t b0 = cb + 8u * lane;
                float g[2][8] = {};
                if (b0 < x.nAct && !a.fold) {
                    if constexpr (INC != 0) {
                        const RowSums<GATHER_BF16> rs =
                            gather_rows_pipe<GATHER_BF16>(a, rb, make_rsrc(x.Bin16, a.sides * ldh * 2u), ldh * 2u, b0 * 2u);
#pragma unroll
                        for (int j = 0; j < 8; ++j) {
                            g[0][j] = rs.s0[j];
                            g[1][j] = rs.s1[j];
                        }
                    } else {
                        gather_h<INC == 0>(a, x.Bin16 + b0, ldh, rb, g);
                    }
                }
                // two groups of 4 bands (the second may lie beyond ldb)
                float lo0[4], lo1[4], hi0[4], hi1[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    lo0[j] = g[0][j];
                    lo1[j] = g[1][j];
                    hi0[j] = g[0][4 + j];
                    hi1[j] = g[1][4 + j];
                }
                if constexpr (INC == 0) {
                    band_update4<MODE_SWEEP>(a, x, s0, b0, lo0, lo1, myAbs, myRel, myAct);
                    if (b0 + 4u < a.ldb) band_update4<MODE_SWEEP>(a, x, s0, b0 + 4u, hi0, hi1, myAbs, myRel, myAct);
                } else {
                    band_update_inc4<INC == 1>(a, x, s0, b0, lo0, lo1, myAbs, myRel, myAct);
                    if (b0 + 4u < a.ldb) band_update_inc4<INC == 1>(a, x, s0, b0 + 4u, hi0, hi1, myAbs, myRel, myAct);
                }
            }
        } else {
            // 4 bands per lane
            const uint32_t s1 = 2u * h;
            const uint32_t rc[3] = {a.rowptr[s1], a.rowptr[s1 + 1], a.rowptr[s1 + 2]};
            const float4 Wf = a.unitW[2u * fl], Wg = a.unitW[2u * fl + 1u];
            for (uint32_t cb = 0; cb < x.nPrev; cb += 4u * bt) {
                const uint32_t b0 = cb + 4u * lane;
                float gf[2][4] = {}, gg[2][4] = {};
                if (b0 < x.nAct && !a.fold) {
                    gather_h4<kPairGroups>(a, x.Bin16 + b0, ldh, rb, gf);
                    gather_h4<kPairGroups>(a, x.Bin16 + b0, ldh, rc, gg);
                }
                pair_update4<INC>(a, x, s0, s1, b0, Wf, Wg, gf[0], gf[1], gg[0], gg[1], myAbs, myRel, myAct);
            }
        }
    }
    band_partials(a, myAbs, myRel, INC == 0 ? 0.0f : myAct);
}

// ---- Staged gathers: the sweeps of the single facets ----------------------
//
// A workgroup owns kStageFacets consecutive facets of the sweep order (one
// team of bt lanes, 8 BF16 bands per lane, per facet) and gathers the
// distinct columns of their kStageRows rows once per sweep through the LDS.
// The block's distinct columns, in column order, are staged kStageSlots at a
// time (a chunk).  A row's entries in one chunk form one segment (slots grow
// along a row).  Per chunk every team sums two whole segments: the rows are
// ranked by segment length and team t takes ranks t and kStageRows - 1 - t
// (long with short), and a block-level copy of the CSR holds, per chunk and
// team, the two segments back to back, each padded to a multiple of 4
// entries of count 0, as 4 bytes: the slot's byte offset in its LDS plane
// (low half) and the entry's ray count in BF16 (high half; counts above 63
// split in parts of at most 63, 2 parts in 10 000 on the terrain graph);
// the single-band gathers read a copy of 2 bytes per entry (slot | count
// << 10) at the same places; a run descriptor per chunk and team gives the
// first entry, the groups of 4 of each segment and the two local rows.  A
// segment sums count x value (two entries at a time: each word of the two
// staged rows regrouped by band with v_perm, then one BF16 dot product per
// band against the pair's counts, FP32 accumulation), and the team adds the
// sum times 1 / den of the row to the row's FP32 sums in the LDS once (a
// row belongs to one team per chunk: no atomics, and the sums do not depend
// on scheduling).  The staged rows sit in one LDS plane per lane of the
// team (16 bytes per slot, planes 16 KB + 96 bytes apart, so the lanes of a
// team read distinct banks).  On the 468 M-entry terrain graph a block of
// 128 facets references about 15 k distinct columns for 60 k entries: the
// lines the sweep requests from L2 drop from 6.2e8 to 2.2e8, and the
// entries read the LDS instead of the vector L1.
constexpr uint32_t kStageFacets = 128u;                       // facets (teams) per workgroup
constexpr uint32_t kStageRows = 2u * kStageFacets;            // local rows
constexpr uint32_t kStageSlots = 1024u;                       // staged columns per chunk
constexpr uint32_t kStagePer = kStageSlots / kStageFacets;    // staged columns per team and chunk
constexpr uint32_t kStageMaxChunks = 32u;                     // chunks per block (build limit)
constexpr uint32_t kStageNone = 0xffffffffu;
constexpr uint32_t kStageMaxBands = 48u;  // bt <= 6: the LDS of stage_lds_bytes fits the CU's 160 KB

struct StageArgs {
    const uint32_t* dcol;   // distinct columns of block b, sorted, from dOff[b]
    const uint32_t* dOff;   // [blocks + 1]
    const uint32_t* mOff;   // [blocks] run descriptor of (chunk j, team t) of block b at mOff[b] + 64 j + t
    const uint2* meta;      // x: first entry; y: groups of row A | groups of row B << 8 | A << 16 | B << 24
    const uint4* key;       // padded entries, 4 per vector: 16 x slot in chunk | BF16 count << 16 (0: padding)
    const uint2* key16;     // the same entries, 2 bytes each: slot in chunk | count << 10
    const uint32_t* order;  // facet of the sweep order's position i (null: facet i)
    const uint32_t* den;    // [S] denominators: an entry's weight is count / den of its row
};

constexpr uint32_t kStagePlane = kStageSlots * 16u + 96u;  // bytes between the LDS planes of two lanes

// dynamic LDS of k_sweep_staged: staged BF16 rows (a plane per lane) and the
// rows' FP32 sums
__host__ __device__ constexpr uint32_t stage_lds_bytes(uint32_t bt) {
    return bt * kStagePlane + kStageRows * bt * 32u + kStageRows * 4u;
}

// max-reduce n floats over the blockDim.x / 64 waves of a block (at most 16)
template <int N>
__device__ __forceinline__ void block_max_waves(float (&v)[N], float (&out)[N]) {
#pragma unroll
    for (int q = 0; q < N; ++q)
#pragma unroll
        for (int m = 32; m > 0; m >>= 1) v[q] = fmaxf(v[q], __shfl_xor(v[q], m, 64));
    __shared__ float sh[16][N];
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < N; ++q) sh[threadIdx.x >> 6][q] = v[q];
    __syncthreads();
    const uint32_t nw = blockDim.x >> 6;
#pragma unroll
    for (int q = 0; q < N; ++q) {
        float r = sh[0][q];
        for (uint32_t w = 1; w < nw; ++w) r = fmaxf(r, sh[w][q]);
        out[q] = r;
    }
}

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

// s += ca * (8 BF16 values of the staged row of key ka) + cb * (those of
// kb): per word (bands 2i, 2i + 1) the two rows' values of band 2i, and of
// band 2i + 1, are paired by v_perm and dotted with the pair (ca, cb) of
// BF16 counts (products exact in FP32)
__device__ __forceinline__ void stage_pair(float (&s)[8], const uint8_t* plane, uint32_t ka, uint32_t kb) {
    const uint4 va = *reinterpret_cast<const uint4*>(plane + (ka & 0xffffu));
    const uint4 vb = *reinterpret_cast<const uint4*>(plane + (kb & 0xffffu));
    const bf16x2_t c = __builtin_bit_cast(bf16x2_t, __builtin_amdgcn_perm(kb, ka, 0x07060302u));
    const uint32_t wa[4] = {va.x, va.y, va.z, va.w}, wb[4] = {vb.x, vb.y, vb.z, vb.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t lo = __builtin_amdgcn_perm(wb[i], wa[i], 0x05040100u);
        const uint32_t hi = __builtin_amdgcn_perm(wb[i], wa[i], 0x07060302u);
        s[2 * i] = __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2_t, lo), c, s[2 * i], false);
        s[2 * i + 1] = __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2_t, hi), c, s[2 * i + 1], false);
    }
}

template <int INC>
__global__ __launch_bounds__(1024) void k_sweep_staged(SweepArgs a, StageArgs st) {
    extern __shared__ uint4 stageLds[];
    BandCtx x;
    if (!band_ctx<MODE_SWEEP>(a, false, x)) return;
    const uint32_t bt = a.bt, nb8 = 8u * bt;
    const uint32_t tid = threadIdx.x, team = tid / bt, lane = tid - team * bt;
    const uint32_t blk = xcd_block(a.xcd);
    uint8_t* plane = reinterpret_cast<uint8_t*>(stageLds) + lane * kStagePlane;  // [kStageSlots] x 16 bytes
    float* acc = reinterpret_cast<float*>(reinterpret_cast<uint8_t*>(stageLds) + bt * kStagePlane);  // [kStageRows][nb8]
    float* rinv = acc + kStageRows * nb8;  // [kStageRows] 1 / den of the local rows
    const uint32_t b0 = 8u * lane;
    const bool act = b0 < x.nAct;
    for (uint32_t i = tid; i < kStageRows * nb8 / 4u; i += blockDim.x)
        reinterpret_cast<float4*>(acc)[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (uint32_t lr = tid; lr < kStageRows; lr += blockDim.x) {
        const uint32_t i = blk * kStageFacets + (lr >> 1);
        const uint32_t d = i < a.F ? st.den[2u * (st.order ? st.order[i] : i) + (lr & 1u)] : 0u;
        rinv[lr] = d ? 1.0f / float(d) : 0.0f;
    }
    const uint32_t d0 = st.dOff[blk], U = st.dOff[blk + 1u] - d0;
    const uint32_t nCh = a.fold ? 0u : (U + kStageSlots - 1u) / kStageSlots;  // (a folding sweep gathers nothing)
    const uint32_t* dc = st.dcol + d0;
    const uint2* meta = st.meta + st.mOff[blk];
    const uint32_t rowBytes = a.ldh * 2u;
    const __amdgpu_buffer_rsrc_t src = make_rsrc(x.Bin16, a.sides * rowBytes);
    // staging: this thread moves 16 bytes (its lane's 8 bands) of the slots
    // team + kStageFacets k of a chunk; the columns are read one chunk ahead
    uint32_t dq[kStagePer];
    uint4 rv[kStagePer];
    const auto load_dcol = [&](uint32_t j) {
#pragma unroll
        for (uint32_t k = 0; k < kStagePer; ++k) {
            const uint32_t s = j * kStageSlots + team + kStageFacets * k;
            dq[k] = s < U ? dc[s] : kStageNone;
        }
    };
    const auto load_rows = [&]() {  // (lanes of retired bands load nothing)
#pragma unroll
        for (uint32_t k = 0; k < kStagePer; ++k) {
            const uint32_t off = dq[k] != kStageNone && act ? dq[k] * rowBytes + lane * 16u : kNoRow;
            rv[k] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(src, off, 0, 0));
        }
    };
    const auto store_rows = [&]() {
#pragma unroll
        for (uint32_t k = 0; k < kStagePer; ++k)
            *reinterpret_cast<uint4*>(plane + (team + kStageFacets * k) * 16u) = rv[k];
    };
    // Everything a chunk needs from global memory (its staged rows, its run
    // descriptor and the first two groups of the run) is loaded during the
    // chunk before, so a chunk waits for one batch of loads, not three
    // dependent ones; only the later groups of a run are loaded in the run.
    uint2 mNext = make_uint2(0u, 0u), mNext2 = mNext;
    uint4 k1 = make_uint4(0u, 0u, 0u, 0u), k2 = k1;
    const auto first_groups = [&](uint2 m) {  // the first two groups of the run of m
        k1 = st.key[m.x >> 2];
        k2 = st.key[(m.x >> 2) + 1u];
    };
    if (nCh) {
        mNext = meta[team];
        load_dcol(0);
        load_rows();
        if (nCh > 1u) {
            load_dcol(1);
            mNext2 = meta[kStageFacets + team];
        }
        store_rows();
        first_groups(mNext);
    }
    __syncthreads();
    for (uint32_t j = 0; j < nCh; ++j) {
        const uint2 m = mNext;
        mNext = mNext2;
        const uint32_t qa = m.y & 255u, qn = qa + ((m.y >> 8) & 255u);
        const uint32_t rA = (m.y >> 16) & 255u, rB = m.y >> 24;
        const uint32_t g0 = m.x >> 2;  // first group of 4 entries
        // segment A (groups [0, qa)), then segment B ([qa, qn)); k1 and k2
        // hold groups q and q + 1, each reloaded (group q + 2, q + 3) once
        // it is consumed, so the loop moves no key registers.  The loads are
        // unconditional (a vmcnt wait then covers only the older load): they
        // read at most two groups past the run, inside the layout or the key
        // array's slack of 4 kCsrPad entries
        uint32_t q = 0;
        for (int seg = 0; seg < 2; ++seg) {
            const uint32_t qe = seg ? qn : qa;
            float s[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) s[t] = 0.0f;
            const bool any = q < qe;
            for (; q + 2u <= qe; q += 2u) {
                if (act) {
                    stage_pair(s, plane, k1.x, k1.y);
                    stage_pair(s, plane, k1.z, k1.w);
                }
                k1 = st.key[g0 + q + 2u];
                if (act) {
                    stage_pair(s, plane, k2.x, k2.y);
                    stage_pair(s, plane, k2.z, k2.w);
                }
                k2 = st.key[g0 + q + 3u];
            }
            if (q < qe) {  // one group left in the segment
                if (act) {
                    stage_pair(s, plane, k1.x, k1.y);
                    stage_pair(s, plane, k1.z, k1.w);
                }
                k1 = k2;
                k2 = st.key[g0 + q + 2u];
                ++q;
            }
            if (act && any) {  // the segment's count-weighted sum times 1 / den into its row
                const float y = rinv[seg ? rB : rA];
                float4* p = reinterpret_cast<float4*>(acc + __umul24(seg ? rB : rA, nb8) + b0);
                const float4 u0 = p[0], u1 = p[1];
                p[0] = make_float4(u0.x + y * s[0], u0.y + y * s[1], u0.z + y * s[2], u0.w + y * s[3]);
                p[1] = make_float4(u1.x + y * s[4], u1.y + y * s[5], u1.z + y * s[6], u1.w + y * s[7]);
            }
        }
        // the next chunk's loads, issued after this chunk's run loads (loads
        // complete in order)
        if (j + 1u < nCh) {
            load_rows();
            if (j + 2u < nCh) {
                load_dcol(j + 2u);
                mNext2 = meta[(j + 2u) * kStageFacets + team];
            }
            first_groups(mNext);
        }
        __syncthreads();
        if (j + 1u < nCh) store_rows();
        __syncthreads();
    }
    float myAbs = 0.0f, myRel = 0.0f, myAct = 0.0f;
    const uint32_t fl = blk * kStageFacets + team;
    if (fl < a.F) {
        const uint32_t f = a.sched ? a.sched[fl] : fl;
        const uint32_t s0 = 2u * f;
        const float4* p0 = reinterpret_cast<const float4*>(acc + (2u * team) * nb8 + b0);
        const float4* p1 = reinterpret_cast<const float4*>(acc + (2u * team + 1u) * nb8 + b0);
        const float4 a0 = p0[0], a1 = p0[1], c0 = p1[0], c1 = p1[1];
        float lo0[4] = {a0.x, a0.y, a0.z, a0.w}, hi0[4] = {a1.x, a1.y, a1.z, a1.w};
        float lo1[4] = {c0.x, c0.y, c0.z, c0.w}, hi1[4] = {c1.x, c1.y, c1.z, c1.w};
        band_update_inc4<INC == 1>(a, x, s0, b0, lo0, lo1, myAbs, myRel, myAct);
        if (b0 + 4u < a.ldb) band_update_inc4<INC == 1>(a, x, s0, b0 + 4u, hi0, hi1, myAbs, myRel, myAct);
    }
    float v[3] = {myAbs, myRel, myAct}, out[3];
    block_max_waves<3>(v, out);
    if (threadIdx.x == 0) {
        a.partial[a.partOff + blockIdx.x] = make_float2(out[0], out[1]);
        if (a.partAct) a.partAct[a.partOff + blockIdx.x] = uint32_t(out[2]);
    }
}

// Single-band staged gathers (the coupled longwave sweep and the single-band
// sweeps of the final longwave solves): the same layout, one lane per team
// (a workgroup of kStageFacets lanes), the chunk's 1024 values of B staged in
// the LDS (4 KB) and every run's count-weighted sum, times 1 / den of its
// row, added to its row's sum there.  The random single-value gathers of
// every entry become one gather per distinct column and workgroup.  On
// return acc[2 t + side] holds the gather of the workgroup's t-th facet.
__device__ __forceinline__ void stage_gather1(const StageArgs& st, const float* __restrict__ B, uint32_t nF,
                                              uint32_t blk, uint32_t t, float* vals, float* acc, float* rinv) {
    for (uint32_t lr = t; lr < kStageRows; lr += kStageFacets) {
        const uint32_t i = blk * kStageFacets + (lr >> 1);
        const uint32_t d = i < nF ? st.den[2u * (st.order ? st.order[i] : i) + (lr & 1u)] : 0u;
        rinv[lr] = d ? 1.0f / float(d) : 0.0f;
        acc[lr] = 0.0f;
    }
    const uint32_t d0 = st.dOff[blk], U = st.dOff[blk + 1u] - d0;
    const uint32_t nCh = (U + kStageSlots - 1u) / kStageSlots;
    const uint32_t* dc = st.dcol + d0;
    const uint2* meta = st.meta + st.mOff[blk];
    // a chunk's distinct columns and run descriptor are read two chunks
    // ahead, its staged values and the first two key groups of the run one
    // chunk ahead; the key loads are unconditional (at most two groups past
    // a run: the key array's slack)
    uint32_t dq[kStagePer];
    float v[kStagePer];
    const auto load_dcol = [&](uint32_t j) {
#pragma unroll
        for (uint32_t k = 0; k < kStagePer; ++k) {
            const uint32_t sl = j * kStageSlots + t + kStageFacets * k;
            dq[k] = sl < U ? dc[sl] : kStageNone;
        }
    };
    const auto load_vals = [&]() {
#pragma unroll
        for (uint32_t k = 0; k < kStagePer; ++k) v[k] = dq[k] != kStageNone ? B[dq[k]] : 0.0f;
    };
    const auto store_vals = [&]() {
#pragma unroll
        for (uint32_t k = 0; k < kStagePer; ++k) vals[t + kStageFacets * k] = v[k];
    };
    uint2 mNext = make_uint2(0u, 0u), mNext2 = mNext, k1 = mNext, k2 = mNext;
    const auto first_groups = [&](uint2 m) {
        k1 = st.key16[m.x >> 2];
        k2 = st.key16[(m.x >> 2) + 1u];
    };
    if (nCh) {
        mNext = meta[t];
        load_dcol(0);
        load_vals();
        if (nCh > 1u) {
            load_dcol(1);
            mNext2 = meta[kStageFacets + t];
        }
        store_vals();
        first_groups(mNext);
    }
    __syncthreads();
    // 2-byte keys, 4 per 8 bytes: slot | count << 10
    const auto term = [&](uint32_t k) { return float(k >> 10) * vals[k & 1023u]; };
    const auto terms = [&](uint2 k) {
        return term(k.x & 0xffffu) + term(k.x >> 16) + term(k.y & 0xffffu) + term(k.y >> 16);
    };
    for (uint32_t j = 0; j < nCh; ++j) {
        const uint2 m = mNext;
        mNext = mNext2;
        const uint32_t qa = m.y & 255u, qn = qa + ((m.y >> 8) & 255u);
        const uint32_t rA = (m.y >> 16) & 255u, rB = m.y >> 24;
        const uint32_t g0 = m.x >> 2;
        uint32_t q = 0;
        for (int seg = 0; seg < 2; ++seg) {  // segment A, then B: count-weighted sums times 1 / den
            const uint32_t qe = seg ? qn : qa;
            const bool any = q < qe;
            float sum = 0.0f;
            for (; q + 2u <= qe; q += 2u) {
                sum += terms(k1);
                k1 = st.key16[g0 + q + 2u];
                sum += terms(k2);
                k2 = st.key16[g0 + q + 3u];
            }
            if (q < qe) {
                sum += terms(k1);
                k1 = k2;
                k2 = st.key16[g0 + q + 2u];
                ++q;
            }
            if (any) acc[seg ? rB : rA] += rinv[seg ? rB : rA] * sum;
        }
        if (j + 1u < nCh) {
            load_vals();
            if (j + 2u < nCh) {
                load_dcol(j + 2u);
                mNext2 = meta[(j + 2u) * kStageFacets + t];
            }
            first_groups(mNext);
        }
        __syncthreads();  // the runs are done with vals
        if (j + 1u < nCh) store_vals();
        __syncthreads();
    }
}

// Iin = fsky L + G for the workgroup's facets (coupled longwave sweep)
__global__ __launch_bounds__(kStageFacets) void k_gather1_staged(StageArgs st, const float* __restrict__ B,
                                                                  const float* __restrict__ fsky, float L, uint32_t nF,
                                                                  const SolveCtrl* ctrl, float* __restrict__ Iin) {
    if (ctrl && ctrl_word(&ctrl->done)) return;
    __shared__ float vals[kStageSlots];
    __shared__ float acc[kStageRows];
    __shared__ float rinv[kStageRows];
    const uint32_t t = threadIdx.x, blk = blockIdx.x;
    stage_gather1(st, B, nF, blk, t, vals, acc, rinv);
    const uint32_t fl = blk * kStageFacets + t;
    if (fl < nF) {
        const uint32_t s0 = 2u * (st.order ? st.order[fl] : fl);
        Iin[s0] = fsky[s0] * L + acc[2u * t];
        Iin[s0 + 1u] = fsky[s0 + 1u] * L + acc[2u * t + 1u];
    }
}

// Single-band FP32 sweep of the staged facets (entry-parallel solves at one
// band: the final longwave solves): the staged gather of Bin, then the
// relaxed Jacobi update of k_sweep (fused incident in judge sweeps) and the
// block's residual partials at partOff + block
__global__ __launch_bounds__(kStageFacets) void k_sweep1_staged(SweepArgs a, StageArgs st) {
    SolveCtrl* ctrl = a.ctrl;
    if (ctrl_word(&ctrl->done)) return;
    const uint32_t it = ctrl_word(&ctrl->iter);
    const float* __restrict__ Bin = (it & 1u) ? a.buf1 : a.buf0;
    float* __restrict__ Bout = (it & 1u) ? a.buf0 : a.buf1;
    __shared__ float vals[kStageSlots];
    __shared__ float acc[kStageRows];
    __shared__ float rinv[kStageRows];
    const uint32_t t = threadIdx.x, blk = blockIdx.x;
    stage_gather1(st, Bin, a.F, blk, t, vals, acc, rinv);
    float myAbs = 0.0f, myRel = 0.0f;
    const uint32_t fl = blk * kStageFacets + t;
    if (fl < a.F) {
        const uint32_t s0 = 2u * (st.order ? st.order[fl] : fl), s1 = s0 + 1u;
        const float g0 = acc[2u * t], g1 = acc[2u * t + 1u];
        const float n0 = a.c[s0] + a.r[s0] * g0 + a.t[s0] * g1;
        const float n1 = a.c[s1] + a.r[s1] * g1 + a.t[s1] * g0;
        if (a.fuseH && a.judge) {  // incident of this sweep's gather
            const float sk = a.skyv[0];
            a.H[s0] = fmaxf(a.fsky[s0] * sk + g0 + a.D[s0], 0.0f);
            a.H[s1] = fmaxf(a.fsky[s1] * sk + g1 + a.D[s1], 0.0f);
        }
        const float o0 = Bin[s0], o1 = Bin[s1];
        const float x0 = o0 + a.relax * (n0 - o0);
        const float x1 = o1 + a.relax * (n1 - o1);
        Bout[s0] = x0;
        Bout[s1] = x1;
        const float d0 = fabsf(x0 - o0), d1 = fabsf(x1 - o1);
        myAbs = fmaxf(d0, d1);
        myRel = fmaxf(d0 / fmaxf(fabsf(x0), 1.0e-30f), d1 / fmaxf(fabsf(x1), 1.0e-30f));
    }
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) {
        myAbs = fmaxf(myAbs, __shfl_xor(myAbs, m, 64));
        myRel = fmaxf(myRel, __shfl_xor(myRel, m, 64));
    }
    __shared__ float sAbs[kStageFacets / 64u], sRel[kStageFacets / 64u];
    if ((t & 63u) == 0u) {
        sAbs[t >> 6] = myAbs;
        sRel[t >> 6] = myRel;
    }
    __syncthreads();
    if (t == 0u) {
        float bA = 0.0f, bR = 0.0f;
        for (uint32_t w = 0; w < kStageFacets / 64u; ++w) {
            bA = fmaxf(bA, sAbs[w]);
            bR = fmaxf(bR, sRel[w]);
        }
        a.partial[a.partOff + blk] = make_float2(bA, bR);
    }
}

// ---- building the staged layout ------------------------------------------
// Blocks of kStageFacets facets of `order` (null: facet i); local row lr =
// 2 t + side of the block's t-th facet.  Built per batch of blocks (one
// workgroup each) with a bitmap over the sides per block: the bits of the
// block's columns, the word ranks and the sorted distinct columns, then the
// segments.  Two passes over the batches: the first sizes every block
// (distinct columns U, padded entries P, chunks), the second writes.

// side of local row lr of block b, or kStageNone past the facets
__device__ __forceinline__ uint32_t stage_side(const uint32_t* order, uint32_t nF, uint32_t b, uint32_t lr) {
    const uint32_t i = b * kStageFacets + (lr >> 1);
    if (i >= nF) return kStageNone;
    return 2u * (order ? order[i] : i) + (lr & 1u);
}

__global__ __launch_bounds__(1024) void k_stage_bits(const uint32_t* __restrict__ rowptr,
                                                     const uint32_t* __restrict__ col,
                                                     const uint32_t* __restrict__ order, uint32_t nF, uint32_t nBlk,
                                                     uint32_t b0, uint32_t words, uint32_t* __restrict__ bitmap) {
    const uint32_t b = b0 + blockIdx.x;
    if (b >= nBlk) return;
    uint32_t* bm = bitmap + size_t(blockIdx.x) * words;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    for (uint32_t lr = wave; lr < kStageRows; lr += 16u) {
        const uint32_t s = stage_side(order, nF, b, lr);
        if (s == kStageNone) continue;
        for (uint32_t e = rowptr[s] + lane; e < rowptr[s + 1u]; e += 64u) {
            const uint32_t c = col[e];
            atomicOr(bm + (c >> 5), 1u << (c & 31u));
        }
    }
}

// word ranks (set bits before the word), U[b], and (dOff set) the sorted
// distinct columns at dOff[b]
__global__ __launch_bounds__(1024) void k_stage_rank(const uint32_t* __restrict__ bitmap, uint32_t nBlk, uint32_t b0,
                                                     uint32_t words, uint32_t* __restrict__ rank,
                                                     uint32_t* __restrict__ U, const uint32_t* __restrict__ dOff,
                                                     uint32_t* __restrict__ dcol) {
    const uint32_t b = b0 + blockIdx.x;
    if (b >= nBlk) return;
    const uint32_t* bm = bitmap + size_t(blockIdx.x) * words;
    uint32_t* rk = rank + size_t(blockIdx.x) * words;
    const uint32_t tid = threadIdx.x, per = (words + 1023u) / 1024u;
    const uint32_t w0 = min(words, tid * per), w1 = min(words, w0 + per);
    uint32_t n = 0;
    for (uint32_t w = w0; w < w1; ++w) n += __popc(bm[w]);
    __shared__ uint32_t sc[1024];
    sc[tid] = n;
    __syncthreads();
    for (uint32_t off = 1; off < 1024u; off <<= 1) {
        const uint32_t v = tid >= off ? sc[tid - off] : 0u;
        __syncthreads();
        sc[tid] += v;
        __syncthreads();
    }
    uint32_t run = sc[tid] - n;
    if (tid == 1023u) U[b] = sc[1023];
    uint32_t* dc = dOff ? dcol + dOff[b] : nullptr;
    for (uint32_t w = w0; w < w1; ++w) {
        rk[w] = run;
        uint32_t m = bm[w];
        if (dc) {
            while (m) {
                dc[run++] = 32u * w + uint32_t(__ffs(m) - 1);
                m &= m - 1u;
            }
        } else {
            run += __popc(m);
        }
    }
}

// slot of column c in its block (rank of c among the block's columns)
__device__ __forceinline__ uint32_t stage_slot(const uint32_t* bm, const uint32_t* rk, uint32_t c) {
    return rk[c >> 5] + uint32_t(__popc(bm[c >> 5] & ((1u << (c & 31u)) - 1u)));
}

// Entries a CSR entry of count c takes in the staged layout: parts of at
// most kStagePart rays (the 6-bit count of the 2-byte keys; exact in BF16),
// entries of the same slot.
constexpr uint32_t kStagePart = 63u;
__device__ __forceinline__ uint32_t stage_parts(uint32_t c) { return (c + kStagePart - 1u) / kStagePart; }

// BF16 bits of an integer with at most 8 significant bits
__device__ __forceinline__ uint32_t stage_bf16_bits(uint32_t n) { return __float_as_uint(float(n)) >> 16; }

// Per row (one wave): the segments along the row.  fn(e, j, slot, at, n,
// lastOfSegment) for every entry e of the row: at is the place of its first
// layout entry in the segment (entries expanded by stage_parts), n its parts.
template <class Fn>
__device__ __forceinline__ void stage_row_runs(const uint32_t* __restrict__ col, const uint32_t* __restrict__ cnt,
                                               const uint32_t* bm, const uint32_t* rk, uint32_t e0, uint32_t e1,
                                               Fn&& fn) {
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t runJ = kStageNone, runCum = 0u, runStartCum = 0u;
    for (uint32_t base = e0; base < e1; base += 64u) {
        const uint32_t e = base + lane;
        const bool valid = e < e1;
        const uint32_t slot = valid ? stage_slot(bm, rk, col[e]) : 0u;
        const uint32_t j = valid ? slot / kStageSlots : kStageNone;
        const uint32_t n = valid ? stage_parts(cnt[e]) : 0u;
        uint32_t incl = n;  // layout entries of this batch up to e
        for (uint32_t o = 1; o < 64u; o <<= 1) {
            const uint32_t v = __shfl_up(incl, o, 64);
            if (lane >= o) incl += v;
        }
        const uint32_t cum = runCum + incl - n;  // layout entries of the row before e
        uint32_t prevJ = __shfl_up(j, 1u, 64);
        if (lane == 0u) prevJ = runJ;
        uint32_t nextJ = __shfl_down(j, 1u, 64);
        if (lane == 63u) nextJ = e + 1u < e1 ? stage_slot(bm, rk, col[e + 1u]) / kStageSlots : kStageNone;
        const bool start = valid && j != prevJ;
        const unsigned long long starts = __ballot(start);
        const unsigned long long upto = lane == 63u ? ~0ull : ((2ull << lane) - 1ull);
        const unsigned long long below = starts & upto;
        const uint32_t cumFirst = __shfl(cum, below ? 63 - __clzll(below) : 0, 64);
        const uint32_t segCum = below ? cumFirst : runStartCum;  // layout entries before the segment
        if (valid) fn(e, j, slot, cum - segCum, n, j != nextJ);
        const uint32_t lastLane = min(63u, e1 - 1u - base);
        runJ = __shfl(j, int(lastLane), 64);
        runStartCum = __shfl(segCum, int(lastLane), 64);
        runCum += __shfl(incl, 63, 64);
    }
}

__host__ __device__ constexpr uint32_t stage_fill_lds_bytes() {
    return 2u * kStageRows * kStageMaxChunks * 2u + 16u * kStageRows + 2u * (kStageMaxChunks + 1u) * 4u;
}

// Pass 1 (write false): U, P (padded entries) and chunks of the block.
// Pass 2 (write true): the run descriptors and the padded entries at
// eOff[b].  Segment lengths per (local row, chunk) in the LDS.
template <bool WRITE>
__global__ __launch_bounds__(1024) void k_stage_fill(const uint32_t* __restrict__ rowptr,
                                                     const uint32_t* __restrict__ col,
                                                     const uint32_t* __restrict__ order, uint32_t nF, uint32_t nBlk,
                                                     uint32_t b0, uint32_t words, const uint32_t* __restrict__ bitmap,
                                                     const uint32_t* __restrict__ rank, const uint32_t* __restrict__ U,
                                                     uint32_t* __restrict__ P, const uint32_t* __restrict__ eOff,
                                                     const uint32_t* __restrict__ mOff, uint2* __restrict__ meta,
                                                     const uint32_t* __restrict__ cnt, uint32_t* __restrict__ key,
                                                     uint16_t* __restrict__ key16, uint32_t* __restrict__ fault) {
    extern __shared__ uint32_t fillLds[];
    const uint32_t b = b0 + blockIdx.x;
    if (b >= nBlk) return;
    const uint32_t* bm = bitmap + size_t(blockIdx.x) * words;
    const uint32_t* rk = rank + size_t(blockIdx.x) * words;
    const uint32_t nCh = (U[b] + kStageSlots - 1u) / kStageSlots;
    if (nCh > kStageMaxChunks) {
        if (threadIdx.x == 0) atomicOr(fault, 1u);
        return;
    }
    uint16_t* len = reinterpret_cast<uint16_t*>(fillLds);   // [kStageRows][kStageMaxChunks] segment lengths
    uint16_t* pos = len + kStageRows * kStageMaxChunks;     // segment place in its chunk
    uint8_t* sorted = reinterpret_cast<uint8_t*>(pos + kStageRows * kStageMaxChunks);  // [16 waves][kStageRows]
    uint32_t* cT = reinterpret_cast<uint32_t*>(sorted + 16u * kStageRows);           // chunk totals, then starts
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63u;
    for (uint32_t i = tid; i < kStageRows * kStageMaxChunks; i += 1024u) len[i] = 0u;
    __syncthreads();
    for (uint32_t lr = wave; lr < kStageRows; lr += 16u) {
        const uint32_t s = stage_side(order, nF, b, lr);
        if (s == kStageNone) continue;
        stage_row_runs(col, cnt, bm, rk, rowptr[s], rowptr[s + 1u],
                       [&](uint32_t, uint32_t j, uint32_t, uint32_t at, uint32_t n, bool last) {
                           if (!last) return;
                           // a segment longer than 1020 entries does not fit the 8-bit group
                           // counts (nor, expanded by large counts, the 16-bit lengths)
                           if (at + n > 1020u) atomicOr(fault, 16u);
                           len[lr * kStageMaxChunks + j] = uint16_t(min(at + n, 1024u));
                       });
    }
    __syncthreads();
    const auto padded = [](uint32_t n) { return (n + 3u) & ~3u; };
    if (!WRITE) {
        uint32_t t = 0;
        for (uint32_t i = tid; i < kStageRows * nCh; i += 1024u) t += padded(len[(i / nCh) * kStageMaxChunks + i % nCh]);
        for (int m = 32; m > 0; m >>= 1) t += __shfl_xor(t, m, 64);
        __shared__ uint32_t red[16];
        if (lane == 0u) red[wave] = t;
        __syncthreads();
        if (tid == 0) {
            uint32_t sum = 0;
            for (uint32_t w = 0; w < 16u; ++w) sum += red[w];
            P[b] = sum;
        }
        return;
    }
    // per chunk (a wave each): rows ranked by padded length (longest first,
    // then by local row), team t takes ranks t and kStageRows - 1 - t;
    // each lane ranks kStageRows / 64 rows and pairs kStageFacets / 64 teams
    constexpr uint32_t kRowsPerLane = kStageRows / 64u, kTeamsPerLane = kStageFacets / 64u;
    constexpr uint32_t kChunksPerWave = (kStageMaxChunks + 15u) / 16u;
    uint2 mine[kChunksPerWave][kTeamsPerLane];
    for (uint32_t k = 0, j = wave; k < kChunksPerWave; ++k, j += 16u) {
        if (j >= nCh) break;
        uint8_t* srt = sorted + wave * kStageRows;
        for (uint32_t h = 0; h < kRowsPerLane; ++h) {
            const uint32_t lr = lane + 64u * h;
            const uint32_t kl = (padded(len[lr * kStageMaxChunks + j]) << 8) | (kStageRows - 1u - lr);
            uint32_t r = 0;
            for (uint32_t o = 0; o < kStageRows; ++o)
                r += ((padded(len[o * kStageMaxChunks + j]) << 8) | (kStageRows - 1u - o)) > kl ? 1u : 0u;
            srt[r] = uint8_t(lr);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        uint32_t base = 0;  // entries of the teams of the earlier lane passes
        for (uint32_t h = 0; h < kTeamsPerLane; ++h) {
            const uint32_t t = lane + 64u * h;
            const uint32_t A = srt[t], B = srt[kStageRows - 1u - t];
            const uint32_t pA = padded(len[A * kStageMaxChunks + j]), pB = padded(len[B * kStageMaxChunks + j]);
            if (pA > 1020u || pB > 1020u) atomicOr(fault, 16u);  // 8-bit group counts
            uint32_t inc = pA + pB;
            for (uint32_t o = 1; o < 64u; o <<= 1) {
                const uint32_t v = __shfl_up(inc, o, 64);
                if (lane >= o) inc += v;
            }
            const uint32_t start = base + inc - pA - pB;
            pos[A * kStageMaxChunks + j] = uint16_t(start);
            pos[B * kStageMaxChunks + j] = uint16_t(start + pA);
            mine[k][h] = make_uint2(start, (pA >> 2) | ((pB >> 2) << 8) | (A << 16) | (B << 24));
            base += __shfl(inc, 63, 64);
        }
        if (lane == 63u) {
            cT[j] = base;
            if (base > 0xffffu) atomicOr(fault, 8u);  // segment places are 16-bit
        }
        __builtin_amdgcn_wave_barrier();  // the next chunk's ranks overwrite this wave's list
    }
    __syncthreads();
    uint32_t* cst = cT + kStageMaxChunks + 1u;
    if (tid == 0) {
        uint32_t run = 0;
        for (uint32_t j = 0; j < nCh; ++j) {
            cst[j] = run;
            run += cT[j];
        }
        cst[nCh] = run;
        if (run != eOff[b + 1u] - eOff[b]) atomicOr(fault, 4u);
    }
    __syncthreads();
    const uint32_t eb = eOff[b], eEnd = eOff[b + 1u];
    for (uint32_t k = 0, j = wave; k < kChunksPerWave; ++k, j += 16u) {
        if (j >= nCh) break;
        for (uint32_t h = 0; h < kTeamsPerLane; ++h)
            meta[mOff[b] + j * kStageFacets + lane + 64u * h] = make_uint2(eb + cst[j] + mine[k][h].x, mine[k][h].y);
    }
    for (uint32_t lr = wave; lr < kStageRows; lr += 16u) {
        const uint32_t s = stage_side(order, nF, b, lr);
        if (s == kStageNone) continue;
        stage_row_runs(col, cnt, bm, rk, rowptr[s], rowptr[s + 1u],
                       [&](uint32_t e, uint32_t j, uint32_t slot, uint32_t p, uint32_t n, bool last) {
                           const uint32_t seg = eb + cst[j] + pos[lr * kStageMaxChunks + j];
                           const uint32_t at = seg + p;
                           const uint32_t end = last ? seg + padded(p + n) : at + n;
                           if (end > eEnd) {  // cannot happen with consistent counts; never write past the block
                               atomicOr(fault, 2u);
                               return;
                           }
                           const uint32_t sl = slot - j * kStageSlots;
                           for (uint32_t q = 0, c = cnt[e]; q < n; ++q, c -= kStagePart) {
                               const uint32_t part = min(c, kStagePart);
                               key[at + q] = (sl << 4) | (stage_bf16_bits(part) << 16);
                               key16[at + q] = uint16_t(sl | (part << 10));
                           }
                           for (uint32_t q = at + n; q < end; ++q) {  // the segment's padding
                               key[q] = 0u;
                               key16[q] = 0u;
                           }
                       });
    }
}

// FP8 increment sweep of the coupled pairs.  A team of bt lanes (a
// multiple of 3, teams never straddle a wave) owns the two facets of a
// unit; a lane owns 4 bands, so the three lanes of one 12-band chunk round
// it together:
// each lane's largest |v| of its side is max-reduced over the three by
// shuffles, the chunk exponent follows, and every lane rounds its 4 values,
// keeps what rounding drops in the carry and stores its 4 bytes (the first
// lane also the exponent).  The update itself is the increment sweep's pair
// block: FP32 weights, accumulated gathers and carries, FP32 Jacobi /
// Chebyshev update, per-band retirement.
//
// emit: this lane has new increments v (sides 0 and 1 of facet s0, 4 bands
// from b0; every lane of the wave calls this, emit or not, for the shuffles)
__device__ __forceinline__ void emit_q8(const SweepArgs& a, const BandCtx& x, uint32_t s0, uint32_t b0, uint32_t lane,
                                        uint32_t bt, bool emit, const float (&v)[8]) {
    float m0 = 0.0f, m1 = 0.0f;
    if (emit) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            m0 = fmaxf(m0, fabsf(v[j]));
            m1 = fmaxf(m1, fabsf(v[4 + j]));
        }
    }
    // the chunk's three lanes: lane - q .. lane - q + 2 of this team
    const uint32_t q = lane % 3u;
    const int wl = int(threadIdx.x & 63u);
    float c0 = m0, c1 = m1;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int src = wl - int(q) + k;
        const bool inTeam = lane - q + uint32_t(k) < bt;
        const float o0 = __shfl(m0, src & 63, 64), o1 = __shfl(m1, src & 63, 64);
        if (inTeam) {
            c0 = fmaxf(c0, o0);
            c1 = fmaxf(c1, o1);
        }
    }
    if (!emit) return;
    const int e0 = chunk_exponent(c0), e1 = chunk_exponent(c1);
    const float d0 = pow2f(-e0), d1 = pow2f(-e1), u0 = pow2f(e0), u1 = pow2f(e1);
    const int w0 = to_fp8x4(v[0] * d0, v[1] * d0, v[2] * d0, v[3] * d0);
    const int w1 = to_fp8x4(v[4] * d1, v[5] * d1, v[6] * d1, v[7] * d1);
    float q0[4], q1[4], k0[4], k1[4];
    fp8x4(w0, q0);
    fp8x4(w1, q1);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        k0[j] = v[j] - q0[j] * u0;
        k1[j] = v[4 + j] - q1[j] * u1;
    }
    const size_t i0 = size_t(s0) * a.ldb + b0;
    st4(a.carry + i0, k0);
    st4(a.carry + i0 + a.ldb, k1);
    const uint32_t chunk = (b0 / kQ8Bands) * 16u, byte = b0 % kQ8Bands;
    uint8_t* r0 = x.Bout8 + size_t(s0) * a.ldq + chunk;
    uint8_t* r1 = r0 + a.ldq;
    *reinterpret_cast<uint32_t*>(r0 + byte) = uint32_t(w0);
    *reinterpret_cast<uint32_t*>(r1 + byte) = uint32_t(w1);
    if (q == 0u) {
        *reinterpret_cast<uint32_t*>(r0 + 12) = uint32_t(e0) & 0xffu;
        *reinterpret_cast<uint32_t*>(r1 + 12) = uint32_t(e1) & 0xffu;
    }
}

__global__ __launch_bounds__(256) void k_sweep_q8_pair(SweepArgs a) {
    BandCtx x;
    if (!band_ctx<MODE_SWEEP>(a, false, x)) return;
    const uint32_t bt = a.bt, tpw = 64u / bt;
    const uint32_t wl = threadIdx.x & 63u, tiw = wl / bt;
    const uint32_t lane = wl - tiw * bt;
    const uint32_t fl = xcd_block(a.xcd) * (4u * tpw) + (threadIdx.x >> 6) * tpw + tiw;
    const bool live = tiw < tpw && fl < a.nUnits;
    float myAbs = 0.0f, myRel = 0.0f, myAct = 0.0f;
    uint32_t s0 = 0u, s1 = 0u;
    uint32_t rb[3] = {0u, 0u, 0u}, rc[3] = {0u, 0u, 0u};
    float4 Wf = make_float4(0.0f, 0.0f, 0.0f, 0.0f), Wg = Wf;
    if (live) {
        const uint2 un = a.units[fl];
        s0 = 2u * un.x;
        s1 = 2u * un.y;
        rc[0] = a.rowptr[s1];
        rc[1] = a.rowptr[s1 + 1];
        rc[2] = a.rowptr[s1 + 2];
        Wf = a.unitW[2u * fl];
        Wg = a.unitW[2u * fl + 1u];
        rb[0] = a.rowptr[s0];
        rb[1] = a.rowptr[s0 + 1];
        rb[2] = a.rowptr[s0 + 2];
    }
    const __amdgpu_buffer_rsrc_t rows = make_rsrc(x.Bin8, a.sides * a.ldq);
    // passes of 4 bt bands, a multiple of 12: chunks never straddle a pass
    for (uint32_t cb = 0; cb < x.nPrev; cb += 4u * bt) {
        const uint32_t b0 = cb + 4u * lane;
        const uint32_t laneBytes = (b0 / kQ8Bands) * 16u, wsel = (b0 % kQ8Bands) / 4u;
        float vf[8] = {}, vg[8] = {};
        bool emit = false;
        if (live && b0 < x.nPrev) {
            float gf[2][4] = {}, gg[2][4] = {};
            if (b0 < x.nAct) {
                const RowSums<GATHER_FP8> rf = gather_rows_pipe<GATHER_FP8>(a, rb, rows, a.ldq, laneBytes, wsel);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    gf[0][j] = rf.s0[j];
                    gf[1][j] = rf.s1[j];
                }
                const RowSums<GATHER_FP8> rg = gather_rows_pipe<GATHER_FP8>(a, rc, rows, a.ldq, laneBytes, wsel);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    gg[0][j] = rg.s0[j];
                    gg[1][j] = rg.s1[j];
                }
            }
            emit = pair_update4<2, true>(a, x, s0, s1, b0, Wf, Wg, gf[0], gf[1], gg[0], gg[1], myAbs, myRel, myAct,
                                         vf, vg);
        }
        emit_q8(a, x, s0, b0, lane, bt, emit, vf);
        emit_q8(a, x, s1, b0, lane, bt, emit, vg);
    }
    band_partials(a, myAbs, myRel, myAct);
}

// FP8 increment sweep of the single facets: a team of bt lanes per facet,
// each lane owns a whole 12-band chunk (one 16-byte gather per CSR entry),
// updates its bands 4 at a time and rounds the chunk with one exponent per
// side.  Fully retired 4-band groups keep their carry and accumulated
// gathers; their bytes are written as zeros, which nobody gathers (a
// group's gathered bands are only used while it is active).
__global__ __launch_bounds__(256) void k_sweep_q8_single(SweepArgs a) {
    BandCtx x;
    if (!band_ctx<MODE_SWEEP>(a, false, x)) return;
    const uint32_t bt = a.bt, tpb = 256u / bt;
    const uint32_t team = threadIdx.x / bt, lane = threadIdx.x - team * bt;
    const uint32_t fl = xcd_block(a.xcd) * tpb + team;
    float myAbs = 0.0f, myRel = 0.0f, myAct = 0.0f;
    if (team < tpb && fl < a.F) {
        const uint32_t s0 = 2u * (a.sched ? a.sched[fl] : fl);
        const uint32_t rb[3] = {a.rowptr[s0], a.rowptr[s0 + 1], a.rowptr[s0 + 2]};
        const __amdgpu_buffer_rsrc_t rows = make_rsrc(x.Bin8, a.sides * a.ldq);
        for (uint32_t cb = 0; cb < x.nPrev; cb += kQ8Bands * bt) {
            const uint32_t b0 = cb + kQ8Bands * lane;
            if (b0 >= x.nPrev) continue;
            float g0[12] = {}, g1[12] = {};
            if (b0 < x.nAct) {
                const RowSums<GATHER_FP8_CHUNK> rs =
                    gather_rows_pipe<GATHER_FP8_CHUNK>(a, rb, rows, a.ldq, (b0 / kQ8Bands) * 16u);
#pragma unroll
                for (int j = 0; j < 12; ++j) {
                    g0[j] = rs.s0[j];
                    g1[j] = rs.s1[j];
                }
            }
            float v0[12] = {}, v1[12] = {};
            bool em[3] = {false, false, false};
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                const uint32_t b = b0 + 4u * uint32_t(t);
                if (b >= a.ldb) continue;
                float h0[4], h1[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    h0[j] = g0[4 * t + j];
                    h1[j] = g1[4 * t + j];
                }
                Upd4 u;
                if (upd_begin<2>(a, x, s0, b, h0, h1, u)) {
                    upd_end<2, true>(a, x, s0, b, u, u.n0, u.n1, myAbs, myRel, myAct);
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        v0[4 * t + j] = u.k0[j];
                        v1[4 * t + j] = u.k1[j];
                    }
                    em[t] = true;
                }
            }
            if (!(em[0] || em[1] || em[2])) continue;
            float m0 = 0.0f, m1 = 0.0f;
#pragma unroll
            for (int j = 0; j < 12; ++j) {
                m0 = fmaxf(m0, fabsf(v0[j]));
                m1 = fmaxf(m1, fabsf(v1[j]));
            }
            const int e0 = chunk_exponent(m0), e1 = chunk_exponent(m1);
            const float d0 = pow2f(-e0), d1 = pow2f(-e1), u0 = pow2f(e0), u1 = pow2f(e1);
            uint32_t w0[4] = {0u, 0u, 0u, uint32_t(e0) & 0xffu}, w1[4] = {0u, 0u, 0u, uint32_t(e1) & 0xffu};
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                if (!em[t]) continue;
                const int q0 = to_fp8x4(v0[4 * t] * d0, v0[4 * t + 1] * d0, v0[4 * t + 2] * d0, v0[4 * t + 3] * d0);
                const int q1 = to_fp8x4(v1[4 * t] * d1, v1[4 * t + 1] * d1, v1[4 * t + 2] * d1, v1[4 * t + 3] * d1);
                w0[t] = uint32_t(q0);
                w1[t] = uint32_t(q1);
                float f0[4], f1[4], k0[4], k1[4];
                fp8x4(q0, f0);
                fp8x4(q1, f1);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    k0[j] = v0[4 * t + j] - f0[j] * u0;
                    k1[j] = v1[4 * t + j] - f1[j] * u1;
                }
                const size_t i0 = size_t(s0) * a.ldb + b0 + 4u * uint32_t(t);
                st4(a.carry + i0, k0);
                st4(a.carry + i0 + a.ldb, k1);
            }
            uint8_t* r0 = x.Bout8 + size_t(s0) * a.ldq + (b0 / kQ8Bands) * 16u;
            *reinterpret_cast<uint4*>(r0) = make_uint4(w0[0], w0[1], w0[2], w0[3]);
            *reinterpret_cast<uint4*>(r0 + a.ldq) = make_uint4(w1[0], w1[1], w1[2], w1[3]);
        }
    }
    band_partials(a, myAbs, myRel, myAct);
}

// Start of an FP8 increment solve: Q_0 = the FP8 rounding of B_0 (rows of
// ldq bytes), carry = B_0 - Q_0, accumulated gathers 0; one thread per
// 12-band chunk of a side.
__global__ void k_to_q8(const float* __restrict__ B, uint32_t S, uint32_t ldb, uint32_t ldq, uint8_t* __restrict__ Q,
                        float* __restrict__ carry, float* __restrict__ gacc) {
    const uint32_t nq = ldq / 16u;
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(S) * nq) return;
    const uint32_t s = uint32_t(i / nq), j = uint32_t(i - size_t(s) * nq);
    const uint32_t b0 = j * kQ8Bands;
    float v[kQ8Bands];
    float m = 0.0f;
#pragma unroll
    for (uint32_t k = 0; k < kQ8Bands; ++k) {
        v[k] = b0 + k < ldb ? B[size_t(s) * ldb + b0 + k] : 0.0f;
        m = fmaxf(m, fabsf(v[k]));
    }
    const int e = chunk_exponent(m);
    const float d = pow2f(-e), u = pow2f(e);
    uint32_t w[4] = {0u, 0u, 0u, uint32_t(e) & 0xffu};
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const int word = to_fp8x4(v[4 * t] * d, v[4 * t + 1] * d, v[4 * t + 2] * d, v[4 * t + 3] * d);
        w[t] = uint32_t(word);
        float qv[4];
        fp8x4(word, qv);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t b = b0 + uint32_t(4 * t + k);
            if (b < ldb) {
                carry[size_t(s) * ldb + b] = v[4 * t + k] - qv[k] * u;
                gacc[size_t(s) * ldb + b] = 0.0f;
            }
        }
    }
    *reinterpret_cast<uint4*>(Q + size_t(s) * ldq + size_t(j) * 16u) = make_uint4(w[0], w[1], w[2], w[3]);
}

// H = max(0, fsky S + G + D) from the accumulated gathers G of an increment
// solve (each band's G is that of its last sweep)
__global__ void k_incident_gathered(const float* __restrict__ G, const float* __restrict__ D,
                                    const float* __restrict__ fsky, const float* __restrict__ skyv, uint32_t S,
                                    uint32_t ldb, float* __restrict__ H) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(S) * ldb) return;
    const uint32_t s = uint32_t(i / ldb), b = uint32_t(i - size_t(s) * ldb);
    H[i] = fmaxf(fsky[s] * skyv[b] + G[i] + D[i], 0.0f);
}

// B -> BF16 copy with row stride ldh (bands >= ldb are zero)
__global__ void k_to_bf16(const float* __restrict__ B, uint32_t S, uint32_t ldb, uint32_t ldh,
                          uint16_t* __restrict__ H) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(S) * ldh) return;
    const size_t s = i / ldh, b = i - s * ldh;
    H[i] = b < ldb ? f32_to_bf16(B[s * ldb + b]) : uint16_t(0);
}

// Reduces the per-block residuals of the sweep just run and advances the
// control block: iteration count, last residuals, done flag.
__global__ __launch_bounds__(1024) void k_finalize(SolveCtrl* ctrl, const float2* __restrict__ partial,
                                                   const uint32_t* __restrict__ partAct, uint32_t n, uint32_t judge,
                                                   uint32_t refresh, uint32_t inner, const float* __restrict__ rho2,
                                                   float* __restrict__ omega, uint32_t ldb) {
    if (*reinterpret_cast<volatile uint32_t*>(&ctrl->done)) return;
    // Chebyshev weights for the next sweep (interval [-rho_b, rho_b]):
    // omega_1 = 1, omega_2 = 1 / (1 - rho^2 / 2), omega_k+1 = 1 / (1 - rho^2 omega_k / 4)
    if (omega)
        for (uint32_t b = threadIdx.x; b < ldb; b += 1024u) {
            const float r2 = rho2[b], w = omega[b];
            omega[b] = ctrl->iter == ctrl->chebStart ? 1.0f / (1.0f - 0.5f * r2) : 1.0f / (1.0f - 0.25f * r2 * w);
        }
    if (inner && *reinterpret_cast<volatile uint32_t*>(&ctrl->innerDone)) return;
    float mA = 0.0f, mR = 0.0f;
    uint32_t act = 0u;
    for (uint32_t i = threadIdx.x; i < n; i += 1024u) {
        const float2 p = partial[i];
        mA = fmaxf(mA, p.x);
        mR = fmaxf(mR, p.y);
        if (partAct) act = max(act, partAct[i]);
    }
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) {
        mA = fmaxf(mA, __shfl_xor(mA, m, 64));
        mR = fmaxf(mR, __shfl_xor(mR, m, 64));
        act = max(act, uint32_t(__shfl_xor(int(act), m, 64)));
    }
    __shared__ float sA[16], sR[16];
    __shared__ uint32_t sN[16];
    if ((threadIdx.x & 63) == 0) {
        sA[threadIdx.x >> 6] = mA;
        sR[threadIdx.x >> 6] = mR;
        sN[threadIdx.x >> 6] = act;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float absd = 0.0f, reld = 0.0f;
        uint32_t na = 0u;
        for (int w = 0; w < 16; ++w) {
            absd = fmaxf(absd, sA[w]);
            reld = fmaxf(reld, sR[w]);
            na = max(na, sN[w]);
        }
        const uint32_t it = ctrl->iter + 1u;
        ctrl->iter = it;
        ctrl->lastAbs = absd;
        ctrl->lastRel = reld;
        const uint32_t old = ctrl->nAct;
        ctrl->nActPrev = old;
        if (judge && partAct && ctrl->tolRel >= 0.0f) ctrl->nAct = min(old, na);
        if (refresh) {
            ctrl->refreshes += 1u;
            ctrl->innerDone = 0u;
        }
        if (inner && ctrl->innerTol >= 0.0f && reld <= ctrl->innerTol) ctrl->innerDone = 1u;
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

// Row stride of the [S][ldb] band arrays: 1 band, else a multiple of 4
// (16-byte vectors).  Padding wider rows to whole 128-byte lines measured
// slower on gfx950 (more row-stream work for the same gathered lines).
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
    if (avg < 192.0) return 16;  // 16 beat 32 at 130 entries per side (single band: 0.81 vs 0.87 ms)
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

// Block order of the sweeps.  Workgroups go round-robin over the 8 XCDs,
// each with its own 4 MB L2; with the remap each XCD works through one
// contiguous range of the (spatially ordered) facets, so the rows it holds
// in flight gather overlapping columns.  Measured on the terrain forest
// (468 M entries): band-parallel FP32 sweep 9.18 -> 8.14 ms (L2 hit 0.51 ->
// 0.74), increment sweep 5.69 -> 5.57 ms; the single-band sweep 1.44 ->
// 1.52 ms (its rows fit the L2 already, hit 0.91), so it keeps the plain
// order.  FH_XCD_REMAP=0|1 forces either order (tuning knob).
static uint32_t xcd_remap(bool bandParallel) {
    static const int forced = [] {
        const char* e = std::getenv("FH_XCD_REMAP");
        return e ? int(std::atoi(e) != 0) : -1;
    }();
    return forced >= 0 ? uint32_t(forced) : uint32_t(bandParallel);
}

bool uses_bf16_gather(uint32_t ldb) { return band_parallel(ldb); }

// BF16 increments save gathered lines only when an FP32 row spans more than
// one 128-byte line (urban block, 31 windows: FP32 0.337 s, BF16 increments
// 0.355 s for the 24-node run; terrain, 47 windows: 0.852 / 0.688 s)
bool bf16_pays(uint32_t ldb) { return band_parallel(ldb) && ldb > 32u; }

uint32_t bf16_stride(uint32_t ldb) { return (ldb + 63u) & ~63u; }

// Facets are processed in a spatial order built with the geometry (the BVH's
// leaf order): rows processed together see overlapping columns (-8% FP32
// shortwave sweep, -4% LW sweep against index order on the terrain forest;
// the order does not change any computed value).  FH_SCHED=0 processes them
// in index order (tuning knob).
static bool use_sched() {
    static const bool v = [] {
        const char* e = std::getenv("FH_SCHED");
        return e ? std::atoi(e) != 0 : true;
    }();
    return v;
}

// FP32-weight gathers read the packed CSR entries when the graph has them
// (FH_CSR_PACKED=0: column and FP32 weight arrays; tuning knob)
bool use_packed() {
    static const bool v = [] {
        const char* e = std::getenv("FH_CSR_PACKED");
        return e ? std::atoi(e) != 0 : true;
    }();
    return v;
}

// BF16 band-parallel solves gather BF16 increments (FH_INCREMENTS=0: a BF16
// phase, then FP32 sweeps; tuning knob)
static bool use_increments() {
    static const bool v = [] {
        const char* e = std::getenv("FH_INCREMENTS");
        return e ? std::atoi(e) != 0 : true;
    }();
    return v;
}

bool increment_solve(uint32_t ldb, uint32_t precision) {
    return band_parallel(ldb) &&
           ((precision == FH_PRECISION_BF16_DC && use_increments()) || precision == FH_PRECISION_FP8_INC);
}

uint32_t q8_stride(uint32_t ldb) { return (ldb + kQ8Bands - 1u) / kQ8Bands * 16u; }

const uint32_t* sweep_schedule(const fh_context* c) { return use_sched() && c->sched.p ? c->sched.p : nullptr; }

// ---- pair blocks ------------------------------------------------------------
// Leaves of a crown often come in close, facing pairs whose mutual exchange
// is the slowest mode of the Jacobi sweep (eigenvalues +-(r + t) F of the
// pair).  The band-parallel sweeps update two facets per team and solve the
// CSR entries between the two exactly (a block-Jacobi sweep with 4-side
// blocks: same fixed point, same Jacobi-residual stopping rule; measured on
// a 70 k-facet terrain forest at 47 bands: 8 -> 6 Chebyshev sweeps for the
// slowest band).  FH_PAIRS=0 keeps one facet per team (tuning knob).
// A facet is matched only with a partner that takes at least this share of
// one of its rows' weight (terrain forest, 1.0 M facets, 47 bands: 4 % of
// the facets matched, 138 -> 118 shortwave sweeps; all mutual best
// partners, 57 % of the facets: 116).
#ifndef FH_PAIR_MIN_WEIGHT
#define FH_PAIR_MIN_WEIGHT 0.2f
#endif
constexpr float kPairMinWeight = FH_PAIR_MIN_WEIGHT;

static bool use_pairs() {
    static const bool v = [] {
        const char* e = std::getenv("FH_PAIRS");
        return e ? std::atoi(e) != 0 : true;
    }();
    return v;
}


// Strongest unmatched partner of each unmatched facet: the facet whose two
// sides collect the largest weight sum in one of f's rows, if that sum is
// at least minStrength (entries of one column facet are adjacent in a
// sorted row; an unsorted row only splits the sums).  Ties go to the
// smaller index.
__global__ void k_pair_best(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ col,
                            const float* __restrict__ w, const uint32_t* __restrict__ pairOf, uint32_t F,
                            float minStrength, uint32_t* __restrict__ best) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    if (pairOf[f] != kNoFacet) {
        best[f] = kNoFacet;
        return;
    }
    float bw = 0.0f;
    uint32_t bh = kNoFacet;
    auto consider = [&](uint32_t h, float s) {
        if (h == kNoFacet || h == f || pairOf[h] != kNoFacet || s < minStrength) return;
        if (s > bw || (s == bw && s > 0.0f && h < bh)) {
            bw = s;
            bh = h;
        }
    };
    for (uint32_t s = 2u * f; s < 2u * f + 2u; ++s) {
        uint32_t cur = kNoFacet;
        float sum = 0.0f;
        for (uint32_t e = rowptr[s], e1 = rowptr[s + 1]; e < e1; ++e) {
            const uint32_t h = col[e] >> 1;
            if (h != cur) {
                consider(cur, sum);
                cur = h;
                sum = 0.0f;
            }
            sum += w[e];
        }
        consider(cur, sum);
    }
    best[f] = bh;
}

// mutual choices become pairs (symmetric: f -> h exactly when h -> f)
__global__ void k_pair_match(const uint32_t* __restrict__ best, uint32_t F, uint32_t* __restrict__ pairOf) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F || pairOf[f] != kNoFacet) return;
    const uint32_t h = best[f];
    if (h != kNoFacet && best[h] == f) pairOf[f] = h;
}

// weights of the entries of facet a's rows whose column is a side of facet
// b: (a0 -> b0, a0 -> b1, a1 -> b0, a1 -> b1), duplicates summed
__device__ float4 block_weights(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ col,
                                const float* __restrict__ w, uint32_t a, uint32_t b) {
    float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t q = 0; q < 2u; ++q) {
        const uint32_t s = 2u * a + q;
        for (uint32_t e = rowptr[s], e1 = rowptr[s + 1]; e < e1; ++e) {
            const uint32_t c = col[e];
            if ((c >> 1) == b) v[2u * q + (c & 1u)] += w[e];
        }
    }
    return make_float4(v[0], v[1], v[2], v[3]);
}

__global__ void k_unit_weights(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ col,
                               const float* __restrict__ w, const uint2* __restrict__ units, uint32_t U,
                               float4* __restrict__ unitW) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= U) return;
    const uint2 p = units[u];
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    unitW[2u * u] = p.y == kNoFacet ? z : block_weights(rowptr, col, w, p.x, p.y);
    unitW[2u * u + 1u] = p.y == kNoFacet ? z : block_weights(rowptr, col, w, p.y, p.x);
}

static void build_pair_units(fh_context* c) {
    c->units.release();
    c->unitW.release();
    c->singles.release();
    c->pairFacets.release();
    c->unitCount = 0u;
    c->singleCount = 0u;
    const uint32_t F = c->F;
    if (!use_pairs() || F < 2u) return;
    DBuf<uint32_t> pairOf, best;
    pairOf.ensure(F);
    best.ensure(F);
    launch_fill_u32(c, pairOf.p, F, kNoFacet);
    const uint32_t grid = (F + 255u) / 256u;
    for (int round = 0; round < 3; ++round) {  // later rounds match what earlier choices left
        hipLaunchKernelGGL(k_pair_best, dim3(grid), dim3(256), 0, c->stream, c->rowptr.p, c->col.p, c->val32.p,
                           pairOf.p, F, kPairMinWeight, best.p);
        hipLaunchKernelGGL(k_pair_match, dim3(grid), dim3(256), 0, c->stream, best.p, F, pairOf.p);
        FH_HIP(hipGetLastError());
    }
    std::vector<uint32_t> partner(F), order(F);
    FH_HIP(hipMemcpyAsync(partner.data(), pairOf.p, sizeof(uint32_t) * F, hipMemcpyDeviceToHost, c->stream));
    const uint32_t* sched = sweep_schedule(c);
    if (sched)
        FH_HIP(hipMemcpyAsync(order.data(), sched, sizeof(uint32_t) * F, hipMemcpyDeviceToHost, c->stream));
    FH_HIP(hipStreamSynchronize(c->stream));
    if (!sched)
        for (uint32_t f = 0; f < F; ++f) order[f] = f;
    // in sweep order: a matched facet and its partner form a unit (at the
    // position of the first of the two), every other facet stays single
    std::vector<uint2> units;
    std::vector<uint32_t> singles;
    singles.reserve(F);
    std::vector<uint8_t> placed(F, 0u);
    for (uint32_t i = 0; i < F; ++i) {
        const uint32_t f = order[i];
        if (placed[f]) continue;
        placed[f] = 1u;
        const uint32_t p = partner[f];
        if (p != kNoFacet && p < F && !placed[p] && partner[p] == f) {
            placed[p] = 1u;
            units.push_back(make_uint2(f, p));
        } else {
            singles.push_back(f);
        }
    }
    const uint32_t U = uint32_t(units.size()), nS = uint32_t(singles.size());
    if (U == 0u) return;
    c->units.ensure(U);
    c->unitW.ensure(2u * size_t(U));
    c->pairFacets.ensure(2u * size_t(U));
    FH_HIP(hipMemcpyAsync(c->pairFacets.p, units.data(), sizeof(uint2) * U, hipMemcpyHostToDevice, c->stream));
    c->singles.ensure(std::max(nS, 1u));
    FH_HIP(hipMemcpyAsync(c->units.p, units.data(), sizeof(uint2) * U, hipMemcpyHostToDevice, c->stream));
    if (nS)
        FH_HIP(hipMemcpyAsync(c->singles.p, singles.data(), sizeof(uint32_t) * nS, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_unit_weights, dim3((U + 255u) / 256u), dim3(256), 0, c->stream, c->rowptr.p, c->col.p,
                       c->val32.p, c->units.p, U, c->unitW.p);
    FH_HIP(hipGetLastError());
    FH_HIP(hipStreamSynchronize(c->stream));  // the host vectors back the copies
    if (!c->pairStream) {
        FH_HIP(hipStreamCreate(&c->pairStream));
        FH_HIP(hipEventCreateWithFlags(&c->pairFork, hipEventDisableTiming));
        FH_HIP(hipEventCreateWithFlags(&c->pairJoin, hipEventDisableTiming));
    }
    c->unitCount = U;
    c->singleCount = nS;
}

static bool use_staged() {  // read at every launch: tests compare both paths in one process
    const char* v = std::getenv("FH_STAGED");
    return !(v && std::atoi(v) == 0);
}

static void release_stage(fh_context* c) {
    c->stBlocks = 0u;
    c->stDcol.release();
    c->stDoff.release();
    c->stMoff.release();
    c->stMeta.release();
    c->stKey.release();
    c->stKey16.release();
}

__global__ void k_stage_chunks(const uint32_t* __restrict__ U, uint32_t nBlk, uint32_t* __restrict__ runs) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nBlk) runs[b] = (U[b] + kStageSlots - 1u) / kStageSlots * kStageFacets;
}

// The staged layout for the order of the sweeps' singles launch: the singles
// list when there are pair blocks, else the sweep schedule.  A block with
// more than kStageMaxChunks chunks of distinct columns, or a segment of more
// than 1020 entries, leaves the graph without a staged layout (the sweeps
// then gather through the vector L1).
static void build_stage(fh_context* c) {
    release_stage(c);
    const bool onSingles = c->unitCount > 0u;
    const uint32_t* order = onSingles ? c->singles.p : sweep_schedule(c);
    const uint32_t nF = onSingles ? c->singleCount : c->F;
    if (nF == 0u || c->nnz == 0u) return;
    static const bool attr = [] {
        FH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_stage_fill<false>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, int(stage_fill_lds_bytes())));
        FH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_stage_fill<true>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, int(stage_fill_lds_bytes())));
        return true;
    }();
    (void)attr;
    const uint32_t nBlk = (nF + kStageFacets - 1u) / kStageFacets;
    const uint32_t words = (c->S + 31u) / 32u;
    const uint32_t G = std::min(nBlk, 512u);
    DBuf<uint32_t> bitmap, rank, U, P, runs, eOff, fault;
    bitmap.ensure(size_t(G) * words);
    rank.ensure(size_t(G) * words);
    U.ensure(nBlk);
    P.ensure(nBlk);
    runs.ensure(nBlk);
    eOff.ensure(size_t(nBlk) + 1u);
    fault.ensure(1);
    launch_fill_u32(c, fault.p, 1, 0u);
    c->stDoff.ensure(size_t(nBlk) + 1u);
    c->stMoff.ensure(size_t(nBlk) + 1u);
    const size_t fillLds = stage_fill_lds_bytes();
    const auto batch = [&](uint32_t b0, bool write) {
        launch_fill_u32(c, bitmap.p, size_t(G) * words, 0u);
        hipLaunchKernelGGL(k_stage_bits, dim3(G), dim3(1024), 0, c->stream, c->rowptr.p, c->col.p, order, nF, nBlk, b0,
                           words, bitmap.p);
        hipLaunchKernelGGL(k_stage_rank, dim3(G), dim3(1024), 0, c->stream, bitmap.p, nBlk, b0, words, rank.p, U.p,
                           write ? static_cast<const uint32_t*>(c->stDoff.p) : nullptr, c->stDcol.p);
        if (write)
            hipLaunchKernelGGL(k_stage_fill<true>, dim3(G), dim3(1024), fillLds, c->stream, c->rowptr.p, c->col.p,
                               order, nF, nBlk, b0, words, bitmap.p, rank.p, U.p, P.p, eOff.p, c->stMoff.p,
                               c->stMeta.p, c->cnt.p, c->stKey.p, c->stKey16.p, fault.p);
        else
            hipLaunchKernelGGL(k_stage_fill<false>, dim3(G), dim3(1024), fillLds, c->stream, c->rowptr.p, c->col.p,
                               order, nF, nBlk, b0, words, bitmap.p, rank.p, U.p, P.p, eOff.p, c->stMoff.p,
                               c->stMeta.p, c->cnt.p, c->stKey.p, c->stKey16.p, fault.p);
        FH_HIP(hipGetLastError());
    };
    // sizes: distinct columns, padded entries and run descriptors per block
    for (uint32_t b0 = 0; b0 < nBlk; b0 += G) batch(b0, false);
    hipLaunchKernelGGL(k_stage_chunks, dim3((nBlk + 255u) / 256u), dim3(256), 0, c->stream, U.p, nBlk, runs.p);
    launch_scan(c, U.p, nBlk, c->stDoff.p);
    launch_scan(c, P.p, nBlk, eOff.p);
    launch_scan(c, runs.p, nBlk, c->stMoff.p);
    uint32_t tot[4] = {0u, 0u, 0u, 0u};
    FH_HIP(hipMemcpyAsync(&tot[0], c->stDoff.p + nBlk, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    FH_HIP(hipMemcpyAsync(&tot[1], eOff.p + nBlk, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    FH_HIP(hipMemcpyAsync(&tot[2], c->stMoff.p + nBlk, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    FH_HIP(hipMemcpyAsync(&tot[3], fault.p, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    FH_HIP(hipStreamSynchronize(c->stream));
    if (tot[3]) {
        release_stage(c);
        return;
    }
    c->stDcol.ensure(std::max<uint32_t>(tot[0], 1u));
    c->stKey.ensure(size_t(tot[1]) + 4u * kCsrPad);
    c->stKey16.ensure(size_t(tot[1]) + 4u * kCsrPad);
    c->stMeta.ensure(std::max<uint32_t>(tot[2], 1u));
    for (uint32_t b0 = 0; b0 < nBlk; b0 += G) batch(b0, true);
    FH_HIP(hipMemcpyAsync(&tot[3], fault.p, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    FH_HIP(hipStreamSynchronize(c->stream));
    if (tot[3]) {
        release_stage(c);
        return;
    }
    c->stBlocks = nBlk;
    c->stOnSingles = onSingles;
}

void build_pairs(fh_context* c) {
    c->incSerial = 0;  // a new graph: no increment state continues
    build_pair_units(c);
    build_stage(c);
}

// the staged singles launch of an increment sweep: kStageFacets teams of bt
// lanes per workgroup, the LDS of stage_lds_bytes(bt)
static void launch_staged(fh_context* c, const SweepArgs& a, uint32_t grid, hipStream_t st) {
    static const bool attr = [] {
        FH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sweep_staged<1>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, int(stage_lds_bytes(kStageMaxBands / 8u))));
        FH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sweep_staged<2>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, int(stage_lds_bytes(kStageMaxBands / 8u))));
        return true;
    }();
    (void)attr;
    StageArgs sa;
    sa.dcol = c->stDcol.p;
    sa.dOff = c->stDoff.p;
    sa.mOff = c->stMoff.p;
    sa.meta = c->stMeta.p;
    sa.key = reinterpret_cast<const uint4*>(c->stKey.p);
    sa.key16 = reinterpret_cast<const uint2*>(c->stKey16.p);
    sa.order = a.sched;
    sa.den = c->den.p;
    const dim3 block(kStageFacets * a.bt);
    if (a.inc == 1u) hipLaunchKernelGGL((k_sweep_staged<1>), dim3(grid), block, stage_lds_bytes(a.bt), st, a, sa);
    else hipLaunchKernelGGL((k_sweep_staged<2>), dim3(grid), block, stage_lds_bytes(a.bt), st, a, sa);
}

bool launch_gather1_staged(fh_context* c, const float* B, float L, const SolveCtrl* ctrl, float* Iin) {
    if (!c->stBlocks || !use_staged()) return false;
    StageArgs sa;
    sa.dcol = c->stDcol.p;
    sa.dOff = c->stDoff.p;
    sa.mOff = c->stMoff.p;
    sa.meta = c->stMeta.p;
    sa.key = reinterpret_cast<const uint4*>(c->stKey.p);
    sa.key16 = reinterpret_cast<const uint2*>(c->stKey16.p);
    sa.order = c->stOnSingles ? c->singles.p : sweep_schedule(c);
    sa.den = c->den.p;
    const uint32_t nF = c->stOnSingles ? c->singleCount : c->F;
    hipLaunchKernelGGL(k_gather1_staged, dim3(c->stBlocks), dim3(kStageFacets), 0, c->stream, sa, B, c->fsky.p, L, nF,
                       ctrl, Iin);
    FH_HIP(hipGetLastError());
    return true;
}

void launch_sweep(fh_context* c, int mode, bool w16, SweepArgs a) {
    a.xcd = xcd_remap(band_parallel(a.ldb));
    a.sched = sweep_schedule(c);
    a.pk = c->pkBits && use_packed() ? c->pk.p : nullptr;
    a.den = c->den.p;
    a.pkBits = c->pkBits;
    a.sides = c->S;
    if (band_parallel(a.ldb)) {
        // the band-parallel gathers address the gathered rows by 32-bit byte
        // offsets below kNoRow
        if (uint64_t(c->S) * a.ldb * 4u >= uint64_t(kNoRow))
            fail(FH_RUNTIME_ERROR, "radiosity array of " + std::to_string(uint64_t(c->S) * a.ldb * 4u) +
                                       " bytes exceeds the 4 GiB the band-parallel sweeps address");
        // lanes per team: one 16-byte vector of the row each, at most 64
        // bands per team pass
        const bool half = (mode == MODE_SWEEP && w16 && a.h0) || a.inc;
        const bool q8 = a.inc && a.q8;
        const uint32_t vw = half && !q8 ? 8u : 4u;
        a.bt = std::min((a.ldb + vw - 1u) / vw, 64u / vw);
        // FP8 increment sweeps: the single facets one 12-band chunk per lane;
        // the pairs 4 bands per lane, a multiple of 3 lanes (the lanes of one
        // chunk), teams inside a wave
        if (q8) a.bt = std::min((a.ldb + kQ8Bands - 1u) / kQ8Bands, 16u);
        const uint32_t perBlock = 256u / a.bt;
        a.F = c->F;
        a.rowptr = c->rowptr.p;
        a.col = c->col.p;
        a.w32 = c->val32.p;
        a.w16 = c->val16.p;
        a.fsky = c->fsky.p;
        a.ctrl = c->ctrl.p;
        // pair blocks (FP32-weight and increment sweeps): the single facets
        // in one launch, the coupled pairs in a second one (4 bands per
        // lane) on the pair stream beside it (fork / join by events), each
        // writing its own range of the per-block partials
        const bool pair =
            c->unitCount && use_pairs() && !a.jacobiOnly && (a.inc || (mode == MODE_SWEEP && !w16 && !half));
        // staged gathers for the single facets of increment sweeps, where the
        // staged layout was built for this launch's order
        const bool staged = a.inc && !q8 && c->stBlocks && use_staged() && a.ldb <= kStageMaxBands &&
                            c->stOnSingles == pair;
        SweepArgs p = a;
        uint32_t g = std::max(1u, (c->F + perBlock - 1u) / perBlock), gp = 0u;
        if (staged) g = c->stBlocks;
        if (pair) {
            a.sched = c->singles.p;
            a.F = c->singleCount;
            g = staged ? c->stBlocks : (a.F + perBlock - 1u) / perBlock;
            p.bt = q8 ? std::min((p.ldb + 3u) / 4u + 2u, 15u) / 3u * 3u : std::min((p.ldb + 3u) / 4u, 16u);
            p.units = c->units.p;
            p.unitW = c->unitW.p;
            p.nUnits = c->unitCount;
            const uint32_t unitsPerBlock = q8 ? 4u * (64u / p.bt) : 256u / p.bt;
            gp = (p.nUnits + unitsPerBlock - 1u) / unitsPerBlock;
            p.partOff = g;
        }
        const uint32_t parts = std::max(1u, g + gp);
        if (a.partialExt) {
            if (parts > a.partialCap)
                fail(FH_INVALID_ARGUMENT, "sweep needs " + std::to_string(parts) + " block partials; the fixed scratch holds " +
                                              std::to_string(a.partialCap));
            a.partial = p.partial = a.partialExt;
        } else {
            c->partial.ensure(parts);
            a.partial = p.partial = c->partial.p;
        }
        if (a.retire) {
            c->partAct.ensure(parts);
            a.partAct = p.partAct = c->partAct.p;
        } else {
            a.partAct = p.partAct = nullptr;
        }
        hipStream_t st = c->stream;
        if (gp) {
            hipStream_t ps = c->pairStream;
            FH_HIP(hipEventRecord(c->pairFork, st));
            FH_HIP(hipStreamWaitEvent(ps, c->pairFork, 0));
            if (q8) hipLaunchKernelGGL(k_sweep_q8_pair, dim3(gp), dim3(256), 0, ps, p);
            else if (p.inc == 1u) hipLaunchKernelGGL((k_sweep_hv<1, true>), dim3(gp), dim3(256), 0, ps, p);
            else if (p.inc) hipLaunchKernelGGL((k_sweep_hv<2, true>), dim3(gp), dim3(256), 0, ps, p);
            else hipLaunchKernelGGL((k_sweep_v<MODE_SWEEP, false, true>), dim3(gp), dim3(256), 0, ps, p);
            FH_HIP(hipEventRecord(c->pairJoin, ps));
        }
        if (g) {
            if (q8) hipLaunchKernelGGL(k_sweep_q8_single, dim3(g), dim3(256), 0, st, a);
            else if (staged) launch_staged(c, a, g, st);
            else if (a.inc == 1u) hipLaunchKernelGGL((k_sweep_hv<1, false>), dim3(g), dim3(256), 0, st, a);
            else if (a.inc) hipLaunchKernelGGL((k_sweep_hv<2, false>), dim3(g), dim3(256), 0, st, a);
            else if (half) hipLaunchKernelGGL((k_sweep_hv<0, false>), dim3(g), dim3(256), 0, st, a);
            else if (mode == MODE_SWEEP && !w16)
                hipLaunchKernelGGL((k_sweep_v<MODE_SWEEP, false, false>), dim3(g), dim3(256), 0, st, a);
            else if (mode == MODE_SWEEP)
                hipLaunchKernelGGL((k_sweep_v<MODE_SWEEP, true, false>), dim3(g), dim3(256), 0, st, a);
            else if (mode == MODE_REFRESH) fail(FH_RUNTIME_ERROR, "refresh sweeps run on the entry-parallel layouts only");
            else hipLaunchKernelGGL((k_sweep_v<MODE_INCIDENT, false, false>), dim3(g), dim3(256), 0, st, a);
        }
        if (gp) FH_HIP(hipStreamWaitEvent(st, c->pairJoin, 0));
        if (mode != MODE_INCIDENT && !a.standalone)
            hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, c->stream, c->ctrl.p, c->partial.p, a.partAct,
                               g + gp, a.judge, uint32_t(mode == MODE_REFRESH),
                               uint32_t(mode == MODE_SWEEP && w16 && !a.inc), a.rho2, const_cast<float*>(a.omega),
                               a.ldb);
        FH_HIP(hipGetLastError());
        return;
    }
    const int nbc = pick_nbc(a.ldb);
    const int team = sweep_team(c);
    a.retire = 0u;
    a.partAct = nullptr;
    a.omega = a.rho2 = nullptr;
    a.nchunks = (a.ldb + uint32_t(nbc) - 1u) / uint32_t(nbc);
    a.F = c->F;
    a.rowptr = c->rowptr.p;
    a.col = c->col.p;
    a.w32 = c->val32.p;
    a.w16 = c->val16.p;
    a.fsky = c->fsky.p;
    a.ctrl = c->ctrl.p;
    // single-band FP32 sweeps of a solve over the staged layout: its facets
    // by k_sweep1_staged, the coupled pairs' facets (outside the layout when
    // it was built on the single facets) by the team kernel beside them,
    // each writing its own range of the per-block partials
    if (a.ldb == 1u && mode == MODE_SWEEP && !w16 && c->stBlocks && use_staged() && !a.standalone &&
        !a.partialExt && a.buf0) {
        StageArgs sa;
        sa.dcol = c->stDcol.p;
        sa.dOff = c->stDoff.p;
        sa.mOff = c->stMoff.p;
        sa.meta = c->stMeta.p;
        sa.key = reinterpret_cast<const uint4*>(c->stKey.p);
        sa.key16 = reinterpret_cast<const uint2*>(c->stKey16.p);
        sa.order = c->stOnSingles ? c->singles.p : sweep_schedule(c);
        sa.den = c->den.p;
        const uint32_t g = c->stBlocks;
        SweepArgs p = a;
        uint32_t gp = 0u;
        if (c->stOnSingles && c->unitCount) {
            p.sched = c->pairFacets.p;
            p.F = 2u * c->unitCount;
            p.xcd = 0u;
            gp = uint32_t((uint64_t(p.F) * uint64_t(team) + 255u) / 256u);
            p.partOff = g;
        }
        a.F = c->stOnSingles ? c->singleCount : c->F;
        c->partial.ensure(size_t(g) + gp);
        a.partial = p.partial = c->partial.p;
        hipLaunchKernelGGL(k_sweep1_staged, dim3(g), dim3(kStageFacets), 0, c->stream, a, sa);
        if (gp) launch_team<1>(team, mode, w16, p, c->stream);
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, c->stream, c->ctrl.p, c->partial.p,
                           static_cast<const uint32_t*>(nullptr), g + gp, a.judge, 0u, 0u,
                           static_cast<const float*>(nullptr), static_cast<float*>(nullptr), a.ldb);
        FH_HIP(hipGetLastError());
        return;
    }
    const uint64_t threads = uint64_t(c->F) * uint64_t(team);
    const uint32_t grid = uint32_t((threads + 255u) / 256u);
    if (a.partialExt) {
        if (std::max(grid, 1u) > a.partialCap)
            fail(FH_INVALID_ARGUMENT, "sweep needs " + std::to_string(grid) + " block partials; the fixed scratch holds " +
                                          std::to_string(a.partialCap));
        a.partial = a.partialExt;
    } else {
        c->partial.ensure(grid ? grid : 1u);
        a.partial = c->partial.p;
    }
    switch (nbc) {
        case 1: launch_team<1>(team, mode, w16, a, c->stream); break;
        case 4: launch_team<4>(team, mode, w16, a, c->stream); break;
        case 8: launch_team<8>(team, mode, w16, a, c->stream); break;
        default: launch_team<16>(team, mode, w16, a, c->stream); break;
    }
    if (mode != MODE_INCIDENT && !a.standalone)
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, c->stream, c->ctrl.p, c->partial.p,
                           static_cast<const uint32_t*>(nullptr), grid ? grid : 1u, a.judge,
                           uint32_t(mode == MODE_REFRESH), uint32_t(mode == MODE_SWEEP && w16),
                           static_cast<const float*>(nullptr), static_cast<float*>(nullptr), a.ldb);
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


__global__ void k_ctrl_reset(SolveCtrl* ctrl, uint32_t maxIter, float tolRel, float tolAbs, uint32_t checkEvery,
                             uint32_t nAct, float innerTol) {
    if (threadIdx.x == 0) {
        SolveCtrl z{};
        z.maxIter = maxIter;
        z.tolRel = tolRel;
        z.tolAbs = tolAbs;
        z.checkEvery = checkEvery ? checkEvery : 1u;
        z.nAct = nAct;
        z.nActPrev = nAct;
        z.innerTol = innerTol;
        *ctrl = z;
    }
}

__global__ void k_ctrl_inner(SolveCtrl* ctrl, float innerTol) {
    if (threadIdx.x == 0) ctrl->innerTol = innerTol;
}

// restarts the Chebyshev sequence at the next sweep (omega = 1)
__global__ void k_cheb_restart(SolveCtrl* ctrl, float* omega, uint32_t ldb) {
    for (uint32_t b = threadIdx.x; b < ldb; b += blockDim.x) omega[b] = 1.0f;
    if (threadIdx.x == 0) ctrl->chebStart = ctrl->iter;
}

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_f32(float* p, size_t n, float v) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void launch_ctrl_reset(fh_context* c, uint32_t maxIter, float tolRel, float tolAbs, uint32_t checkEvery,
                       uint32_t nAct, float innerTol) {
    hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(64), 0, c->stream, c->ctrl.p, maxIter, tolRel, tolAbs, checkEvery,
                       nAct, innerTol);
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

static float env_float(const char* name, float dflt) {
    const char* e = std::getenv(name);
    return e ? float(std::atof(e)) : dflt;
}

int run_jacobi(fh_context* c, const JacobiOptions& o, const float* cc, const float* rr, const float* tt,
               float* bufs[2], fh_solve_stats* stats, bool* incidentDone) {
    const bool bp = band_parallel(o.ldb);
    // FP8 increments on band-parallel strides (below them the solve is FP32)
    const bool q8 = o.precision == FH_PRECISION_FP8_INC && bp;
    const bool bf16 = o.precision == FH_PRECISION_BF16_DC || o.precision == FH_PRECISION_BF16_F || q8;
    const bool retire = o.retire && bp && o.tolRel >= 0.0f;
    const bool half = bf16 && bp;
    // FH_DC_FIRST overrides the end of the BF16 phase (tuning knob)
    const float firstInner = env_float("FH_DC_FIRST", o.firstInner);
    // the accumulated gathers of increment solves are a running FP32 sum, so
    // their last-ulp changes do not settle: such solves resolve relative
    // changes down to kIncrementTolFloor (2.5 FP32 ulps)
    constexpr float kIncrementTolFloor = 3.0e-7f;
    const bool incr = half && (q8 || (o.precision == FH_PRECISION_BF16_DC && use_increments()));
    // the FP32 sweeps that follow a BF16 phase settle to within a few ulps
    // of the FP32 fixed point, but their last-bit changes need not die out
    // (urban block, 31 windows: the change stays near 2e-7 for 1000 sweeps
    // at a 1e-7 tolerance), so they take the same floor
    const bool floored = incr || (half && o.precision == FH_PRECISION_BF16_F);
    const float tolRel = floored && o.tolRel >= 0.0f ? std::max(o.tolRel, kIncrementTolFloor) : o.tolRel;
    launch_ctrl_reset(c, o.maxIter ? o.maxIter : 1u, tolRel, o.tolAbs, o.checkEvery, o.ldb,
                      half ? std::max(firstInner, o.tolRel) : o.tolRel);
    SweepArgs a{};
    a.r = rr;
    a.t = tt;
    a.ldb = o.ldb;
    a.relax = o.relax;
    a.retire = retire ? 1u : 0u;
    a.buf0 = bufs[0];
    a.buf1 = bufs[1];
    // Chebyshev acceleration of the band-parallel solves (FP32, and the BF16
    // phase of BF16 solves)
    if (bp && o.rho2) {
        c->omega.ensure(o.ldb);
        launch_fill_f32(c, c->omega.p, o.ldb, 1.0f);
        a.omega = c->omega.p;
        a.rho2 = o.rho2;
    }
    // incident fused into the judge sweeps: every band's H is that of its
    // last judge sweep, complete once every band converged (retirement)
    if (bp && o.fmat) {
        a.fmat = o.fmat;
        a.rtab = o.rtab;
        a.ttab = o.ttab;
        a.dirf = o.dirf;
        a.dniv = o.dniv;
        a.skyv = c->skyv.p;
    }
    // (increment solves hold the gathers of every band's last sweep in gacc
    // and derive H from them once after the solve)
    // (entry-parallel FP32 solves: every sweep judges, so the last one run
    // leaves H of its gather)
    const bool fuse = o.H && o.D && !incr && (retire || (!bp && !bf16));
    if (fuse) {
        a.fuseH = 1u;
        a.H = o.H;
        a.D = o.D;
        a.skyv = c->skyv.p;
    }
    uint64_t launched = 0;
    uint32_t batch = 4;
    if (bf16 && !half) c->c2.ensure(size_t(c->S) * o.ldb);
    // continue the increment state of the solve before (see below)?
    // FH_SW_CONTINUE=0: always a base sweep (read per solve)
    const char* contEnv = std::getenv("FH_SW_CONTINUE");
    const bool contOn = !(contEnv && std::atoi(contEnv) == 0);
    const bool cont = contOn && incr && !q8 && o.incSerial && *o.incSerial && *o.incSerial == c->incSerial &&
                      c->incLdb == o.ldb;
    uint64_t serial = 0;
    if (incr) {
        // every sweep gathers low-precision increments: BF16 -- the base
        // sweep Q_0 = bf16(B_0), then the BF16 increments of B with their
        // rounding carried to the next one (band_update_inc4) -- or FP8 --
        // Q_0 = the FP8 rounding of B_0 with its carry and zero accumulated
        // gathers set up first, then FP8 increment sweeps (k_sweep_q8_*); FP32
        // weights and accumulation, FP32 fixed point
        const size_t nf = size_t(c->S) * o.ldb;
        c->gacc.ensure(nf);
        c->carry.ensure(nf);
        if (q8) {
            const uint32_t ldq = q8_stride(o.ldb);
            const size_t nq = size_t(c->S) * ldq;
            c->X8a.ensure(nq);
            c->X8b.ensure(nq);
            const size_t chunks = size_t(c->S) * (ldq / 16u);
            hipLaunchKernelGGL(k_to_q8, dim3(uint32_t((chunks + 255) / 256)), dim3(256), 0, c->stream, bufs[0], c->S,
                               o.ldb, ldq, c->X8a.p, c->carry.p, c->gacc.p);
            FH_HIP(hipGetLastError());
            a.q0 = c->X8a.p;
            a.q1 = c->X8b.p;
            a.ldq = ldq;
            a.q8 = 1u;
        } else {
            const uint32_t ldh = bf16_stride(o.ldb);
            const size_t nh = size_t(c->S) * ldh;
            c->X16a.ensure(nh);
            c->X16b.ensure(nh);
            if (!cont) {
                hipLaunchKernelGGL(k_to_bf16, dim3(uint32_t((nh + 255) / 256)), dim3(256), 0, c->stream, bufs[0],
                                   c->S, o.ldb, ldh, c->X16a.p);
                FH_HIP(hipGetLastError());
            }
            a.h0 = c->X16a.p;
            a.h1 = c->X16b.p;
            a.ldh = ldh;
        }
        a.gacc = c->gacc.p;
        a.carry = c->carry.p;
        a.c = cc;
        a.judge = 1u;
        // the first sweep: the base sweep gathering Q_0 = bf16(B_0), or,
        // continuing the increment state of the solve before (B_0 its final
        // B, the same operator: gacc = M (B_0 - carry) with the increments it
        // wrote last pending in X16a), a sweep without gathers that adds the
        // pending increments to the carries
        while (true) {
            for (uint32_t k = 0; k < batch; ++k) {
                a.inc = launched == 0 && !q8 && !cont ? 1u : 2u;
                a.fold = launched == 0 && cont ? 1u : 0u;
                launch_sweep(c, MODE_SWEEP, false, a);
                ++launched;
            }
            read_ctrl(c);
            if (c->hctrl->done) break;
            if (batch < 32u) batch *= 2u;
        }
        if (!q8) {  // the increments written last in X16a, the state this solve's serial
            if (c->hctrl->iter & 1u) c->X16a.swap(c->X16b);
            c->incSerial = ++c->solveSerial;
            c->incLdb = o.ldb;
            serial = c->incSerial;
        } else {
            c->incSerial = 0;
        }
        if (o.H && o.D) {
            const size_t n = size_t(c->S) * o.ldb;
            hipLaunchKernelGGL(k_incident_gathered, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, c->stream,
                               c->gacc.p, o.D, c->fsky.p, c->skyv.p, c->S, o.ldb, o.H);
            FH_HIP(hipGetLastError());
        }
    } else if (half) {
        c->incSerial = 0;  // X16a / X16b overwritten
        // Phase 1: BF16 sweeps gathering BF16 copies of B (X16a = bf16(bufs[0])
        // first), ending at relative change firstInner (about the distance of
        // the BF16 operator's fixed point) or after innerBf16 sweeps.
        // Phase 2: FP32 sweeps from there, which judge convergence per band
        // and remove the BF16 defect (the FP32 operator's fixed point).
        const uint32_t ldh = bf16_stride(o.ldb);
        const size_t nh = size_t(c->S) * ldh;
        c->X16a.ensure(nh);
        c->X16b.ensure(nh);
        hipLaunchKernelGGL(k_to_bf16, dim3(uint32_t((nh + 255) / 256)), dim3(256), 0, c->stream, bufs[0], c->S, o.ldb,
                           ldh, c->X16a.p);
        FH_HIP(hipGetLastError());
        a.h0 = c->X16a.p;
        a.h1 = c->X16b.p;
        a.ldh = ldh;
        a.c = cc;
        a.judge = 0u;
        for (uint32_t k = 0; k < o.innerBf16; ++k) launch_sweep(c, MODE_SWEEP, true, a);
        a.h0 = a.h1 = nullptr;
        if (a.omega) {
            hipLaunchKernelGGL(k_cheb_restart, dim3(1), dim3(64), 0, c->stream, c->ctrl.p, c->omega.p, o.ldb);
            FH_HIP(hipGetLastError());
        }
        a.judge = 1u;
        while (true) {
            for (uint32_t k = 0; k < batch; ++k) launch_sweep(c, MODE_SWEEP, false, a);
            read_ctrl(c);
            if (c->hctrl->done) break;
            if (batch < 32u) batch *= 2u;
        }
    } else {
        while (true) {
            for (uint32_t k = 0; k < batch; ++k) {
                if (!bf16) {
                    a.c = cc;
                    a.judge = 1u;
                    launch_sweep(c, MODE_SWEEP, false, a);
                } else {
                    // groups of BF16-weight sweeps (stopping early once their
                    // relative change is within tolRel) closed by an FP32
                    // refresh that judges convergence and recomputes the
                    // defect correction; the first group runs on the
                    // uncorrected source
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
    }
    if (stats) {
        stats->iterations = c->hctrl->iter;
        stats->maxDelta = c->hctrl->lastAbs;
        stats->maxRelDelta = c->hctrl->lastRel;
        stats->refreshes = c->hctrl->refreshes;
    }
    if (o.incSerial) *o.incSerial = serial;
    if (incidentDone)
        *incidentDone = (fuse && (!retire || c->hctrl->nAct == 0u)) || (incr && o.H && o.D);
    return int(c->hctrl->iter & 1u);
}

// per-band maximum of a non-negative [S][ldb] array (float bits compared as
// unsigned integers) into out[ldb] (zeroed by the caller), ldb <= 4096
__global__ __launch_bounds__(256) void k_band_max(const float* __restrict__ v, uint32_t S, uint32_t ldb,
                                                  uint32_t* __restrict__ out) {
    __shared__ uint32_t sh[4096];
    for (uint32_t b = threadIdx.x; b < ldb; b += 256u) sh[b] = 0u;
    __syncthreads();
    const size_t n = size_t(S) * ldb;
    for (size_t i = size_t(blockIdx.x) * 256u + threadIdx.x; i < n; i += size_t(gridDim.x) * 256u)
        atomicMax(&sh[i % ldb], __float_as_uint(fmaxf(v[i], 0.0f)));
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < ldb; b += 256u)
        if (sh[b]) atomicMax(&out[b], sh[b]);
}

void estimate_rho2(fh_context* c, uint32_t ldb, uint32_t nb, const float* rr, const float* tt, float* v0, float* v1,
                   float* zero, float* rho2) {
    if (ldb > 4096u) fail(FH_INVALID_ARGUMENT, "too many bands for the spectral estimate");
    const size_t n = size_t(c->S) * ldb;
    launch_fill_f32(c, v0, n, 1.0f);
    launch_fill_f32(c, zero, n, 0.0f);
    c->bandMax.ensure(2u * ldb);
    launch_fill_u32(c, c->bandMax.p, 2u * ldb, 0u);
    // B_k+1 = K B_k (c = 0); maxima after m and m + 2 sweeps (both land in
    // v0); FH_RHO_SWEEPS overrides m (tuning knob)
    const int m = std::max(2, int(env_float("FH_RHO_SWEEPS", 8.0f))) & ~1;
    for (int k = 0; k < m + 2; ++k) {
        launch_single_sweep(c, ldb, zero, rr, tt, (k & 1) ? v1 : v0, (k & 1) ? v0 : v1, false, nullptr, nullptr,
                            nullptr, nullptr, true);  // the operator of the solves: pair blocks
        if (k == m - 1 || k == m + 1) {
            hipLaunchKernelGGL(k_band_max, dim3(512), dim3(256), 0, c->stream, v0, c->S, ldb,
                               c->bandMax.p + (k == m - 1 ? 0u : ldb));
            FH_HIP(hipGetLastError());
        }
    }
    std::vector<uint32_t> h(2u * ldb);
    FH_HIP(hipMemcpyAsync(h.data(), c->bandMax.p, sizeof(uint32_t) * h.size(), hipMemcpyDeviceToHost, c->stream));
    FH_HIP(hipStreamSynchronize(c->stream));
    std::vector<float> r2(ldb, 0.0f);
    for (uint32_t b = 0; b < ldb && b < nb; ++b) {
        float a, z;
        std::memcpy(&a, &h[b], 4);
        std::memcpy(&z, &h[ldb + b], 4);
        if (a > 0.0f && z > 0.0f && std::isfinite(a) && std::isfinite(z)) {
            // two-step ratio (eigenvalues come in +-lambda pairs), 2 % margin
            const float rho = std::min(0.98f, env_float("FH_RHO_MARGIN", 1.02f) * std::sqrt(z / a));
            r2[b] = rho * rho;
        }
    }
    FH_HIP(hipMemcpyAsync(rho2, r2.data(), sizeof(float) * ldb, hipMemcpyHostToDevice, c->stream));
    FH_HIP(hipStreamSynchronize(c->stream));
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
                         const float* Bin, float* Bout, bool w16, uint16_t* h16in, uint16_t* h16out, float* gacc,
                         float* carry, bool pairs, float2* partials, uint32_t partialCap,
                         const SweepTables* tables, uint8_t* q8in, uint8_t* q8out) {
    SweepArgs a{};
    a.partialExt = partials;
    a.partialCap = partialCap;
    if (q8in && q8out && gacc && carry && band_parallel(ldb)) {
        a.q8 = 1u;
        a.q0 = q8in;
        a.q1 = q8out;
        a.ldq = q8_stride(ldb);
        w16 = true;  // an increment sweep below
        h16in = h16out = nullptr;
    }
    if (tables && band_parallel(ldb)) {
        a.fmat = tables->fmat;
        a.rtab = tables->rtab;
        a.ttab = tables->ttab;
        a.dirf = tables->dirf;
        a.dniv = tables->dniv;
        a.skyv = c->skyv.p;
        a.omega = tables->omega;
    }
    a.jacobiOnly = pairs ? 0u : 1u;
    if (w16 && (h16in || a.q8) && gacc && carry && band_parallel(ldb)) {
        a.inc = 2u;
        a.gacc = gacc;
        a.carry = carry;
        w16 = false;
    }
    a.h0 = h16in;
    a.h1 = h16out;
    a.ldh = bf16_stride(ldb);
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
