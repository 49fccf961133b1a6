// eb.hip -- FacetEB on the device: one meteorological node of the coupled
// shortwave / longwave radiosity and surface energy balance.
//
// Follows runFacetEBCore (models/histream/src/facetrt/facetrtcore.cpp,
// "core"):
//   shortwave windows -> optics, solve, incident, absorbed   core:3495-3565
//   energy-balance constants (h_a, rho_a, r_a)              core:3567-3578
//   fluxes per side (leaf facets share side 0's T)          core:3592-3720
//   residual / thermal response                             core:3726-3743
//   relaxed Newton update, clamp [220, 360] K               core:3758-3827
//   convergence: max dT <= tol and max res/resp <= tol      core:3829-3831
//   final LW solve + residual, converged flag               core:3834-3846
// with the band loop fused into the multi-band SpMM, every optics array
// built on the device from material tables, and the Newton step and the
// scene-level energy sums in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cmath>
#include <cstdlib>

#include "hapke.hpp"
#include "solver.hpp"
#include "sweep_common.hpp"

namespace fh {

constexpr double kSigma = 5.670374419e-8;
constexpr double kPiD = 3.14159265358979323846;

struct EbConst {
    double Ta, ea, P, ha, rho, ra, dTime, relax;
    uint32_t vegMethod;
    uint32_t leafCount;
    uint32_t adaptive;  // 1: full Newton steps, relax only where a step reverses direction
};

// Relaxation of one Newton step: the configured one, or with adaptive
// relaxation 1 unless the unit's full step changed sign since its previous
// step (an oscillation), which then takes the configured relaxation; prev
// holds the unit's previous full step.
__device__ __forceinline__ double step_relaxation(const EbConst& k, float* prev, double step) {
    if (!k.adaptive || !prev) return k.relax;
    const double w = double(*prev) * step < 0.0 ? k.relax : 1.0;
    *prev = float(step);
    return w;
}

__device__ __forceinline__ double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

// core:854-867
__device__ __forceinline__ double sat_vp(double T) {
    const double c = clampd(T - 273.15, -80.0, 80.0);
    return 6.107 * pow(10.0, 7.5 * c / (237.3 + c));
}
__device__ __forceinline__ double sat_vp_slope(double T) {
    const double c = clampd(T - 273.15, -80.0, 80.0);
    return sat_vp(T) * 2.3026 * 7.5 * 237.3 / ((237.3 + c) * (237.3 + c));
}
// core:869-905
__device__ __forceinline__ void latent(double T, const EbConst& k, double rs, double& le, double& dle) {
    const double c = T - 273.15;
    const double lam = (2.501 - 0.002361 * c) * 1.0e6;
    const double Pm = fmax(100.0, k.P);
    const double deficit = 0.622 * fmax(0.0, sat_vp(T) - k.ea) / Pm;
    const double inv = 1.0 / fmax(1.0, k.ra + rs);
    le = k.rho * lam * deficit * inv;
    dle = fmax(0.0, k.rho * inv * (-2361.0 * deficit + lam * 0.622 * sat_vp_slope(T) / Pm));
}
// core:907-915
__device__ __forceinline__ double soil_storage(const fh_eb_material& m, double dT) {
    const double inertia = sqrt(fmax(0.0, double(m.specificHeat) * m.density * m.thermalConductivity));
    return 2.0 * inertia / sqrt(kPiD * fmax(1.0, dT));
}

// facet areas (FP64), for the scene-level energy sums
__global__ void k_areas(const float* __restrict__ fv, uint32_t F, double* __restrict__ area) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const float* t = fv + 9 * size_t(f);
    const double ux = double(t[3]) - t[0], uy = double(t[4]) - t[1], uz = double(t[5]) - t[2];
    const double vx = double(t[6]) - t[0], vy = double(t[7]) - t[1], vz = double(t[8]) - t[2];
    const double cx = uy * vz - uz * vy, cy = uz * vx - ux * vz, cz = ux * vy - uy * vx;
    area[f] = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
}

// per-side LW optics constants (makeEnergyLongwaveOptics, core:1005-1036,
// angular strength 0)
// ground sides (facets from leafCount on) take the Hapke emissivity at
// |n_y| when strength > 0 (core:1025-1027)
__global__ void k_lw_setup(uint32_t S, const uint32_t* __restrict__ mat, const fh_eb_material* __restrict__ mats,
                           uint32_t leafCount, const float4* __restrict__ nrm, float strength,
                           float* __restrict__ r, float* __restrict__ t, float* __restrict__ eps) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const fh_eb_material m = mats[mat[s >> 1]];
    const float rr = fminf(fmaxf(m.tirReflectance, 0.0f), 1.0f);
    const float tt = fminf(fmaxf(m.tirTransmittance, 0.0f), 1.0f - rr);
    const float e = fmaxf(0.0f, 1.0f - rr - tt);
    if (strength > 0.0f && (s >> 1) >= leafCount) {
        const float eh = hapke::emissivity(e, fabsf(nrm[s >> 1].y), strength);
        r[s] = fminf(fmaxf(1.0f - tt - eh, 0.0f), 1.0f);
        t[s] = tt;
        eps[s] = eh;
        return;
    }
    r[s] = fminf(fmaxf(1.0f - tt - e, 0.0f), 1.0f);
    t[s] = tt;
    eps[s] = fminf(fmaxf(1.0f - r[s] - tt, 0.0f), 1.0f);
}

// LW source c = e + r fsky_s L + t fsky_o L with e = eps sigma T^4 (FP64)
__global__ void k_lw_source(uint32_t S, const float* __restrict__ T, const float* __restrict__ eps,
                            const float* __restrict__ r, const float* __restrict__ t, const float* __restrict__ fsky,
                            float L, float* __restrict__ e, float* __restrict__ c) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const double Tc = clampd(double(T[s]), 1.0, 1000.0);
    const float em = float(double(eps[s]) * kSigma * Tc * Tc * Tc * Tc);
    e[s] = em;
    c[s] = em + r[s] * (fsky[s] * L) + t[s] * (fsky[s ^ 1u] * L);
}

// SW window optics per side and band (makeEnergyShortwaveOptics,
// core:952-1003): the material's reflectance times the facet's texture
// factor (when given), clamped to [0, 1]; the ground sides' reflectance is
// the Hapke value of that for this sun (core:988-996): the Lambertian value
// clamped to 0.999 at strength 0.
// Internal band b is window order[b]; dni is given in internal order.
__global__ void k_sw_optics(uint32_t S, uint32_t nb, uint32_t ldb, const uint32_t* __restrict__ mat,
                            const float* __restrict__ refl, const float* __restrict__ trans,
                            const uint32_t* __restrict__ order, const float4* __restrict__ nrm,
                            const float* __restrict__ sunlit, float sx, float sy, float sz,
                            const float* __restrict__ dni, uint32_t leafCount, float strength,
                            const float* __restrict__ factor, float* __restrict__ r, float* __restrict__ t,
                            float* __restrict__ D, float* __restrict__ dir) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(S) * ldb) return;
    const uint32_t s = uint32_t(i / ldb), b = uint32_t(i - size_t(s) * ldb);
    if (b == 0) {
        const float4 n = nrm[s >> 1];
        const float sg = (s & 1u) ? -1.0f : 1.0f;
        const float cosv = fmaxf(0.0f, (sg * n.x) * sx + (sg * n.y) * sy + (sg * n.z) * sz);
        dir[s] = __fmul_rn(fminf(fmaxf(sunlit[s], 0.0f), 1.0f), cosv);
    }
    if (b >= nb) {
        r[i] = 0.0f;
        t[i] = 0.0f;
        D[i] = 0.0f;
        return;
    }
    const uint32_t m = mat[s >> 1], w = order[b];
    const float mean = refl[size_t(m) * nb + w];
    float rr = fminf(fmaxf(factor ? mean * factor[s >> 1] : mean, 0.0f), 1.0f);
    const float tt = fminf(fmaxf(trans[size_t(m) * nb + w], 0.0f), 1.0f - rr);
    if ((s >> 1) >= leafCount) {
        // ground sides: backgroundHapkeReflectance for every side (core:989-992);
        // at strength 0 it is the Lambertian value clamped to 0.999
        if (strength > 0.0f) {
            const float4 nf = nrm[s >> 1];
            const float sg = (s & 1u) ? -1.0f : 1.0f;
            const float sn[3] = {nf.x * sg, nf.y * sg, nf.z * sg};
            const float sun[3] = {sx, sy, sz};
            rr = hapke::reflectance(rr, sun, sn, sn, strength);
        } else {
            rr = fminf(rr, 0.999f);
        }
    }
    r[i] = rr;
    t[i] = fminf(tt, 1.0f - rr);
    const float4 n = nrm[s >> 1];
    const float sg = (s & 1u) ? -1.0f : 1.0f;
    const float cosv = fmaxf(0.0f, (sg * n.x) * sx + (sg * n.y) * sy + (sg * n.z) * sz);
    D[i] = __fmul_rn(__fmul_rn(fminf(fmaxf(sunlit[s], 0.0f), 1.0f), cosv), dni[b]);
}

// absorbed SW = sum_b alpha H, incident = sum_b H; FP64 block partials of
// sum A (D + fsky S), sum A alpha H, sum A fsky B for the energy closure.
// A team of 16 lanes owns one side, each lane 4 contiguous bands per pass
// (16-byte loads of the [S][ldb] rows), summed over the team by shuffles.
// per-side direct factor clamp(sunlit, 0, 1) max(0, cos(normal, sun)) (the
// b = 0 part of k_sw_optics), for solves that read the material tables
__global__ void k_sun_dir(uint32_t S, const float4* __restrict__ nrm, const float* __restrict__ sunlit, float sx,
                          float sy, float sz, float* __restrict__ dir) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const float4 n = nrm[s >> 1];
    const float sg = (s & 1u) ? -1.0f : 1.0f;
    const float cosv = fmaxf(0.0f, (sg * n.x) * sx + (sg * n.y) * sy + (sg * n.z) * sz);
    dir[s] = __fmul_rn(fminf(fmaxf(sunlit[s], 0.0f), 1.0f), cosv);
}

// Absorbed / incident SW per side.  TAB: r, t from the material tables and
// D = dir dni (the values k_sw_optics would store); FROMG: the incident
// H = max(0, fsky S + G + D) from the accumulated gathers G of an increment
// solve (else the H array).
struct AbsorbTables {
    const uint32_t* mat;
    const float* rtab;
    const float* ttab;
    const float* dir;
    const float* dniv;
    const float* G;
};

template <bool TAB, bool FROMG>
__global__ __launch_bounds__(256) void k_absorb(uint32_t S, uint32_t nb, uint32_t ldb, const float* __restrict__ r,
                                                const float* __restrict__ t, const float* __restrict__ H,
                                                const float* __restrict__ D, const float* __restrict__ B,
                                                const float* __restrict__ fsky, const float* __restrict__ skyv,
                                                const double* __restrict__ area, float* __restrict__ absorbed,
                                                float* __restrict__ incident, double* __restrict__ part,
                                                AbsorbTables tb) {
    const uint32_t team = threadIdx.x >> 4, lane = threadIdx.x & 15u;
    const uint32_t s = blockIdx.x * 16u + team;
    double pin = 0.0, pabs = 0.0, pesc = 0.0;
    float a = 0.0f, h = 0.0f;
    if (s < S) {
        const double A = area[s >> 1];
        const float fs = fsky[s];
        // ldb is 1 or a multiple of 4 (band_stride); the single band is lane 0's
        // (TAB: band-parallel strides, ldb >= 16)
        const bool vec = ldb >= 4u;
        const size_t mrow = TAB ? size_t(tb.mat[s >> 1]) * ldb : 0u;
        const float dsd = TAB ? tb.dir[s] : 0.0f;
        for (uint32_t b0 = 4u * lane; b0 < nb; b0 += 64u) {
            const size_t i = size_t(s) * ldb + b0;
            float rr[4], tt[4], hh[4], dd[4];
            if constexpr (TAB) {
                const float4 R4 = *reinterpret_cast<const float4*>(tb.rtab + mrow + b0);
                const float4 T4 = *reinterpret_cast<const float4*>(tb.ttab + mrow + b0);
                const float4 N4 = *reinterpret_cast<const float4*>(tb.dniv + b0);
                const float nn[4] = {N4.x, N4.y, N4.z, N4.w};
                rr[0] = R4.x; rr[1] = R4.y; rr[2] = R4.z; rr[3] = R4.w;
                tt[0] = T4.x; tt[1] = T4.y; tt[2] = T4.z; tt[3] = T4.w;
#pragma unroll
                for (int j = 0; j < 4; ++j) dd[j] = __fmul_rn(dsd, nn[j]);
            } else {
                const float4 R4 = vec ? *reinterpret_cast<const float4*>(r + i) : make_float4(r[i], 0.f, 0.f, 0.f);
                const float4 T4 = vec ? *reinterpret_cast<const float4*>(t + i) : make_float4(t[i], 0.f, 0.f, 0.f);
                const float4 D4 = vec ? *reinterpret_cast<const float4*>(D + i) : make_float4(D[i], 0.f, 0.f, 0.f);
                rr[0] = R4.x; rr[1] = R4.y; rr[2] = R4.z; rr[3] = R4.w;
                tt[0] = T4.x; tt[1] = T4.y; tt[2] = T4.z; tt[3] = T4.w;
                dd[0] = D4.x; dd[1] = D4.y; dd[2] = D4.z; dd[3] = D4.w;
            }
            if constexpr (FROMG) {
                const float4 G4 = *reinterpret_cast<const float4*>(tb.G + i);
                const float gg[4] = {G4.x, G4.y, G4.z, G4.w};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint32_t b = b0 + uint32_t(j);
                    hh[j] = b < nb ? fmaxf(fs * skyv[b] + gg[j] + dd[j], 0.0f) : 0.0f;
                }
            } else {
                const float4 H4 = vec ? *reinterpret_cast<const float4*>(H + i) : make_float4(H[i], 0.f, 0.f, 0.f);
                hh[0] = H4.x; hh[1] = H4.y; hh[2] = H4.z; hh[3] = H4.w;
            }
            const float4 B4 = vec ? *reinterpret_cast<const float4*>(B + i) : make_float4(B[i], 0.f, 0.f, 0.f);
            const float bb[4] = {B4.x, B4.y, B4.z, B4.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t b = b0 + uint32_t(j);
                if (b >= nb) break;
                const float alpha = fminf(fmaxf(1.0f - rr[j] - tt[j], 0.0f), 1.0f);
                a += alpha * hh[j];
                h += hh[j];
                pin += A * (double(dd[j]) + double(fs) * skyv[b]);
                pabs += A * double(alpha) * hh[j];
                pesc += A * double(fs) * bb[j];
            }
        }
    }
#pragma unroll
    for (int m = 8; m > 0; m >>= 1) {
        a += __shfl_xor(a, m, 16);
        h += __shfl_xor(h, m, 16);
    }
    if (lane == 0 && s < S) {
        absorbed[s] = a;
        incident[s] = h;
    }
    __shared__ double sh[3][256];
    sh[0][threadIdx.x] = pin;
    sh[1][threadIdx.x] = pabs;
    sh[2][threadIdx.x] = pesc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (int(threadIdx.x) < off)
            for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) part[size_t(blockIdx.x) * 4 + k] = sh[k][0];
}

// LW emission of a side at temperature T (FP64, stored FP32)
__device__ __forceinline__ float lw_emission(float eps, float T) {
    const double Tc = clampd(double(T), 1.0, 1000.0);
    return float(double(eps) * kSigma * Tc * Tc * Tc * Tc);
}

struct NewtonOut {
    float dT, eq;  // temperature step, |residual| / max(1, thermal response)
    float res;     // |residual|, W m-2
    double rn;
};

// Fluxes, residual and relaxed Newton update of one facet (FP64) given the
// incident LW H0, H1 of its two sides.  Leaf facets share side 0's T.
//
// Photovoltaic object facets (core:3655-3663, 3772-3778): power P =
// clamp(eta25 (1 + gamma (T - 298.15)) (I_0 + bifaciality I_1), 0, A_0 + A_1)
// with I the incident and A the absorbed shortwave of the two sides leaves
// the balance (residual - P), and while unclamped its slope eta25 gamma I
// joins the Newton derivative.
__device__ NewtonOut newton_facet(uint32_t f, const EbConst& k, const fh_eb_material& m,
                                  const float* __restrict__ absorbed, const float* __restrict__ incidentSW, float H0,
                                  float H1, const float* __restrict__ eps, float* __restrict__ T,
                                  const float* __restrict__ Tprev, bool update, float* __restrict__ Rn_out,
                                  float* __restrict__ H_out, float* __restrict__ LE_out, float* __restrict__ G_out,
                                  float* __restrict__ pv_out, const double* __restrict__ area,
                                  float* __restrict__ stepPrev = nullptr) {
    NewtonOut o{0.0f, 0.0f, 0.0f, 0.0};
    const uint32_t s0 = 2u * f;
    const float Hs[2] = {H0, H1};
    double rn[2];
    for (int q = 0; q < 2; ++q) {
        const uint32_t s = s0 + q;
        rn[q] = double(absorbed[s]) + double(eps[s]) * Hs[q] - double(lw_emission(eps[s], T[s]));
    }
    if (f < k.leafCount) {
        const double Tl = T[s0];
        const double net = rn[0] + rn[1];
        const double hs = 2.0 * k.ha * m.convectiveScale * (Tl - k.Ta);
        double rs = m.surfaceResistance;
        if (k.vegMethod == 1u) {
            const double par = fmax(0.0, double(absorbed[s0]) + absorbed[s0 + 1]);
            const double open = 0.25 + 0.75 * par / (par + 180.0);
            const double tc = (Tl - 273.15 - 25.0) / 18.0;
            rs = clampd(m.surfaceResistance / fmax(0.08, open * exp(-(tc * tc))), 40.0, 100000.0);
        }
        double le = 0.0, dle = 0.0;
        if (m.vegetation) {
            latent(Tl, k, rs, le, dle);
            le *= 2.0;
            dle *= 2.0;
        }
        const double gd = fmax(0.0, double(m.heatCapacityPerArea)) / k.dTime;
        const double g = gd * (Tl - double(Tprev[s0]));
        double pv = 0.0, dpv = 0.0;
        if (m.photovoltaic) {
            const double inc = double(incidentSW[s0]) + double(m.pvBifaciality) * double(incidentSW[s0 + 1]);
            const double avail = double(absorbed[s0]) + double(absorbed[s0 + 1]);
            const double raw = double(m.pvEta25) * (1.0 + double(m.pvGamma) * (Tl - 298.15)) * inc;
            pv = clampd(raw, 0.0, avail);
            dpv = raw > 0.0 && raw < avail ? double(m.pvEta25) * double(m.pvGamma) * inc : 0.0;
        }
        const double res = net - hs - le - g - pv;
        const double der =
            8.0 * double(eps[s0]) * kSigma * Tl * Tl * Tl + 2.0 * k.ha * m.convectiveScale + dle + gd + dpv;
        o.eq = float(fabs(res) / fmax(1.0, der));
        o.res = float(fabs(res));
        o.rn = area[f] * net;
        if (Rn_out)
            for (int q = 0; q < 2; ++q) {
                Rn_out[s0 + q] = float(net);
                H_out[s0 + q] = float(hs);
                LE_out[s0 + q] = float(le);
                G_out[s0 + q] = float(g);
            }
        if (pv_out) pv_out[s0] = pv_out[s0 + 1] = float(pv);
        if (update) {
            const double cand = clampd(Tl + res / fmax(1.0, der), 220.0, 360.0);
            const double w = step_relaxation(k, stepPrev ? stepPrev + s0 : nullptr, cand - Tl);
            const double upd = Tl + w * (cand - Tl);
            o.dT = float(fabs(upd - Tl));
            T[s0] = float(upd);
            T[s0 + 1] = float(upd);
        }
    } else {
        const bool dry = m.soilMoisture < 0.03f;
        for (int q = 0; q < 2; ++q) {
            const uint32_t s = s0 + q;
            const double Ts = T[s];
            const double hs = k.ha * (Ts - k.Ta);
            double le = 0.0, dle = 0.0;
            if (!dry) latent(Ts, k, m.surfaceResistance, le, dle);
            double g, gd;
            if (m.soilTemperatureMethod == 0) {
                gd = 0.0;
                g = 0.35 * rn[q];
            } else {
                gd = soil_storage(m, k.dTime) * (m.soilTemperatureMethod == 2 ? 1.5 : 1.0);
                g = gd * (Ts - double(Tprev[s]));
            }
            const double res = rn[q] - hs - le - g;
            const double der = 4.0 * double(eps[s]) * kSigma * Ts * Ts * Ts + k.ha + dle + gd;
            o.eq = fmaxf(o.eq, float(fabs(res) / fmax(1.0, der)));
            o.res = fmaxf(o.res, float(fabs(res)));
            o.rn += area[f] * rn[q];
            if (Rn_out) {
                Rn_out[s] = float(rn[q]);
                H_out[s] = float(hs);
                LE_out[s] = float(le);
                G_out[s] = float(g);
            }
            if (pv_out) pv_out[s] = 0.0f;
            if (update) {
                const double cand = clampd(Ts + res / fmax(1.0, der), 220.0, 360.0);
                const double w = step_relaxation(k, stepPrev ? stepPrev + s : nullptr, cand - Ts);
                const double upd = Ts + w * (cand - Ts);
                o.dT = fmaxf(o.dT, float(fabs(upd - Ts)));
                T[s] = float(upd);
            }
        }
    }
    return o;
}

// Newton step over all facets from the incident LW H (one lane per facet).
// Writes per-block (max dT, max residual-equivalent) and the FP64 partial of
// sum A Rn over energy-balance sides.
__global__ __launch_bounds__(256) void k_newton(uint32_t F, EbConst k, const uint32_t* __restrict__ mat,
                                                const fh_eb_material* __restrict__ mats,
                                                const float* __restrict__ absorbed,
                                                const float* __restrict__ incidentSW, const float* __restrict__ H,
                                                const float* __restrict__ eps, float* __restrict__ T,
                                                const float* __restrict__ Tprev, int update, float* __restrict__ Rn_out,
                                                float* __restrict__ H_out, float* __restrict__ LE_out,
                                                float* __restrict__ G_out, float* __restrict__ pv_out,
                                                const double* __restrict__ area, float4* __restrict__ part,
                                                double* __restrict__ dpart) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    float v[3] = {0.0f, 0.0f, 0.0f};
    double rnSum = 0.0;
    if (f < F) {
        const NewtonOut o = newton_facet(f, k, mats[mat[f]], absorbed, incidentSW, H[2u * f], H[2u * f + 1], eps, T,
                                         Tprev, update != 0, Rn_out, H_out, LE_out, G_out, pv_out, area);
        v[0] = o.dT;
        v[1] = o.eq;
        v[2] = o.res;
        rnSum = o.rn;
    }
#pragma unroll
    for (int msk = 32; msk > 0; msk >>= 1) rnSum += __shfl_xor(rnSum, msk, 64);
    __shared__ double srn[4];
    if ((threadIdx.x & 63) == 0) srn[threadIdx.x >> 6] = rnSum;
    float out[3];
    block_max<3>(v, out);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = make_float4(out[0], out[1], out[2], 0.0f);
        dpart[size_t(blockIdx.x) * 4 + 3] = (srn[0] + srn[1]) + (srn[2] + srn[3]);
    }
}

// ---- fused coupled sweep ----------------------------------------------------
// One nonlinear Jacobi-Newton iteration of the coupled LW radiosity and
// energy balance per launch: each facet team gathers G = F B for its two
// sides, forms the incident LW H = max(0, fsky L + G), takes the relaxed
// Newton step of the facet's temperature from H (FP64), recomputes the LW
// emission at the new temperature and writes the Jacobi update
// B' = e(T') + r (fsky_s L + G_s) + t (fsky_o L + G_o).  The fixed point is
// the reference's converged coupling (a LW solve converged for T and a zero
// energy residual); the loop stops when max dT, max residual-equivalent and
// the largest relative change of B are all within tolerance.
struct CoupledArgs {
    const uint32_t* rowptr;
    const uint32_t* col;
    const float* w32;
    const uint16_t* w16;
    const float* Bin;
    float* Bout;
    const float* fsky;
    const float* r;
    const float* t;
    const float* eps;
    const float* absorbed;
    const float* incidentSW;  // incident shortwave per side (photovoltaic power)
    float* T;
    const float* Tprev;
    float* Hout;
    const uint32_t* mat;
    const fh_eb_material* mats;
    const double* area;
    float L;
    uint32_t F;
    EbConst k;
    SolveCtrl* ctrl;
    float4* partial;
    float* Iin;  // per side: unclamped incident fsky L + G of the gather half
    float* stepPrev;  // per side: previous full Newton step (adaptive relaxation)
    const uint32_t* sched;  // facet gathered by logical team i (null: facet i)
    const uint32_t* pk;     // packed CSR entries (FP32 gathers), or null
    const uint32_t* den;
    uint32_t pkBits;
};

// Gather half of a coupled sweep: a team of TEAM lanes per facet sums
// G = F B over both sides' rows and stores the unclamped incident
// fsky L + G per side in Iin.
template <int TEAM, bool W16>
__global__ __launch_bounds__(256) void k_coupled_gather(CoupledArgs a) {
    if (ctrl_word(&a.ctrl->done)) return;
    const int lane = threadIdx.x % TEAM;
    const uint32_t fl = (blockIdx.x * 256u + threadIdx.x) / TEAM;
    if (fl >= a.F) return;
    const uint32_t f = a.sched ? a.sched[fl] : fl;
    const uint32_t s0 = 2u * f;
    const uint32_t e0 = a.rowptr[s0], e1 = a.rowptr[s0 + 1], e2 = a.rowptr[s0 + 2];
    float g[2][1] = {{0.0f}, {0.0f}};
    if (!W16 && a.pk)
        team_rows_dot2_pk<TEAM, (TEAM >= 32 ? 4 : 8)>(a.pk, a.pkBits, row_den(a.den[s0]), row_den(a.den[s0 + 1]),
                                                      a.Bin, e0, e1, e1, e2, lane, g[0][0], g[1][0]);
    else
        team_rows_dot2<TEAM, (TEAM >= 32 ? 4 : 8), W16>(a.col, a.w32, a.w16, a.Bin, e0, e1, e1, e2, lane, g[0][0],
                                                        g[1][0]);
    team_reduce<1, TEAM>(g[0]);
    team_reduce<1, TEAM>(g[1]);
    if (lane == 0) {
        a.Iin[s0] = a.fsky[s0] * a.L + g[0][0];
        a.Iin[s0 + 1] = a.fsky[s0 + 1] * a.L + g[1][0];
    }
}

// Update half, one thread per facet (all lanes busy with the FP64 Newton
// step): H = max(0, incident), relaxed Newton step of the facet's
// temperature, LW emission at the new temperature and the Jacobi update
// B' = e(T') + r in_s + t in_o; per-block maxima of |dB|, |dB|/|B|, dT and
// the residual-equivalent.
__global__ __launch_bounds__(256) void k_coupled_update(CoupledArgs a) {
    if (ctrl_word(&a.ctrl->done)) return;
    const uint32_t f = blockIdx.x * 256u + threadIdx.x;
    float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};  // |dB|, |dB|/|B|, dT, residual-equivalent
    if (f < a.F) {
        const uint32_t s0 = 2u * f;
        const float in0 = a.Iin[s0], in1 = a.Iin[s0 + 1];
        const float H0 = fmaxf(in0, 0.0f), H1 = fmaxf(in1, 0.0f);
        a.Hout[s0] = H0;
        a.Hout[s0 + 1] = H1;
        const NewtonOut o = newton_facet(f, a.k, a.mats[a.mat[f]], a.absorbed, a.incidentSW, H0, H1, a.eps, a.T,
                                         a.Tprev, true, nullptr, nullptr, nullptr, nullptr, nullptr, a.area,
                                         a.stepPrev);
        v[2] = o.dT;
        v[3] = o.eq;
        const float ins[2] = {in0, in1};
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const uint32_t s = s0 + q;
            const float em = lw_emission(a.eps[s], a.T[s]);
            const float nb = em + a.r[s] * ins[q] + a.t[s] * ins[q ^ 1];
            const float old = a.Bin[s];
            a.Bout[s] = nb;
            const float d = fabsf(nb - old);
            v[0] = fmaxf(v[0], d);
            v[1] = fmaxf(v[1], d / fmaxf(fabsf(nb), 1.0e-30f));
        }
    }
    float out[4];
    block_max<4>(v, out);
    if (threadIdx.x == 0) a.partial[blockIdx.x] = make_float4(out[0], out[1], out[2], out[3]);
}

// Reduces the coupled-sweep partials; done when max dT <= tolT, the max
// residual-equivalent <= tolT and max |dB|/|B| <= tolRel (or the budget ends).
__global__ __launch_bounds__(1024) void k_coupled_finalize(SolveCtrl* ctrl, const float4* __restrict__ part,
                                                           uint32_t n, float tolT, float* __restrict__ last4) {
    if (*reinterpret_cast<volatile uint32_t*>(&ctrl->done)) return;
    float m[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t i = threadIdx.x; i < n; i += 1024u) {
        const float4 p = part[i];
        m[0] = fmaxf(m[0], p.x);
        m[1] = fmaxf(m[1], p.y);
        m[2] = fmaxf(m[2], p.z);
        m[3] = fmaxf(m[3], p.w);
    }
    __shared__ float sh[1024][4];
    for (int q = 0; q < 4; ++q) sh[threadIdx.x][q] = m[q];
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (int(threadIdx.x) < off)
            for (int q = 0; q < 4; ++q) sh[threadIdx.x][q] = fmaxf(sh[threadIdx.x][q], sh[threadIdx.x + off][q]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const uint32_t it = ctrl->iter + 1u;
        ctrl->iter = it;
        ctrl->lastAbs = sh[0][0];
        ctrl->lastRel = sh[0][1];
        for (int q = 0; q < 4; ++q) last4[q] = sh[0][q];
        const bool conv = sh[0][2] <= tolT && sh[0][3] <= tolT && sh[0][1] <= ctrl->tolRel;
        if (conv || it >= ctrl->maxIter) ctrl->done = 1u;
    }
}

// max over n float4 partials, component-wise, into out[0..3]
__global__ __launch_bounds__(1024) void k_reduce_max4(const float4* __restrict__ part, uint32_t n,
                                                      float* __restrict__ out) {
    float m[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t i = threadIdx.x; i < n; i += 1024u) {
        const float4 p = part[i];
        m[0] = fmaxf(m[0], p.x);
        m[1] = fmaxf(m[1], p.y);
        m[2] = fmaxf(m[2], p.z);
        m[3] = fmaxf(m[3], p.w);
    }
    float r[4];
    block_max_1024<4>(m, r);
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; ++q) out[q] = r[q];
}

__global__ __launch_bounds__(1024) void k_reduce_max2(const float2* __restrict__ part, uint32_t n,
                                                      float2* __restrict__ out) {
    float a = 0.0f, b = 0.0f;
    for (uint32_t i = threadIdx.x; i < n; i += 1024u) {
        a = fmaxf(a, part[i].x);
        b = fmaxf(b, part[i].y);
    }
    __shared__ float sa[1024], sb[1024];
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (int(threadIdx.x) < off) {
            sa[threadIdx.x] = fmaxf(sa[threadIdx.x], sa[threadIdx.x + off]);
            sb[threadIdx.x] = fmaxf(sb[threadIdx.x], sb[threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = make_float2(sa[0], sb[0]);
}

// sums columns [k0, k1) (k1 <= 4) of n rows of 4 doubles in one pass (each
// thread a fixed set of rows, 32 coalesced bytes per row, then a fixed tree):
// deterministic
__global__ __launch_bounds__(1024) void k_reduce_sum4(const double* __restrict__ part, uint32_t n, uint32_t k0,
                                                      uint32_t k1, double* __restrict__ out) {
    __shared__ double sh[4][1024];
    const uint32_t t = threadIdx.x;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t i = t; i < n; i += 1024u) {
        const double2 lo = *reinterpret_cast<const double2*>(part + size_t(i) * 4);
        const double2 hi = *reinterpret_cast<const double2*>(part + size_t(i) * 4 + 2);
        acc[0] += lo.x;
        acc[1] += lo.y;
        acc[2] += hi.x;
        acc[3] += hi.y;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) sh[k][t] = acc[k];
    __syncthreads();
    for (uint32_t off = 512u; off > 0u; off >>= 1) {
        if (t < off)
#pragma unroll
            for (int k = 0; k < 4; ++k) sh[k][t] += sh[k][t + off];
        __syncthreads();
    }
    if (t >= k0 && t < k1) out[t] = sh[t][0];
}

// per side: sum of the nb bands of a [S][ldb] array, in FP64
__global__ void k_band_sum(const float* __restrict__ B, uint32_t S, uint32_t nb, uint32_t ldb,
                           double* __restrict__ out) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    double acc = 0.0;
    for (uint32_t j = 0; j < nb; ++j) acc += double(B[size_t(s) * ldb + j]);
    out[s] = acc;
}

__global__ void k_copy_f32(const float* __restrict__ a, float* __restrict__ b, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

static inline dim3 blocks_for(size_t n, uint32_t bs = 256) { return dim3(uint32_t((n + bs - 1) / bs)); }

static void copy_f32(fh_context* c, const float* a, float* b, size_t n) {
    if (!n) return;
    hipLaunchKernelGGL(k_copy_f32, blocks_for(n), dim3(256), 0, c->stream, a, b, n);
    FH_HIP(hipGetLastError());
}

// surrogate warm start and feature capture (mlp.hip)
void surrogate_guess(fh_context* c, const EbConst& k, float windSpeed, float skyLongwave, float Ta);
void eb_features(fh_context* c, float Ta, float wind, float sky);

void eb_set_scene(fh_context* c, uint32_t leafCount, const uint32_t* facetMaterial, uint32_t nMat,
                  const fh_eb_material* mats, uint32_t swNb, const float* refl, const float* trans) {
    EbState& e = c->eb;
    const uint32_t F = c->F, S = c->S;
    e.ready = false;
    e.facets = F;
    e.leafCount = leafCount;
    e.nMat = nMat;
    e.swNb = swNb;
    e.swLdb = band_stride(swNb);
    e.material.ensure(F);
    e.mats.ensure(nMat);
    e.swRefl.ensure(size_t(nMat) * swNb);
    e.swTrans.ensure(size_t(nMat) * swNb);
    e.area.ensure(F);
    FH_HIP(hipMemcpyAsync(e.material.p, facetMaterial, sizeof(uint32_t) * F, hipMemcpyHostToDevice, c->stream));
    FH_HIP(hipMemcpyAsync(e.mats.p, mats, sizeof(fh_eb_material) * nMat, hipMemcpyHostToDevice, c->stream));
    if (swNb) {
        FH_HIP(hipMemcpyAsync(e.swRefl.p, refl, sizeof(float) * nMat * swNb, hipMemcpyHostToDevice, c->stream));
        FH_HIP(hipMemcpyAsync(e.swTrans.p, trans, sizeof(float) * nMat * swNb, hipMemcpyHostToDevice, c->stream));
    }
    // internal band order: windows by decreasing largest single-scattering
    // albedo r + t over the materials (the slowest to converge first), so
    // the bands a solve still iterates form a prefix
    std::vector<float> key(swNb, 0.0f);
    for (uint32_t b = 0; b < swNb; ++b)
        for (uint32_t m = 0; m < nMat; ++m) {
            const float r = std::min(std::max(refl[size_t(m) * swNb + b], 0.0f), 1.0f);
            const float t = std::min(std::max(trans[size_t(m) * swNb + b], 0.0f), 1.0f - r);
            key[b] = std::max(key[b], r + t);
        }
    // a ground side's shortwave reflectance is clamped to 0.999 (the Hapke
    // model's base, core:989-992); the material tables hold the values of
    // leaves, so a ground material above the clamp takes the per-side arrays
    std::vector<uint8_t> groundMat(nMat, 0u);
    for (uint32_t f = leafCount; f < F; ++f) groundMat[facetMaterial[f]] = 1u;
    e.groundClamp = false;
    for (uint32_t m = 0; m < nMat; ++m)
        for (uint32_t b = 0; b < swNb && groundMat[m]; ++b)
            if (refl[size_t(m) * swNb + b] > 0.999f) e.groundClamp = true;
    e.swOrder.resize(swNb);
    for (uint32_t b = 0; b < swNb; ++b) e.swOrder[b] = b;
    std::stable_sort(e.swOrder.begin(), e.swOrder.end(), [&](uint32_t x, uint32_t y) { return key[x] > key[y]; });
    e.swOrderD.ensure(swNb ? swNb : 1u);
    if (swNb)
        FH_HIP(hipMemcpyAsync(e.swOrderD.p, e.swOrder.data(), sizeof(uint32_t) * swNb, hipMemcpyHostToDevice,
                              c->stream));
    for (DBuf<float>* b : {&e.T, &e.Tprev, &e.absorbed, &e.incident, &e.pvPower, &e.sunlit, &e.rlw, &e.tlw, &e.eps,
                           &e.elw, &e.clw, &e.Hlw, &e.zero, &e.Blw0, &e.Blw1, &e.Rn, &e.Hs, &e.LE, &e.G})
        b->ensure(S);
    const size_t nsw = size_t(S) * (e.swLdb ? e.swLdb : 1u);
    for (DBuf<float>* b : {&e.Bsw0, &e.Bsw1, &e.rsw, &e.tsw, &e.Dsw, &e.csw, &e.Hsw}) b->ensure(nsw);
    e.swDir.ensure(S);
    if (swNb) {
        // r, t per material in the internal band order, clamped as the
        // per-side arrays of k_sw_optics (the band-parallel sweeps read these)
        const uint32_t ldb = e.swLdb;
        std::vector<float> rt(size_t(nMat) * ldb, 0.0f), tt(size_t(nMat) * ldb, 0.0f);
        for (uint32_t m = 0; m < nMat; ++m)
            for (uint32_t b = 0; b < swNb; ++b) {
                const uint32_t w = e.swOrder[b];
                const float r = std::fmin(std::fmax(refl[size_t(m) * swNb + w], 0.0f), 1.0f);
                const float t = std::fmin(std::fmax(trans[size_t(m) * swNb + w], 0.0f), 1.0f - r);
                rt[size_t(m) * ldb + b] = r;
                tt[size_t(m) * ldb + b] = std::fmin(t, 1.0f - r);
            }
        e.swRtab.ensure(rt.size());
        e.swTtab.ensure(tt.size());
        FH_HIP(hipMemcpyAsync(e.swRtab.p, rt.data(), sizeof(float) * rt.size(), hipMemcpyHostToDevice, c->stream));
        FH_HIP(hipMemcpyAsync(e.swTtab.p, tt.data(), sizeof(float) * tt.size(), hipMemcpyHostToDevice, c->stream));
    }
    e.dni.ensure(e.swLdb ? e.swLdb : 1u);
    e.lit.ensure(S);
    e.dsum.ensure(4);
    e.fout.ensure(1);
    hipLaunchKernelGGL(k_areas, blocks_for(F), dim3(256), 0, c->stream, c->facetVerts.p, F, e.area.p);
    hipLaunchKernelGGL(k_lw_setup, blocks_for(S), dim3(256), 0, c->stream, S, e.material.p, e.mats.p, leafCount,
                       c->facetNormal.p, c->hapkeStrength, e.rlw.p, e.tlw.p, e.eps.p);
    launch_fill_f32(c, e.zero.p, S, 0.0f);
    launch_fill_f32(c, e.absorbed.p, S, 0.0f);
    launch_fill_f32(c, e.incident.p, S, 0.0f);
    launch_fill_f32(c, e.pvPower.p, S, 0.0f);
    launch_fill_f32(c, e.Bsw0.p, nsw, 0.0f);
    e.swIncSerial = 0;
    launch_fill_f32(c, e.Blw0.p, S, 0.0f);
    for (DBuf<float>* b : {&e.Rn, &e.Hs, &e.LE, &e.G, &e.sunlit}) launch_fill_f32(c, b->p, S, 0.0f);
    FH_HIP(hipGetLastError());
    FH_HIP(hipStreamSynchronize(c->stream));
    e.swValid = e.lwValid = e.tValid = false;
    e.swRhoValid = false;
    e.guessPrevValid = false;
    e.lastNodeValid = false;
    e.textured = false;
    if (swNb && c->hasGraph) {
        // spectral radius per internal band of this scene's SW operator (the
        // Chebyshev parameter): r and t do not depend on the sun, so the
        // estimate runs here, once per scene, with the direct term zero
        const uint32_t ldb = e.swLdb;
        launch_fill_f32(c, e.dni.p, ldb, 0.0f);
        hipLaunchKernelGGL(k_sw_optics, blocks_for(nsw), dim3(256), 0, c->stream, S, swNb, ldb, e.material.p,
                           e.swRefl.p, e.swTrans.p, e.swOrderD.p, c->facetNormal.p, e.sunlit.p, 0.0f, 1.0f, 0.0f,
                           e.dni.p, e.leafCount, c->hapkeStrength, nullptr, e.rsw.p, e.tsw.p, e.Dsw.p, e.swDir.p);
        FH_HIP(hipGetLastError());
        e.swRho2.ensure(ldb);
        estimate_rho2(c, ldb, swNb, e.rsw.p, e.tsw.p, e.Bsw1.p, e.Hsw.p, e.csw.p, e.swRho2.p);
        e.swRhoValid = true;
    }
    e.ready = true;
}

static void newton(fh_context* c, const EbConst& k, int update, float& maxDT, float& maxEq, double* rnSum,
                   float* maxRes = nullptr) {
    EbState& e = c->eb;
    const uint32_t F = c->F;
    const uint32_t nbk = (F + 255u) / 256u;
    e.part4.ensure(nbk);
    e.last4.ensure(4);
    e.dpart.ensure(size_t(nbk) * 4u);
    hipLaunchKernelGGL(k_newton, dim3(nbk), dim3(256), 0, c->stream, F, k, e.material.p, e.mats.p, e.absorbed.p,
                       e.incident.p, e.Hlw.p, e.eps.p, e.T.p, e.Tprev.p, update, e.Rn.p, e.Hs.p, e.LE.p, e.G.p,
                       e.pvPower.p, e.area.p, e.part4.p, e.dpart.p);
    hipLaunchKernelGGL(k_reduce_max4, dim3(1), dim3(1024), 0, c->stream, e.part4.p, nbk, e.last4.p);
    if (rnSum) hipLaunchKernelGGL(k_reduce_sum4, dim3(1), dim3(1024), 0, c->stream, e.dpart.p, nbk, 3u, 4u, e.dsum.p);
    FH_HIP(hipGetLastError());
    float h[4];
    FH_HIP(hipMemcpyAsync(h, e.last4.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    double d4[4] = {0, 0, 0, 0};
    if (rnSum) FH_HIP(hipMemcpyAsync(d4, e.dsum.p, sizeof(double) * 4, hipMemcpyDeviceToHost, c->stream));
    FH_HIP(hipStreamSynchronize(c->stream));
    maxDT = h[0];
    maxEq = h[1];
    if (maxRes) *maxRes = h[2];
    if (rnSum) *rnSum = d4[3];
}

// LW: source from T, warm or zero start, solve, incident (into e.Hlw)
static uint32_t lw_solve(fh_context* c, const fh_eb_node_params& p, uint32_t precision, float* maxDelta = nullptr) {
    EbState& e = c->eb;
    const uint32_t S = c->S;
    hipLaunchKernelGGL(k_lw_source, blocks_for(S), dim3(256), 0, c->stream, S, e.T.p, e.eps.p, e.rlw.p, e.tlw.p,
                       c->fsky.p, p.skyLongwave, e.elw.p, e.clw.p);
    FH_HIP(hipGetLastError());
    if (!e.lwValid || p.zeroStart) launch_fill_f32(c, e.Blw0.p, S, 0.0f);
    const float L = p.skyLongwave;
    set_sky(c, &L, 1u, 1u);
    JacobiOptions o;
    o.ldb = 1u;
    o.precision = precision;
    o.maxIter = p.maxSweeps;
    o.tolRel = p.radiosityTolerance;
    o.H = e.Hlw.p;
    o.D = e.zero.p;
    float* bufs[2] = {e.Blw0.p, e.Blw1.p};
    fh_solve_stats st{};
    bool hDone = false;
    const int fin = run_jacobi(c, o, e.clw.p, e.rlw.p, e.tlw.p, bufs, &st, &hDone);
    if (fin == 1) e.Blw0.swap(e.Blw1);
    e.lwValid = true;
    if (!hDone) run_incident(c, 1u, e.Blw0.p, e.zero.p, e.Hlw.p);
    if (maxDelta) *maxDelta = st.maxDelta;
    return st.iterations;
}

// Fused coupled sweeps until the three tolerances hold or maxIter; the
// result is left in Blw0 and the last (|dB|, |dB|/|B|, dT, residual-eq).
static uint32_t coupled_loop(fh_context* c, const fh_eb_node_params& p, const EbConst& k, bool w16,
                             uint32_t maxIter, float last[4]) {
    EbState& e = c->eb;
    const int team = sweep_team(c);
    const uint32_t grid = uint32_t((uint64_t(c->F) * uint64_t(team) + 255u) / 256u);
    const uint32_t gridU = (c->F + 255u) / 256u;
    e.part4.ensure(gridU);
    e.last4.ensure(4);
    e.Iin.ensure(c->S);
    launch_ctrl_reset(c, maxIter ? maxIter : 1u, p.radiosityTolerance, -1.0f, 1u);
    CoupledArgs a{};
    a.rowptr = c->rowptr.p;
    a.col = c->col.p;
    a.w32 = c->val32.p;
    a.w16 = c->val16.p;
    a.fsky = c->fsky.p;
    a.r = e.rlw.p;
    a.t = e.tlw.p;
    a.eps = e.eps.p;
    a.absorbed = e.absorbed.p;
    a.incidentSW = e.incident.p;
    a.T = e.T.p;
    a.Tprev = e.Tprev.p;
    a.Hout = e.Hlw.p;
    a.mat = e.material.p;
    a.mats = e.mats.p;
    a.area = e.area.p;
    a.L = p.skyLongwave;
    a.F = c->F;
    a.k = k;
    a.ctrl = c->ctrl.p;
    a.partial = e.part4.p;
    a.Iin = e.Iin.p;
    a.sched = sweep_schedule(c);
    a.pk = c->pkBits && use_packed() ? c->pk.p : nullptr;
    a.den = c->den.p;
    a.pkBits = c->pkBits;
    e.stepPrev.ensure(c->S);
    launch_fill_f32(c, e.stepPrev.p, c->S, 0.0f);
    a.stepPrev = e.stepPrev.p;
    uint64_t launched = 0;
    uint32_t batch = 4;
    while (true) {
        for (uint32_t q = 0; q < batch; ++q) {
            a.Bin = (launched & 1u) ? e.Blw1.p : e.Blw0.p;
            a.Bout = (launched & 1u) ? e.Blw0.p : e.Blw1.p;
            // FP32 gathers through the staged layout (every distinct column
            // of a workgroup gathered once; the pair blocks' facets, which
            // the layout leaves out, in the team kernel), else the team
            // kernel over all facets
            const bool staged = !w16 && launch_gather1_staged(c, a.Bin, a.L, c->ctrl.p, a.Iin);
            CoupledArgs g = a;
            uint32_t gg = grid;
            if (staged) {
                g.sched = c->pairFacets.p;
                g.F = 2u * c->unitCount;
                gg = uint32_t((uint64_t(g.F) * uint64_t(team) + 255u) / 256u);
            }
#define FH_COUPLED(T, W) hipLaunchKernelGGL((k_coupled_gather<T, W>), dim3(gg), dim3(256), 0, c->stream, g)
            if (gg) {
                if (team == 4) { if (w16) FH_COUPLED(4, true); else FH_COUPLED(4, false); }
                else if (team == 8) { if (w16) FH_COUPLED(8, true); else FH_COUPLED(8, false); }
                else if (team == 16) { if (w16) FH_COUPLED(16, true); else FH_COUPLED(16, false); }
                else if (team == 32) { if (w16) FH_COUPLED(32, true); else FH_COUPLED(32, false); }
                else { if (w16) FH_COUPLED(64, true); else FH_COUPLED(64, false); }
            }
#undef FH_COUPLED
            hipLaunchKernelGGL(k_coupled_update, dim3(gridU), dim3(256), 0, c->stream, a);
            hipLaunchKernelGGL(k_coupled_finalize, dim3(1), dim3(1024), 0, c->stream, c->ctrl.p, e.part4.p, gridU,
                               p.temperatureTolerance, e.last4.p);
            ++launched;
        }
        FH_HIP(hipGetLastError());
        FH_HIP(hipMemcpyAsync(c->hctrl, c->ctrl.p, sizeof(SolveCtrl), hipMemcpyDeviceToHost, c->stream));
        FH_HIP(hipMemcpyAsync(last, e.last4.p, sizeof(float) * 4, hipMemcpyDeviceToHost, c->stream));
        FH_HIP(hipStreamSynchronize(c->stream));
        if (c->hctrl->done) break;
        if (batch < 16u) batch *= 2u;
    }
    const uint32_t iters = c->hctrl->iter;
    if (iters & 1u) e.Blw0.swap(e.Blw1);
    return iters;
}

void eb_run_node(fh_context* c, const fh_eb_node_params& p, const float* dniHost, const float* diffHost,
                 fh_eb_node_stats* out) {
    EbState& e = c->eb;
    const uint32_t S = c->S;
    fh_eb_node_stats st{};
    // --- shortwave ---------------------------------------------------------
    const float sunLen = std::sqrt(p.sun.x * p.sun.x + p.sun.y * p.sun.y + p.sun.z * p.sun.z);
    double swTotal = 0.0;
    for (uint32_t b = 0; b < e.swNb; ++b) swTotal += double(dniHost[b]) + double(diffHost[b]);
    if (p.sun.y > 0.0f && sunLen > 0.0f) {
        launch_fill_u32(c, e.lit.p, S, 0u);
        launch_sun(c, p.sun.x, p.sun.y, p.sun.z, e.lit.p, e.sunlit.p);
    } else {
        launch_fill_f32(c, e.sunlit.p, S, 0.0f);
    }
    if (e.swNb && swTotal > 0.0) {
        const uint32_t nb = e.swNb, ldb = e.swLdb;
        // per-band inputs in the internal band order
        std::vector<float> dpad(ldb, 0.0f), diffInt(nb);
        for (uint32_t b = 0; b < nb; ++b) {
            dpad[b] = p.sun.y > 0.0f ? dniHost[e.swOrder[b]] : 0.0f;
            diffInt[b] = diffHost[e.swOrder[b]];
        }
        FH_HIP(hipMemcpyAsync(e.dni.p, dpad.data(), sizeof(float) * ldb, hipMemcpyHostToDevice, c->stream));
        e.lastDni = dpad;
        e.lastDiffuse = diffInt;
        const size_t n = size_t(S) * ldb;
        static const bool noTables = [] {
            const char* v = std::getenv("FH_SW_ARRAYS");
            return v && std::atoi(v) != 0;
        }();
        // band-parallel solves read r, t from the material tables and form c,
        // D from the per-side direct factors, and the absorbed pass does the
        // same: the [S][ldb] optics arrays and the folded source are only
        // built for the entry-parallel layouts (and the spectral estimate)
        // a Hapke ground's reflectance follows the sun: per-side arrays, and
        // the spectral radius estimated again for every node
        const bool hapkeOn = c->hapkeStrength > 0.0f;
        if (hapkeOn) e.swRhoValid = false;
        // per-facet reflectance factors (textures) also take the per-side arrays
        const bool tab = !noTables && !hapkeOn && !e.groundClamp && !e.textured && uses_bf16_gather(ldb);
        const uint32_t precision = bf16_pays(ldb) ? p.precision : FH_PRECISION_FP32;
        // an increment solve keeps the gathers of every band's last sweep:
        // the absorbed pass forms the incident from them
        const bool fromG = tab && increment_solve(ldb, precision);
        const auto optics_arrays = [&] {
            hipLaunchKernelGGL(k_sw_optics, blocks_for(n), dim3(256), 0, c->stream, S, nb, ldb, e.material.p,
                               e.swRefl.p, e.swTrans.p, e.swOrderD.p, c->facetNormal.p, e.sunlit.p, p.sun.x, p.sun.y,
                               p.sun.z, e.dni.p, e.leafCount, c->hapkeStrength, e.textured ? e.reflFactor.p : nullptr,
                               e.rsw.p, e.tsw.p, e.Dsw.p, e.swDir.p);
            FH_HIP(hipGetLastError());
        };
        if (!tab || !e.swRhoValid) {
            optics_arrays();
        } else {
            hipLaunchKernelGGL(k_sun_dir, blocks_for(S), dim3(256), 0, c->stream, S, c->facetNormal.p, e.sunlit.p,
                               p.sun.x, p.sun.y, p.sun.z, e.swDir.p);
            FH_HIP(hipGetLastError());
        }
        set_sky(c, diffInt.data(), nb, ldb);
        if (!e.swRhoValid) {
            // spectral radius per band of this scene's SW operator (Chebyshev)
            e.swRho2.ensure(ldb);
            estimate_rho2(c, ldb, nb, e.rsw.p, e.tsw.p, e.Bsw1.p, e.Hsw.p, e.csw.p, e.swRho2.p);
            e.swRhoValid = true;
        }
        if (!tab) {
            launch_fill_f32(c, e.Hsw.p, n, 0.0f);  // emission is zero in the SW
            launch_fold(c, e.Hsw.p, e.rsw.p, e.tsw.p, e.Dsw.p, ldb, e.csw.p);
        }
        if (!e.swValid || p.zeroStart) {
            launch_fill_f32(c, e.Bsw0.p, n, 0.0f);
            e.swIncSerial = 0;
        }
        JacobiOptions o;
        // from the previous node's final radiosity, an increment solve
        // continues that node's increment state (no base sweep) while it is
        // the context's latest
        o.incSerial = &e.swIncSerial;
        o.ldb = ldb;
        // BF16 gathers only where they save lines: below 16 windows a side's
        // FP32 row is at most one 64-byte gather, and the entry-parallel
        // BF16 scheme (defect correction) takes more sweeps than FP32
        o.precision = precision;
        o.maxIter = p.maxSweeps;
        o.tolRel = p.radiosityTolerance;
        if (!fromG) {
            o.H = e.Hsw.p;
            o.D = e.Dsw.p;
        }
        if (tab) {
            o.fmat = e.material.p;
            o.rtab = e.swRtab.p;
            o.ttab = e.swTtab.p;
            o.dirf = e.swDir.p;
            o.dniv = e.dni.p;
        }
        static const bool noCheb = [] {
            const char* v = std::getenv("FH_NO_CHEBYSHEV");
            return v && std::atoi(v) != 0;
        }();
        if (!noCheb) o.rho2 = e.swRho2.p;
        float* bufs[2] = {e.Bsw0.p, e.Bsw1.p};
        fh_solve_stats ss{};
        bool hDone = false;
        const int fin = run_jacobi(c, o, e.csw.p, e.rsw.p, e.tsw.p, bufs, &ss, &hDone);
        if (fin == 1) e.Bsw0.swap(e.Bsw1);
        st.swSweeps = ss.iterations;
        st.swRefreshes = ss.refreshes;
        st.swMaxDelta = ss.maxDelta;
        if (!fromG && !hDone) {
            if (tab) optics_arrays();  // the incident sweep reads the D array
            run_incident(c, ldb, e.Bsw0.p, e.Dsw.p, e.Hsw.p);
        }
        const uint32_t nbk = (S + 15u) / 16u;  // 16 sides per block
        e.dpart.ensure(size_t(nbk) * 4u);
        AbsorbTables tb{e.material.p, e.swRtab.p, e.swTtab.p, e.swDir.p, e.dni.p, c->gacc.p};
        if (fromG)
            hipLaunchKernelGGL((k_absorb<true, true>), dim3(nbk), dim3(256), 0, c->stream, S, nb, ldb, e.rsw.p,
                               e.tsw.p, e.Hsw.p, e.Dsw.p, e.Bsw0.p, c->fsky.p, c->skyv.p, e.area.p, e.absorbed.p,
                               e.incident.p, e.dpart.p, tb);
        else if (tab)
            hipLaunchKernelGGL((k_absorb<true, false>), dim3(nbk), dim3(256), 0, c->stream, S, nb, ldb, e.rsw.p,
                               e.tsw.p, e.Hsw.p, e.Dsw.p, e.Bsw0.p, c->fsky.p, c->skyv.p, e.area.p, e.absorbed.p,
                               e.incident.p, e.dpart.p, tb);
        else
            hipLaunchKernelGGL((k_absorb<false, false>), dim3(nbk), dim3(256), 0, c->stream, S, nb, ldb, e.rsw.p,
                               e.tsw.p, e.Hsw.p, e.Dsw.p, e.Bsw0.p, c->fsky.p, c->skyv.p, e.area.p, e.absorbed.p,
                               e.incident.p, e.dpart.p, tb);
        hipLaunchKernelGGL(k_reduce_sum4, dim3(1), dim3(1024), 0, c->stream, e.dpart.p, nbk, 0u, 3u, e.dsum.p);
        FH_HIP(hipGetLastError());
        double d4[4];
        FH_HIP(hipMemcpyAsync(d4, e.dsum.p, sizeof(double) * 4, hipMemcpyDeviceToHost, c->stream));
        FH_HIP(hipStreamSynchronize(c->stream));
        st.swInput = d4[0];
        st.swAbsorbed = d4[1];
        st.swEscaped = d4[2];
        e.swValid = true;
    } else {
        launch_fill_f32(c, e.absorbed.p, S, 0.0f);
        launch_fill_f32(c, e.incident.p, S, 0.0f);
        if (e.swNb) launch_fill_f32(c, e.Bsw0.p, size_t(S) * e.swLdb, 0.0f);
        e.swValid = false;
        e.swIncSerial = 0;
    }
    // --- temperature state ----------------------------------------------------
    EbConst k{};
    k.Ta = p.airTemperatureK;
    k.ea = p.vaporPressure;
    k.P = std::max(100.0, double(p.pressure));
    k.ha = 5.8 + 4.1 * std::sqrt(std::max(0.1, double(p.windSpeed)));
    k.rho = k.P * 100.0 / (287.05 * k.Ta);
    k.ra = k.rho * 1005.0 / std::max(0.1, k.ha);
    k.dTime = std::max(1.0, double(p.dTime));
    k.relax = p.temperatureRelaxation;
    k.adaptive = p.adaptiveRelaxation ? 1u : 0u;
    k.vegMethod = p.vegetationMethod;
    k.leafCount = e.leafCount;
    if (p.firstNode || !e.tValid) {
        launch_fill_f32(c, e.T.p, S, p.airTemperatureK);
        e.guessPrevValid = false;
    }
    copy_f32(c, e.T.p, e.Tprev.p, S);
    e.tValid = true;
    if (p.warmStart == FH_WARM_SURROGATE) {
        surrogate_guess(c, k, p.windSpeed, p.skyLongwave, p.airTemperatureK);
    } else {
        e.guessPrevValid = false;
        if (p.recordFeatures) eb_features(c, p.airTemperatureK, p.windSpeed, p.skyLongwave);
    }
    // --- coupling: fused Jacobi-Newton sweeps, FP32 refinement, final check ---
    if (!e.lwValid || p.zeroStart) st.lwSweeps += lw_solve(c, p, FH_PRECISION_FP32);
    float last[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    // The single-band LW sweep is bound by gather latency, not bytes: BF16
    // form factors only add coupling iterations there, so FH_PRECISION_BF16_DC
    // applies to the multi-band SW operator and the coupling runs in FP32
    // (FH_LW_BF16=1 forces BF16 coupled sweeps, a tuning knob).
    static const bool lwBf16 = [] {
        const char* v = std::getenv("FH_LW_BF16");
        return v && std::atoi(v) != 0;
    }();
    const bool bf16 = p.precision == FH_PRECISION_BF16_DC && lwBf16;
    uint32_t iters = coupled_loop(c, p, k, bf16, std::max(1u, p.maxCoupling), last);
    float maxDT = last[2];
    // final LW solve with the converged temperatures and the residual check
    // (core:3834-3846), always with FP32 form factors: in BF16 mode this is
    // the defect check, and a failed check continues with FP32 sweeps
    st.lwFinalSweeps = lw_solve(c, p, FH_PRECISION_FP32, &st.lwMaxDelta);
    st.lwSweeps += st.lwFinalSweeps;
    float dtFinal = 0.0f, maxEq = 0.0f;
    double rn = 0.0;
    float maxRes = 0.0f;
    newton(c, k, 0, dtFinal, maxEq, &rn, &maxRes);
    if (bf16 && !(maxDT <= p.temperatureTolerance && maxEq <= p.temperatureTolerance)) {
        iters += coupled_loop(c, p, k, false, std::max(1u, p.maxCoupling), last);
        maxDT = last[2];
        st.lwFinalSweeps = lw_solve(c, p, FH_PRECISION_FP32, &st.lwMaxDelta);
        st.lwSweeps += st.lwFinalSweeps;
        newton(c, k, 0, dtFinal, maxEq, &rn, &maxRes);
    }
    st.lwSweeps += iters;
    st.couplingIterations = iters;
    st.maxTemperatureDelta = maxDT;
    st.maxResidualEquivalent = maxEq;
    st.maxEnergyResidual = maxRes;
    st.converged = (maxDT <= p.temperatureTolerance && maxEq <= p.temperatureTolerance) ? 1u : 0u;
    st.netRadiation = rn;
    e.lastSun[0] = p.sun.x;
    e.lastSun[1] = p.sun.y;
    e.lastSun[2] = p.sun.z;
    e.lastTolerance = p.radiosityTolerance;
    e.lastMaxSweeps = p.maxSweeps;
    e.lastNodeValid = true;
    if (out) *out = st;
}

// Diffuse-only shortwave radiosity of the last node: the reference solves
// every window a second time with the direct irradiance removed
// (withoutDirectLight, core:1561-1569, core:3533-3536) for the light
// enhancement max(0, B - B_diffuse) of the result (core:1571-1583,
// core:4046-4047).  The node's optics (a Hapke ground at its sun) and sky
// radiance with the direct term zero, FP32 sweeps from zero to the node's
// radiosity tolerance; the node's own state (its radiosity, temperatures
// and fluxes) is left as it is.  sumHost [S]: the sum over the windows;
// bandHost [S][nb]: per window in window order (either may be null).
void eb_diffuse_solve(fh_context* c, float* sumHost, float* bandHost, fh_solve_stats* out) {
    EbState& e = c->eb;
    const uint32_t S = c->S;
    fh_solve_stats ss{};
    const uint32_t nb = e.swNb, ldb = e.swLdb;
    if (!nb || !e.swValid) {  // no shortwave at the last node: no window was solved
        if (sumHost) std::fill(sumHost, sumHost + S, 0.0f);
        if (bandHost) std::fill(bandHost, bandHost + size_t(S) * nb, 0.0f);
        if (out) *out = ss;
        return;
    }
    const size_t n = size_t(S) * ldb;
    e.Bdf0.ensure(n);
    e.Bdf1.ensure(n);
    e.dfSum.ensure(S);
    launch_fill_f32(c, e.dni.p, ldb, 0.0f);
    hipLaunchKernelGGL(k_sw_optics, blocks_for(n), dim3(256), 0, c->stream, S, nb, ldb, e.material.p, e.swRefl.p,
                       e.swTrans.p, e.swOrderD.p, c->facetNormal.p, e.sunlit.p, e.lastSun[0], e.lastSun[1],
                       e.lastSun[2], e.dni.p, e.leafCount, c->hapkeStrength, e.textured ? e.reflFactor.p : nullptr,
                       e.rsw.p, e.tsw.p, e.Dsw.p, e.swDir.p);
    FH_HIP(hipGetLastError());
    set_sky(c, e.lastDiffuse.data(), nb, ldb);
    launch_fill_f32(c, e.Hsw.p, n, 0.0f);
    launch_fold(c, e.Hsw.p, e.rsw.p, e.tsw.p, e.Dsw.p, ldb, e.csw.p);
    launch_fill_f32(c, e.Bdf0.p, n, 0.0f);
    JacobiOptions o;
    o.ldb = ldb;
    o.precision = FH_PRECISION_FP32;
    o.maxIter = e.lastMaxSweeps;
    o.tolRel = e.lastTolerance;
    if (e.swRhoValid) o.rho2 = e.swRho2.p;  // the operator (r, t) of the node
    float* bufs[2] = {e.Bdf0.p, e.Bdf1.p};
    const int fin = run_jacobi(c, o, e.csw.p, e.rsw.p, e.tsw.p, bufs, &ss, nullptr);
    const float* B = fin == 1 ? e.Bdf1.p : e.Bdf0.p;
    // the node's direct normal irradiance back in place
    FH_HIP(hipMemcpyAsync(e.dni.p, e.lastDni.data(), sizeof(float) * ldb, hipMemcpyHostToDevice, c->stream));
    if (sumHost) {
        hipLaunchKernelGGL(k_band_sum, blocks_for(S), dim3(256), 0, c->stream, B, S, nb, ldb, e.dfSum.p);
        FH_HIP(hipGetLastError());
        std::vector<double> sum(S);
        FH_HIP(hipMemcpyAsync(sum.data(), e.dfSum.p, sizeof(double) * S, hipMemcpyDeviceToHost, c->stream));
        FH_HIP(hipStreamSynchronize(c->stream));
        for (uint32_t s = 0; s < S; ++s) sumHost[s] = float(sum[s]);
    }
    if (bandHost) {
        std::vector<float> all(n);
        FH_HIP(hipMemcpyAsync(all.data(), B, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
        FH_HIP(hipStreamSynchronize(c->stream));
        for (uint32_t s = 0; s < S; ++s)
            for (uint32_t j = 0; j < nb; ++j) bandHost[size_t(s) * nb + e.swOrder[j]] = all[size_t(s) * ldb + j];
    }
    FH_HIP(hipStreamSynchronize(c->stream));
    if (out) *out = ss;
}

}  // namespace fh

using fh::fail;

namespace {
template <class Fn>
int guarded_eb(fh_context* c, Fn&& body) {
    if (!c) return FH_INVALID_ARGUMENT;
    try {
        FH_HIP(hipSetDevice(c->device));
        body();
        c->err.clear();
        return FH_OK;
    } catch (const fh::Error& e) {
        c->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        c->err = e.what();
        return FH_RUNTIME_ERROR;
    }
}
}  // namespace

extern "C" {

// scene.backgroundAngularStrength (core:2904-2910), clamped to [0, 1]
int fh_eb_set_background_angular_strength(fh_context* c, float strength) {
    return guarded_eb(c, [&] {
        if (!std::isfinite(strength)) fail(FH_INVALID_ARGUMENT, "background angular strength must be finite");
        c->hapkeStrength = std::min(std::max(strength, 0.0f), 1.0f);
        fh::EbState& e = c->eb;
        if (e.ready) {
            hipLaunchKernelGGL(fh::k_lw_setup, fh::blocks_for(c->S), dim3(256), 0, c->stream, c->S, e.material.p,
                               e.mats.p, e.leafCount, c->facetNormal.p, c->hapkeStrength, e.rlw.p, e.tlw.p,
                               e.eps.p);
            FH_HIP(hipGetLastError());
            FH_HIP(hipStreamSynchronize(c->stream));
            e.swRhoValid = false;
        }
    });
}

int fh_eb_set_scene(fh_context* c, uint32_t leafFacetCount, const uint32_t* facetMaterial, uint32_t materialCount,
                    const fh_eb_material* materials, uint32_t swBands, const float* swReflectance,
                    const float* swTransmittance) {
    return guarded_eb(c, [&] {
        if (!c->hasGraph) fail(FH_LOGIC_ERROR, "buildVisibilityGraph() must be called before the energy balance");
        if (leafFacetCount > c->F) fail(FH_INVALID_ARGUMENT, "leafFacetCount exceeds the facet count");
        if (!facetMaterial || !materials || materialCount == 0u) fail(FH_INVALID_ARGUMENT, "materials must be given");
        if (swBands > 4096u || (swBands && (!swReflectance || !swTransmittance)))
            fail(FH_INVALID_ARGUMENT, "shortwave band tables must be given");
        for (uint32_t f = 0; f < c->F; ++f)
            if (facetMaterial[f] >= materialCount) fail(FH_INVALID_ARGUMENT, "facet material index out of range");
        for (uint32_t m = 0; m < materialCount; ++m) {
            const fh_eb_material& x = materials[m];
            if (!std::isfinite(x.tirReflectance) || !std::isfinite(x.tirTransmittance) ||
                !std::isfinite(x.surfaceResistance) || !std::isfinite(x.heatCapacityPerArea) ||
                !std::isfinite(x.convectiveScale) || x.soilTemperatureMethod < 0 || x.soilTemperatureMethod > 2)
                fail(FH_INVALID_ARGUMENT, "invalid energy-balance material");
        }
        for (size_t i = 0; i < size_t(materialCount) * swBands; ++i)
            if (!std::isfinite(swReflectance[i]) || !std::isfinite(swTransmittance[i]))
                fail(FH_INVALID_ARGUMENT, "shortwave optics must be finite");
        fh::eb_set_scene(c, leafFacetCount, facetMaterial, materialCount, materials, swBands, swReflectance,
                         swTransmittance);
    });
}

int fh_eb_run_node(fh_context* c, const fh_eb_node_params* p, const float* swDirectNormal, const float* swDiffuse,
                   fh_eb_node_stats* stats) {
    return guarded_eb(c, [&] {
        if (!c->eb.ready) fail(FH_LOGIC_ERROR, "fh_eb_set_scene() must be called before fh_eb_run_node()");
        if (!c->hasGraph || c->eb.facets != c->F)
            fail(FH_LOGIC_ERROR, "buildVisibilityGraph() and fh_eb_set_scene() must follow setGeometry()");
        if (!p) fail(FH_INVALID_ARGUMENT, "params must not be null");
        if (!std::isfinite(p->sun.x) || !std::isfinite(p->sun.y) || !std::isfinite(p->sun.z))
            fail(FH_INVALID_ARGUMENT, "sun direction must be finite");
        if (c->hapkeStrength > 0.0f && !(p->sun.x * p->sun.x + p->sun.y * p->sun.y + p->sun.z * p->sun.z > 0.0f))
            fail(FH_INVALID_ARGUMENT, "the Hapke ground needs a non-zero sun direction");
        if (c->eb.swNb && (!swDirectNormal || !swDiffuse)) fail(FH_INVALID_ARGUMENT, "shortwave inputs must be given");
        for (uint32_t b = 0; b < c->eb.swNb; ++b)
            if (!std::isfinite(swDirectNormal[b]) || !std::isfinite(swDiffuse[b]) || swDirectNormal[b] < 0.0f ||
                swDiffuse[b] < 0.0f)
                fail(FH_INVALID_ARGUMENT, "shortwave inputs must be finite and non-negative");
        if (!std::isfinite(p->airTemperatureK) || p->airTemperatureK < 150.0f || p->airTemperatureK > 400.0f)
            fail(FH_INVALID_ARGUMENT, "air temperature out of range");
        if (!std::isfinite(p->skyLongwave) || p->skyLongwave < 0.0f) fail(FH_INVALID_ARGUMENT, "sky longwave invalid");
        if (!(p->temperatureRelaxation > 0.0f && p->temperatureRelaxation <= 1.0f))
            fail(FH_INVALID_ARGUMENT, "temperature relaxation must be in (0,1]");
        if (!(p->temperatureTolerance > 0.0f)) fail(FH_INVALID_ARGUMENT, "temperature tolerance must be positive");
        if (p->precision > FH_PRECISION_BF16_F || p->warmStart > FH_WARM_SURROGATE)
            fail(FH_INVALID_ARGUMENT, "unknown precision or warm-start mode");
        fh::eb_run_node(c, *p, swDirectNormal, swDiffuse, stats);
    });
}

int fh_eb_get_state(fh_context* c, float* T, float* absorbed, float* Blw, float* sunlit, float* Rn, float* Hs,
                    float* LE, float* G, float* swRadiosity) {
    return guarded_eb(c, [&] {
        fh::EbState& e = c->eb;
        if (!e.ready) fail(FH_LOGIC_ERROR, "fh_eb_set_scene() must be called before fh_eb_get_state()");
        const size_t S = c->S;
        auto get = [&](float* dst, const float* src, size_t n) {
            if (dst) FH_HIP(hipMemcpyAsync(dst, src, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        };
        get(T, e.T.p, S);
        get(absorbed, e.absorbed.p, S);
        get(Blw, e.Blw0.p, S);
        get(sunlit, e.sunlit.p, S);
        get(Rn, e.Rn.p, S);
        get(Hs, e.Hs.p, S);
        get(LE, e.LE.p, S);
        get(G, e.G.p, S);
        if (swRadiosity && e.swNb) {
            // [S][swNb] in window order (internal band b is window swOrder[b])
            const uint32_t nb = e.swNb;
            fh::DBuf<float> tmp;
            tmp.ensure(S * nb);
            fh::launch_unpad(c, e.Bsw0.p, nb, e.swLdb, tmp.p);
            std::vector<float> h(S * nb);
            FH_HIP(hipMemcpyAsync(h.data(), tmp.p, h.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
            FH_HIP(hipStreamSynchronize(c->stream));
            for (size_t s = 0; s < S; ++s)
                for (uint32_t b = 0; b < nb; ++b) swRadiosity[s * nb + e.swOrder[b]] = h[s * nb + b];
        }
        FH_HIP(hipStreamSynchronize(c->stream));
    });
}

int fh_eb_get_photovoltaic(fh_context* c, float* incidentShortwave, float* absorbedShortwave, float* power) {
    return guarded_eb(c, [&] {
        fh::EbState& e = c->eb;
        if (!e.ready) fail(FH_LOGIC_ERROR, "fh_eb_set_scene() must be called before fh_eb_get_photovoltaic()");
        const size_t S = c->S;
        auto get = [&](float* dst, const float* src) {
            if (dst) FH_HIP(hipMemcpyAsync(dst, src, S * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        };
        get(incidentShortwave, e.incident.p);
        get(absorbedShortwave, e.absorbed.p);
        get(power, e.pvPower.p);
        FH_HIP(hipStreamSynchronize(c->stream));
    });
}

int fh_eb_set_temperature(fh_context* c, const float* T) {
    return guarded_eb(c, [&] {
        if (!c->eb.ready) fail(FH_LOGIC_ERROR, "fh_eb_set_scene() must be called before fh_eb_set_temperature()");
        if (!T) fail(FH_INVALID_ARGUMENT, "temperature must not be null");
        for (uint32_t s = 0; s < c->S; ++s)
            if (!std::isfinite(T[s]) || T[s] < 1.0f || T[s] > 1000.0f) fail(FH_INVALID_ARGUMENT, "temperature out of range");
        FH_HIP(hipMemcpyAsync(c->eb.T.p, T, sizeof(float) * c->S, hipMemcpyHostToDevice, c->stream));
        FH_HIP(hipStreamSynchronize(c->stream));
        c->eb.tValid = true;
        c->eb.guessPrevValid = false;  // T no longer follows the warm-started chain
    });
}

int fh_eb_set_reflectance_factors(fh_context* c, const float* factors) {
    return guarded_eb(c, [&] {
        fh::EbState& e = c->eb;
        if (!e.ready || e.facets != c->F)
            fail(FH_LOGIC_ERROR, "fh_eb_set_scene() must be called before fh_eb_set_reflectance_factors()");
        if (factors)
            for (uint32_t f = 0; f < c->F; ++f)
                if (!std::isfinite(factors[f]) || factors[f] < 0.0f)
                    fail(FH_INVALID_ARGUMENT, "reflectance factors must be finite and non-negative");
        e.textured = factors != nullptr;
        if (factors) {
            e.reflFactor.ensure(c->F);
            FH_HIP(hipMemcpyAsync(e.reflFactor.p, factors, sizeof(float) * c->F, hipMemcpyHostToDevice, c->stream));
            FH_HIP(hipStreamSynchronize(c->stream));
        }
        e.swRhoValid = false;  // the operator's reflectances changed
    });
}

int fh_eb_diffuse_radiosity(fh_context* c, float* radiosity, float* bandRadiosity, fh_solve_stats* stats) {
    return guarded_eb(c, [&] {
        if (!c->eb.ready || c->eb.facets != c->F || !c->hasGraph)
            fail(FH_LOGIC_ERROR, "fh_eb_set_scene() must be called before fh_eb_diffuse_radiosity()");
        if (!c->eb.lastNodeValid) fail(FH_LOGIC_ERROR, "fh_eb_run_node() must be called before fh_eb_diffuse_radiosity()");
        fh::eb_diffuse_solve(c, radiosity, bandRadiosity, stats);
    });
}

}  // extern "C"
