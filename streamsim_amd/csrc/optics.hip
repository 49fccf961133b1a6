// optics.hip -- per-side, per-band optics of FacetRT's band solves built on
// the device: makeSpectralOptics (facetrtcore.cpp:1057-1126, the scene's
// bands with the materials' sunlit / shaded temperatures) and
// makeStateSpectralOptics (facetrtcore.cpp:1135-1199, the energy balance's
// sensor bands at the node's side temperatures).
//
// One thread per (side, band) of the [S][nb] arrays, band fastest: a side's
// facet material, texture factor, normal and sunlit fraction are shared by
// its nb threads (L1 hits), the per-band values by every side.  The inputs
// are S + F + materials x bands floats; the outputs 16 bytes per side and
// band, written once (an HBM-bound fill).  Compiled without FMA contraction,
// so the float products match the host formulas term by term.
#include <hip/hip_runtime.h>

#include <cmath>

#include "engine.hpp"
#include "hapke.hpp"
#include "solver.hpp"

namespace fh {

namespace {

// planckRadiance (core:524-536): W m-2 sr-1 um-1, in FP64
__device__ float planck(float wavelength, float temperature) {
    const double c1 = 1.1910439340652e8, c2 = 14388.291040407;
    const double um = wavelength > 50.0f ? double(wavelength) / 1000.0 : double(wavelength);
    const double x = c2 / (fmax(1.0, double(temperature)) * fmax(1.0e-6, um));
    return float(c1 / (pow(um, 5.0) * expm1(x)));
}

__global__ __launch_bounds__(256) void k_spectral_optics(SpectralOpticsArgs a) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(a.S) * a.nb) return;
    const uint32_t s = uint32_t(i / a.nb), b = uint32_t(i - size_t(s) * a.nb);
    const uint32_t f = s >> 1, m = a.mat[f];
    const bool object = f < a.leafCount;
    const float wl = a.wavelength[b];
    // FH_OPTICS_SCENE: thermal above 2500 (nm); FH_OPTICS_STATE:
    // isOpticalWavelength (core:1127-1133), micrometres when <= 2.5
    const float nm = wl <= 2.5f ? wl * 1000.0f : wl;
    const bool thermal = a.mode == FH_OPTICS_SCENE ? wl > 2500.0f : !(nm <= 2500.0f);
    const float meanR = a.reflectance[size_t(m) * a.nb + b];
    float rr = hapke::clampf(__fmul_rn(meanR, thermal || !a.factor ? 1.0f : a.factor[f]), 0.0f, 1.0f);
    const float tt = hapke::clampf(a.transmittance[size_t(m) * a.nb + b], 0.0f, 1.0f - rr);
    const float4 n = a.nrm[f];
    const float sg = (s & 1u) ? -1.0f : 1.0f;
    const float sn[3] = {__fmul_rn(n.x, sg), __fmul_rn(n.y, sg), __fmul_rn(n.z, sg)};
    float eps = fmaxf(0.0f, 1.0f - rr - tt);
    if (!object && thermal) {
        eps = hapke::emissivity(eps, fabsf(sn[1]), a.strength);
        rr = hapke::clampf(1.0f - tt - eps, 0.0f, 1.0f);
    } else if (!object) {
        const float sun[3] = {a.sx, a.sy, a.sz};
        rr = hapke::reflectance(rr, sun, sn, sn, a.strength);
    }
    a.r[i] = rr;
    a.t[i] = fminf(tt, 1.0f - rr);
    if (thermal) {
        float e;
        if (a.mode == FH_OPTICS_SCENE) {
            const float ill = hapke::clampf(a.sunlit[s], 0.0f, 1.0f);
            e = __fmul_rn(eps, __fadd_rn(__fmul_rn(ill, planck(wl, a.Tsun[m])),
                                         __fmul_rn(1.0f - ill, planck(wl, a.Tshade[m]))));
        } else {
            e = __fmul_rn(eps, planck(wl, hapke::clampf(a.T[s], 1.0f, 2000.0f)));
        }
        a.e[i] = e;
        a.D[i] = 0.0f;
    } else {
        const float cosv =
            fmaxf(0.0f, __fadd_rn(__fadd_rn(__fmul_rn(sn[0], a.sx), __fmul_rn(sn[1], a.sy)), __fmul_rn(sn[2], a.sz)));
        // the scene's bands take sunlit as given, the state's clamped (core:1119, 1188)
        const float lit = a.mode == FH_OPTICS_SCENE ? a.sunlit[s] : hapke::clampf(a.sunlit[s], 0.0f, 1.0f);
        a.e[i] = 0.0f;
        a.D[i] = __fmul_rn(__fmul_rn(lit, cosv), a.dni[b]);
    }
}

}  // namespace

void launch_spectral_optics(fh_context* c, const SpectralOpticsArgs& a) {
    const size_t n = size_t(a.S) * a.nb;
    hipLaunchKernelGGL(k_spectral_optics, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, c->stream, a);
    FH_HIP(hipGetLastError());
}

}  // namespace fh
