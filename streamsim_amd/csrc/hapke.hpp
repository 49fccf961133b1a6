// hapke.hpp -- the reference's Hapke background model (facetrtcore.cpp:
// 546-605), for host code (g++) and device code (hipcc) alike.
//
// Ground sides only.  Shortwave: the directional reflectance of a particulate
// surface (isotropic single scattering with an opposition surge, the
// two-stream H function for multiple scattering), single-scattering albedo
// from the Lambertian value, normalised by its value at normal incidence
// without the surge, blended with the Lambertian value by the strength.
// Longwave: the emissivity's angular factor of the same H function at
// cosine |n_y|, normalised to the normal direction, blended the same way.
// At strength 0 the shortwave value is the Lambertian one clamped to 0.999
// and the longwave value the Lambertian one.
#pragma once

#include <cmath>

#ifdef __HIPCC__
#define FH_HAPKE_FN __host__ __device__ inline
#else
#define FH_HAPKE_FN inline
#endif

namespace fh::hapke {

constexpr float kPi = 3.14159265358979323846f;

FH_HAPKE_FN float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
FH_HAPKE_FN float dot3(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// unit vector (the scene's vectors are never degenerate here: normals are
// checked at setGeometry and the sun direction is a unit vector)
FH_HAPKE_FN void unit3(const float* a, float* o) {
    const float len = sqrtf(dot3(a, a));
    const float inv = 1.0f / len;
    o[0] = a[0] * inv;
    o[1] = a[1] * inv;
    o[2] = a[2] * inv;
}

// two-stream H function of the single-scattering albedo w (core:546-550)
FH_HAPKE_FN float h_function(float mu, float w) {
    const float g = sqrtf(fmaxf(1.0f - w, 1.0e-6f));
    return (1.0f + 2.0f * mu) / (1.0f + 2.0f * mu * g);
}

// bidirectional kernel, surge amplitude 1 and width 0.1 (core:552-563)
FH_HAPKE_FN float kernel(float mu0, float mu, float phase, float w) {
    const float surge = 1.0f / (1.0f + tanf(0.5f * phase) / 0.1f);
    const float multiple = h_function(mu0, w) * h_function(mu, w) - 1.0f;
    return w / (4.0f * kPi) * mu0 / fmaxf(mu0 + mu, 1.0e-5f) * (1.0f + surge + multiple);
}

// backgroundHapkeReflectance (core:565-587): reflectance of a side with
// normal `normal` lit from `incoming`, seen from `outgoing`
FH_HAPKE_FN float reflectance(float lambert, const float* incoming, const float* outgoing, const float* normal,
                              float strength) {
    float n[3], in[3], out[3];
    unit3(normal, n);
    unit3(incoming, in);
    unit3(outgoing, out);
    if (dot3(in, n) < 0.0f)
        for (int k = 0; k < 3; ++k) in[k] = in[k] * -1.0f;
    if (dot3(out, n) < 0.0f)
        for (int k = 0; k < 3; ++k) out[k] = out[k] * -1.0f;
    const float mu0 = fmaxf(dot3(n, in), 1.0e-4f);
    const float mu = fmaxf(dot3(n, out), 1.0e-4f);
    const float phase = acosf(clampf(dot3(in, out), -1.0f, 1.0f));
    const float base = clampf(lambert, 0.0f, 0.999f);
    const float onePlus = 1.0f + base;
    const float w = clampf(4.0f * base / (onePlus * onePlus), 1.0e-5f, 0.999f);
    const float directional = kernel(mu0, mu, phase, w);
    const float h1 = h_function(1.0f, w);
    const float normalIncidence = w / (8.0f * kPi) * h1 * h1;
    const float value = clampf(base * directional / fmaxf(normalIncidence, 1.0e-6f), 0.0f, 1.0f);
    return base + (value - base) * clampf(strength, 0.0f, 1.0f);
}

// infraredHapkeEmissivity (core:589-605) at cosine = |n_y| of the side
FH_HAPKE_FN float ir_factor(float mu, float w) {
    const float root = sqrtf(1.0f - w);
    return root * (1.0f + 2.0f * mu / 1.5f) / (1.0f + 2.0f * mu / 1.5f * root);
}

FH_HAPKE_FN float emissivity(float lambert, float cosine, float strength) {
    const float base = clampf(lambert, 0.0f, 1.0f);
    const float w = clampf(1.0f - base * base, 1.0e-5f, 0.999f);
    const float value =
        clampf(base * ir_factor(clampf(cosine, 1.0e-4f, 1.0f), w) / fmaxf(ir_factor(1.0f, w), 1.0e-6f), 0.0f, 1.0f);
    return base + (value - base) * clampf(strength, 0.0f, 1.0f);
}

}  // namespace fh::hapke
