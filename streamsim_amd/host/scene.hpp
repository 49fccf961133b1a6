// scene.hpp -- the facet scene facet_engine hands to the engine: one
// triangle per facet (three vertices carrying the facet normal), a material
// index per facet, the object facets first (leafFacetCount), then ground.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <vector>

#include "facet_backend.hpp"

namespace facetcli {

using Vec3 = std::array<float, 3>;

struct Scene {
    std::vector<facethip::Vertex> vertices;
    std::vector<uint32_t> material;
    uint32_t leafFacetCount = 0;
    float extentX = 1.0f, extentZ = 1.0f;  // ground extent (the periodic tile)
    float backgroundAngularStrength = 0.0f;  // Hapke ground (core:2904-2910); 0: Lambert
    // visibility raster of the reference's graph (visibilityRasterSize /
    // visibilityRasterLayers, core:159-160): 1024 pixels square and 64
    // fragments per pixel, 2048 and 16 (then adapted) for large projects
    uint32_t rasterSize = 1024u, rasterLayers = 64u;
    // per facet: factor on the material's shortwave reflectance from its
    // reflectance-variance texture (reflectanceTextureFactors, core:152);
    // empty: 1 everywhere
    std::vector<float> reflectanceFactor;
    uint32_t facetCount() const { return uint32_t(vertices.size() / 3); }
    float reflectanceFactorOf(uint32_t f) const { return f < reflectanceFactor.size() ? reflectanceFactor[f] : 1.0f; }
};

// appendTriangle (core:288-306): normal of cross(b - a, c - a) in float;
// degenerate triangles are dropped
inline bool appendTriangle(Scene& s, const Vec3& a, const Vec3& b, const Vec3& c, uint32_t mat) {
    const Vec3 u{b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const Vec3 v{c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const Vec3 n{u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
    const float dd = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    if (dd <= 1.0e-16f) return false;
    const float inv = 1.0f / std::sqrt(dd);
    for (const Vec3* p : {&a, &b, &c}) {
        facethip::Vertex vx;
        for (int k = 0; k < 3; ++k) {
            vx.position[k] = (*p)[k];
            vx.normal[k] = n[k] * inv;
        }
        s.vertices.push_back(vx);
    }
    s.material.push_back(mat);
    return true;
}

}  // namespace facetcli
