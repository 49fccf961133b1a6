// Compile check of include/facet_backend.hpp against the way the reference's
// host code drives facetvk::FacetrtVulkan (runGpu, facetrtcore.cpp:1625-1640,
// and the FacetEB set-up, facetrtcore.cpp:3329-3340), after the type swap
// INTEGRATION.md describes.  Built with -fsyntax-only by test_abi.py; the
// scene struct below stands in for the reference's Scene fields these
// call sites read.
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "facet_backend.hpp"

namespace facetvk = facethip;
using FacetrtVulkan = facethip::FacetrtHip;

namespace {

struct SceneFields {
    std::vector<facetvk::Vertex> vertices;
    uint32_t visibilityRasterSize{1024};
    uint32_t visibilityRasterLayers{64};
    uint32_t periodicNeighborCount{0};
    float periodicSizeX{1.0f};
    float periodicSizeZ{1.0f};
};

// the FacetRT GPU route: raster fields, validation off, periodic tiles
[[maybe_unused]] float facetrt_route(const SceneFields& scene, const std::string& shaderDirectory,
                                     const std::vector<facetvk::Direction>& directions,
                                     const facetvk::Direction& sun,
                                     const std::vector<facetvk::SurfaceOptics>& optics) {
    facetvk::Config config{};
    config.rasterWidth = scene.visibilityRasterSize;
    config.rasterHeight = scene.visibilityRasterSize;
    config.maxFragmentsPerPixel = scene.visibilityRasterLayers;
    config.enableValidation = false;
    config.periodicNeighborCount = scene.periodicNeighborCount;
    config.periodicSizeX = scene.periodicSizeX;
    config.periodicSizeZ = scene.periodicSizeZ;
    FacetrtVulkan model;
    model.initialize(config, shaderDirectory);
    model.setGeometry(scene.vertices);
    const facetvk::GraphDiagnostics graph = model.buildVisibilityGraph(directions);
    const std::vector<float> sunlit = model.computeSunlitFraction(sun);
    const facetvk::SolveResult first = model.solve(optics, 100.0f);
    const facetvk::SolveResult second = model.solveAccelerated(optics, 50.0f, &first.radiosity);
    const std::vector<float> incident = model.incidentIrradiance(second, optics, 50.0f);
    model.destroy();
    return graph.maxClosureError + float(graph.hashOverflow + graph.fragmentOverflow) + sunlit[0] + incident[0] +
           float(model.facetCount() + model.surfaceCount()) + float(model.initialized());
}

// the FacetEB set-up: the same fields plus the GPU index from the input
[[maybe_unused]] void faceteb_setup(const SceneFields& scene, float gpuTag) {
    facetvk::Config config{};
    config.rasterWidth = scene.visibilityRasterSize;
    config.rasterHeight = scene.visibilityRasterSize;
    config.maxFragmentsPerPixel = scene.visibilityRasterLayers;
    config.periodicNeighborCount = scene.periodicNeighborCount;
    config.periodicSizeX = scene.periodicSizeX;
    config.periodicSizeZ = scene.periodicSizeZ;
    config.gpuIndex = static_cast<uint32_t>(std::max(0, static_cast<int>(std::lround(gpuTag))));
    config.edgeHashCapacity = 0;
    FacetrtVulkan model;
    model.initialize(config, std::string{});
}

// defaults are the reference's
static_assert(facetvk::Config{}.rasterWidth == 512 && facetvk::Config{}.rasterHeight == 512, "raster default");
static_assert(facetvk::Config{}.maxFragmentsPerPixel == 32 && facetvk::Config{}.edgeHashCapacity == 0,
              "fragment and hash defaults");
static_assert(!facetvk::Config{}.enableValidation, "validation default");
static_assert(facetvk::Config{}.visibilityMethod == facethip::VisibilityMethod::Raster,
              "a default Config rasterises, as the reference does");

}  // namespace
