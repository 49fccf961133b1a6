// facet_backend.hpp -- C++ mirror of facetvk::FacetrtVulkan over the C-ABI.
//
// Same nested types, method names, argument meaning, defaults and exception
// classes as models/histream/src/facetrt/facetrt_vulkan.h:15-104, so the
// reference's host loops (runGpu, the FacetEB lambdas) compile against it by
// swapping the class name.  Header-only; link libfacet_hip.so.
//
// Config keeps the reference's fields, order and defaults (a default
// Config rasterises 512 x 512 pixels with 32 fragments per pixel, as the
// reference does); the fields after them select this engine's ray-cast form
// factors.  Differences a caller can observe:
//   * with Config::visibilityMethod == VisibilityMethod::RayCast,
//     buildVisibilityGraph() casts Config::raysPerSide seeded cosine-weighted
//     rays per side instead of rasterising the given directions (the
//     argument is accepted and only checked for being non-empty), and
//     computeSunlitFraction() samples Config::sunRaysPerSide points per side;
//   * edgeHashCapacity is accepted and unused (the graph is built from
//     sorted pair lists, not a hash table, so it cannot overflow), and
//     hashOverflow is always 0;
//   * enableValidation is accepted and unused (no Vulkan layers).
#pragma once

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "facet_hip.h"

namespace facethip {

struct Vertex {  // facetvk::Vertex (facetrt_vulkan.h:15-18)
    float position[3]{};
    float normal[3]{};
};

struct Direction {  // facetvk::Direction (facetrt_vulkan.h:22-26)
    float x{};
    float y{};
    float z{1.0f};
};

struct SurfaceOptics {  // facetvk::SurfaceOptics (facetrt_vulkan.h:30-35)
    float reflectance{};
    float transmittance{};
    float emission{};
    float directIrradiance{};
};

// How buildVisibilityGraph() and computeSunlitFraction() see the scene.
enum class VisibilityMethod : uint32_t {
    Raster = 0,   // the reference's A-buffer over the given directions
    RayCast = 1,  // seeded cosine-weighted form-factor rays and sun rays
};

struct Config {  // facetvk::Config (facetrt_vulkan.h:37-47), then the ray-cast fields
    uint32_t rasterWidth{512};
    uint32_t rasterHeight{512};
    uint32_t maxFragmentsPerPixel{32};
    uint32_t edgeHashCapacity{0};  // accepted, unused: no hash table
    bool enableValidation{false};  // accepted, unused
    uint32_t gpuIndex{0};
    uint32_t periodicNeighborCount{0};
    float periodicSizeX{1.0f};
    float periodicSizeZ{1.0f};
    // ray-cast form factors (this engine's extension)
    VisibilityMethod visibilityMethod{VisibilityMethod::Raster};
    uint32_t raysPerSide{256};
    uint32_t sunRaysPerSide{64};
    uint32_t seed{1};
    float rayEpsilon{1.0e-5f};
};

struct GraphDiagnostics {  // facetvk::GraphDiagnostics (facetrt_vulkan.h:49-56)
    uint32_t facetCount{};
    uint32_t surfaceCount{};
    uint32_t directedEdgeCount{};
    uint32_t fragmentOverflow{};
    uint32_t hashOverflow{};
    float maxClosureError{};
};

struct SolveResult {  // facetvk::SolveResult (facetrt_vulkan.h:58-62)
    std::vector<float> radiosity;
    uint32_t iterations{};
    float maxDelta{};
};

static_assert(sizeof(Vertex) == sizeof(fh_vertex), "Vertex layout");
static_assert(sizeof(SurfaceOptics) == sizeof(fh_surface_optics), "SurfaceOptics layout");

// Rethrows an fh_status as the exception class the reference throws.
inline void check(int status, const fh_context* ctx) {
    if (status == FH_OK) return;
    const std::string msg = fh_last_error(ctx);
    switch (status) {
        case FH_INVALID_ARGUMENT: throw std::invalid_argument(msg);
        case FH_LOGIC_ERROR: throw std::logic_error(msg);
        case FH_OVERFLOW_ERROR: throw std::overflow_error(msg);
        default: throw std::runtime_error(msg);
    }
}

class FacetrtHip {
public:
    FacetrtHip() = default;
    ~FacetrtHip() { destroy(); }

    FacetrtHip(const FacetrtHip&) = delete;
    FacetrtHip& operator=(const FacetrtHip&) = delete;

    void initialize(const Config& config, const std::string& /*shaderDirectory*/ = {}) {
        destroy();
        fh_config c{};
        c.raysPerSide = config.raysPerSide;
        c.sunRaysPerSide = config.sunRaysPerSide;
        c.seed = config.seed;
        c.gpuIndex = config.gpuIndex;
        c.rayEpsilon = config.rayEpsilon;
        c.periodicNeighborCount = config.periodicNeighborCount;
        c.periodicSizeX = config.periodicSizeX;
        c.periodicSizeZ = config.periodicSizeZ;
        check(fh_create(&c, &m_ctx), nullptr);
        m_config = config;
    }

    void setGeometry(const std::vector<Vertex>& vertices) {
        require();
        check(fh_set_geometry(m_ctx, reinterpret_cast<const fh_vertex*>(vertices.data()),
                              static_cast<uint32_t>(vertices.size())),
              m_ctx);
    }

    GraphDiagnostics buildVisibilityGraph(const std::vector<Direction>& directions) {
        require();
        if (directions.empty()) throw std::invalid_argument("At least one visibility direction is required");
        fh_graph_diagnostics d{};
        if (m_config.visibilityMethod == VisibilityMethod::Raster) {
            std::vector<fh_direction> dirs;
            dirs.reserve(directions.size());
            for (const Direction& v : directions) dirs.push_back({v.x, v.y, v.z});
            check(fh_build_raster_graph(m_ctx, dirs.data(), uint32_t(dirs.size()), m_config.rasterWidth,
                                        m_config.rasterHeight, m_config.maxFragmentsPerPixel, &d),
                  m_ctx);
        } else {
            check(fh_build_form_factors(m_ctx, &d), m_ctx);
        }
        return {d.facetCount, d.surfaceCount, d.directedEdgeCount, d.fragmentOverflow, d.hashOverflow,
                d.maxClosureError};
    }

    SolveResult solve(const std::vector<SurfaceOptics>& optics, float skyRadiosity, uint32_t iterations = 64,
                      float relaxation = 1.0f) {
        require();
        check_size(optics);
        SolveResult r;
        r.radiosity.resize(surfaceCount());
        fh_solve_stats st{};
        check(fh_solve(m_ctx, reinterpret_cast<const fh_surface_optics*>(optics.data()), skyRadiosity, iterations,
                       relaxation, r.radiosity.data(), &st),
              m_ctx);
        r.iterations = st.iterations;
        r.maxDelta = st.maxDelta;
        return r;
    }

    SolveResult solveAccelerated(const std::vector<SurfaceOptics>& optics, float skyRadiosity,
                                 const std::vector<float>* initialRadiosity = nullptr,
                                 uint32_t maximumIterations = 64, uint32_t batchIterations = 8,
                                 float tolerance = 1.0e-4f, float relaxation = 1.0f) {
        require();
        check_size(optics);
        if (initialRadiosity && initialRadiosity->size() != surfaceCount())
            throw std::invalid_argument("Initial radiosity must contain one finite non-negative value per surface");
        SolveResult r;
        r.radiosity.resize(surfaceCount());
        fh_solve_stats st{};
        check(fh_solve_accelerated(m_ctx, reinterpret_cast<const fh_surface_optics*>(optics.data()), skyRadiosity,
                                   initialRadiosity ? initialRadiosity->data() : nullptr, maximumIterations,
                                   batchIterations, tolerance, relaxation, r.radiosity.data(), &st),
              m_ctx);
        r.iterations = st.iterations;
        r.maxDelta = st.maxDelta;
        return r;
    }

    std::vector<float> incidentIrradiance(const SolveResult& solution, const std::vector<SurfaceOptics>& optics,
                                          float skyRadiosity) const {
        require();
        if (solution.radiosity.size() != surfaceCount() || optics.size() != surfaceCount())
            throw std::invalid_argument("Incident irradiance requires a completed solve and visibility graph");
        std::vector<float> out(surfaceCount());
        check(fh_incident_irradiance(m_ctx, solution.radiosity.data(),
                                     reinterpret_cast<const fh_surface_optics*>(optics.data()), skyRadiosity,
                                     out.data()),
              m_ctx);
        return out;
    }

    std::vector<float> computeSunlitFraction(const Direction& sunDirection) {
        require();
        // normalised like facetrt_vulkan.cpp:1068
        const float len = std::sqrt(sunDirection.x * sunDirection.x + sunDirection.y * sunDirection.y +
                                    sunDirection.z * sunDirection.z);
        if (!(len > 1.0e-8f)) throw std::runtime_error("Zero-length direction or degenerate triangle");
        std::vector<float> out(surfaceCount());
        if (m_config.visibilityMethod == VisibilityMethod::Raster) {
            // makeProjection takes the direction as given (vk:1071)
            const fh_direction raw{sunDirection.x, sunDirection.y, sunDirection.z};
            check(fh_compute_raster_sunlit_fraction(m_ctx, raw, m_config.rasterWidth, m_config.rasterHeight,
                                                    m_config.maxFragmentsPerPixel, out.data()),
                  m_ctx);
            return out;
        }
        const float inv = 1.0f / len;
        fh_direction d{sunDirection.x * inv, sunDirection.y * inv, sunDirection.z * inv};
        check(fh_compute_sunlit_fraction(m_ctx, d, out.data()), m_ctx);
        return out;
    }

    void destroy() {
        if (m_ctx) fh_destroy(m_ctx);
        m_ctx = nullptr;
    }

    [[nodiscard]] bool initialized() const noexcept { return m_ctx != nullptr; }
    [[nodiscard]] uint32_t facetCount() const noexcept { return m_ctx ? fh_facet_count(m_ctx) : 0u; }
    [[nodiscard]] uint32_t surfaceCount() const noexcept { return m_ctx ? fh_surface_count(m_ctx) : 0u; }
    [[nodiscard]] fh_context* handle() const noexcept { return m_ctx; }

private:
    void require() const {
        if (!m_ctx) throw std::logic_error("initialize() must be called first");
    }
    void check_size(const std::vector<SurfaceOptics>& optics) const {
        if (optics.size() != surfaceCount())
            throw std::invalid_argument("Optics array must contain exactly two sides per facet");
    }

    fh_context* m_ctx{nullptr};
    Config m_config{};
};

}  // namespace facethip
