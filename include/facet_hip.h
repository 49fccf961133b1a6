/*
 * facet_hip.h -- C-ABI of the MI355X facet radiosity / energy-balance engine.
 *
 * Drop-in boundary for the reference's device backend class
 * facetvk::FacetrtVulkan (models/histream/src/facetrt/facetrt_vulkan.h,
 * "vkh" below) and the host loops of models/histream/src/facetrt/
 * facetrtcore.cpp ("core").  Plain C types, pointers and sizes only.
 * Every entry point returns an fh_status; fh_last_error() holds the message.
 * The C++ shim (streamsim_amd/csrc/facet_backend.hpp) rethrows the codes as
 * the exception types the reference throws (invalid_argument, logic_error,
 * runtime_error, overflow_error).
 *
 * Threading: a context is single-threaded and owns one HIP stream at the
 * default priority on device config.gpuIndex (vkh:75-76: non-copyable, one
 * queue).  Host-pointer entry points are synchronous like the reference's
 * submit + vkQueueWaitIdle; *_device entry points take HBM pointers.
 */
#ifndef FACET_HIP_H
#define FACET_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    FH_OK = 0,
    FH_INVALID_ARGUMENT = 1, /* std::invalid_argument in the reference   */
    FH_LOGIC_ERROR = 2,      /* std::logic_error (call order)            */
    FH_RUNTIME_ERROR = 3,    /* std::runtime_error (device, overflow)    */
    FH_OVERFLOW_ERROR = 4    /* std::overflow_error (sizing)             */
} fh_status;

typedef struct fh_context fh_context;

/* facetvk::Vertex (vkh:15-18): non-indexed triangle list, 3 vertices per
 * facet carrying the same unit normal. */
typedef struct {
    float position[3];
    float normal[3];
} fh_vertex;

/* facetvk::Direction (vkh:22-26): points from the scene to the sun. */
typedef struct {
    float x, y, z;
} fh_direction;

/* facetvk::SurfaceOptics (vkh:30-35): one entry per side; side 2f is the
 * geometric-normal side, side 2f+1 the opposite side. */
typedef struct {
    float reflectance;
    float transmittance;
    float emission;
    float directIrradiance;
} fh_surface_optics;

/* facetvk::Config (vkh:37-47).  The raster fields are replaced by the
 * ray-cast form-factor fields. */
typedef struct {
    uint32_t raysPerSide;    /* cosine-weighted form-factor rays per side  */
    uint32_t sunRaysPerSide; /* sun visibility rays per side               */
    uint32_t seed;           /* counter-RNG seed                           */
    uint32_t gpuIndex;       /* HIP device ordinal                         */
    float rayEpsilon;        /* minimum hit distance (m)                   */
    uint32_t reorder;        /* 1: Morton-order facets inside the engine   */
    uint32_t reserved[2];
} fh_config;

/* facetvk::GraphDiagnostics (vkh:49-56) */
typedef struct {
    uint32_t facetCount;
    uint32_t surfaceCount;
    uint32_t directedEdgeCount;
    uint32_t fragmentOverflow; /* always 0: rays cannot overflow      */
    uint32_t hashOverflow;     /* always 0: no hash table            */
    float maxClosureError;
} fh_graph_diagnostics;

/* iterations / maxDelta of facetvk::SolveResult (vkh:58-62) */
typedef struct {
    uint32_t iterations;  /* Jacobi sweeps run                          */
    float maxDelta;       /* max |B_new - B_old| of the last sweep      */
    float maxRelDelta;    /* max |B_new - B_old| / |B_new| of the last  */
    uint32_t refreshes;   /* FP32 defect-correction passes (BF16 mode)  */
} fh_solve_stats;

typedef enum {
    FH_PRECISION_FP32 = 0,   /* FP32 F, FP32 accumulation                     */
    FH_PRECISION_BF16_DC = 1 /* BF16 F sweeps + FP32-F defect correction      */
} fh_precision;

const char *fh_status_string(int status);
const char *fh_last_error(const fh_context *ctx);

/* FacetrtVulkan() + initialize(Config, shaderDir) (vkh:78) */
int fh_create(const fh_config *config, fh_context **out);
/* destroy() + ~FacetrtVulkan() (vkh:100) */
void fh_destroy(fh_context *ctx);

/* setGeometry(vector<Vertex>) (vkh:79): validates (vk:425-443), uploads and
 * builds the BVH.  Drops any graph built before (vk:451-453). */
int fh_set_geometry(fh_context *ctx, const fh_vertex *vertices, uint32_t vertexCount);

/* buildVisibilityGraph(directions) (vkh:80): the form factors F as a CSR of
 * integer hit counts of config.raysPerSide seeded cosine-weighted rays per
 * side; denominators = raysPerSide, skyCounts = escaped rays. */
int fh_build_form_factors(fh_context *ctx, fh_graph_diagnostics *out);

/* Upload an integer-count visibility graph built elsewhere (the reference's
 * CSR + denominators + skyCounts, vk:979-1061) in place of
 * fh_build_form_factors. */
int fh_set_graph(fh_context *ctx, uint32_t directedEdgeCount, const uint32_t *rowOffsets,
                 const uint32_t *neighbors, const uint32_t *counts,
                 const uint32_t *denominators, const uint32_t *skyCounts,
                 fh_graph_diagnostics *out);

/* Download the current graph (caller order; rows sorted by neighbour). */
uint32_t fh_directed_edge_count(const fh_context *ctx);
int fh_get_graph(fh_context *ctx, uint32_t *rowOffsets, uint32_t *neighbors, uint32_t *counts,
                 uint32_t *denominators, uint32_t *skyCounts);

/* solve(optics, sky, iterations, relaxation) (vkh:82-85): zero start, fixed
 * sweeps.  radiosity receives surfaceCount floats. */
int fh_solve(fh_context *ctx, const fh_surface_optics *optics, float skyRadiosity,
             uint32_t iterations, float relaxation, float *radiosity, fh_solve_stats *stats);

/* solveAccelerated(...) (vkh:86-93): optional warm start, batches of
 * batchIterations sweeps, stop when the absolute maxDelta <= tolerance. */
int fh_solve_accelerated(fh_context *ctx, const fh_surface_optics *optics, float skyRadiosity,
                         const float *initialRadiosity, uint32_t maximumIterations,
                         uint32_t batchIterations, float tolerance, float relaxation,
                         float *radiosity, fh_solve_stats *stats);

/* incidentIrradiance(solution, optics, sky) (vkh:94-97) */
int fh_incident_irradiance(fh_context *ctx, const float *radiosity,
                           const fh_surface_optics *optics, float skyRadiosity,
                           float *irradiance);

/* computeSunlitFraction(direction) (vkh:98): fraction of
 * config.sunRaysPerSide seeded sample points of a side that see the sun. */
int fh_compute_sunlit_fraction(fh_context *ctx, fh_direction sun, float *fraction);

/* facetCount() / surfaceCount() (vkh:103-104) */
uint32_t fh_facet_count(const fh_context *ctx);
uint32_t fh_surface_count(const fh_context *ctx);

/* ---- multi-band extension (fuses the band loop of core:1690-1763) ------ */
/* Arrays are [surfaceCount][nb] (band fastest); sky is [nb]; init may be
 * NULL.  Runs until the largest relative change of any side and band is
 * <= tolRel (checked on every FP32 sweep) or maxIter sweeps. */
int fh_solve_bands(fh_context *ctx, uint32_t nb, const float *reflectance,
                   const float *transmittance, const float *emission, const float *direct,
                   const float *sky, const float *init, uint32_t maxIter, float tolRel,
                   float relaxation, uint32_t precision, float *radiosity,
                   fh_solve_stats *stats);

/* Same with every array already in HBM (device pointers; sky stays host). */
int fh_solve_bands_device(fh_context *ctx, uint32_t nb, const float *d_reflectance,
                          const float *d_transmittance, const float *d_emission,
                          const float *d_direct, const float *sky, const float *d_init,
                          uint32_t maxIter, float tolRel, float relaxation, uint32_t precision,
                          float *d_radiosity, fh_solve_stats *stats);

/* H = max(0, in(B) + D) per side and band (core:1531-1559, multi-band) */
int fh_incident_bands(fh_context *ctx, uint32_t nb, const float *radiosity, const float *direct,
                      const float *sky, float *irradiance);

/* One FP32 Jacobi sweep on device pointers (benchmark / roofline entry).
 * d_c, d_r, d_t, d_Bin, d_Bout are [surfaceCount][ldb]; ldb = nb rounded
 * up to a multiple of 4 (1 when nb == 1).  precision selects the F values
 * read (FP32 or BF16). */
int fh_sweep_device(fh_context *ctx, uint32_t nb, const float *d_c, const float *d_r,
                    const float *d_t, const float *d_Bin, float *d_Bout, uint32_t precision);
uint32_t fh_band_stride(uint32_t nb);

#ifdef __cplusplus
}
#endif
#endif
