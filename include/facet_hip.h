/*
 * facet_hip.h -- C-ABI of the MI355X facet radiosity / energy-balance engine.
 *
 * Drop-in boundary for the reference's device backend class
 * facetvk::FacetrtVulkan (models/histream/src/facetrt/facetrt_vulkan.h,
 * "vkh" below) and the host loops of models/histream/src/facetrt/
 * facetrtcore.cpp ("core").  Plain C types, pointers and sizes only.
 * Every entry point returns an fh_status; fh_last_error() holds the message.
 * The C++ shim (include/facet_backend.hpp) rethrows the codes as
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
 * ray-cast form-factor fields; the periodic fields keep their meaning
 * (sh/visibility.vert:15-35): the scene is repeated at the first
 * periodicNeighborCount tiles of the neighbour list (+-X, +-Z, corners,
 * then the ring at distance 2), offsets periodicSizeX / periodicSizeZ, and
 * rays see every copy (a hit on a copy counts for the original facet). */
typedef struct {
    uint32_t raysPerSide;    /* cosine-weighted form-factor rays per side  */
    uint32_t sunRaysPerSide; /* sun visibility rays per side               */
    uint32_t seed;           /* counter-RNG seed                           */
    uint32_t gpuIndex;       /* HIP device ordinal                         */
    float rayEpsilon;        /* minimum hit distance (m)                   */
    uint32_t periodicNeighborCount; /* 0 .. 20                             */
    float periodicSizeX;     /* tile period along X (m), > 0 when periodic */
    float periodicSizeZ;     /* tile period along Z (m)                    */
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
    FH_PRECISION_FP32 = 0,    /* FP32 F, FP32 accumulation                     */
    FH_PRECISION_BF16_DC = 1, /* BF16 F sweeps + FP32-F defect correction; from
                                 16 bands on: BF16 increments of B gathered   */
    FH_PRECISION_FP8_INC = 2, /* from 16 bands on: FP8 increments of B
                                 gathered (12-band chunks with a shared
                                 power-of-two exponent), FP32 weights,
                                 accumulated gathers and carries; below 16
                                 bands FP32                                   */
    FH_PRECISION_BF16_F = 3   /* BF16 F and a BF16 copy of B gathered until
                                 the BF16 operator's fixed point, then FP32
                                 sweeps to the FP32 one (defect correction),
                                 stopping test max(tolRel, 3e-7); below 16
                                 bands as FH_PRECISION_BF16_DC                */
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

/* buildVisibilityGraph(directions) (vkh:80, vk:890-1061) by the reference's
 * raster method, in place of fh_build_form_factors: per direction an
 * orthographic A-buffer of rasterWidth x rasterHeight pixels and at most
 * maxFragmentsPerPixel fragments per pixel (Config, vkh:38-40) over the
 * scene and its periodic copies; in each pixel the fragments in depth order
 * add to both sides' denominators, the first camera-facing side and the
 * last back side see the sky, and consecutive surfaces of different facets
 * exchange (both directions).  Fragment overflow fails with
 * FH_RUNTIME_ERROR (diagnostics still filled) and leaves no graph. */
int fh_build_raster_graph(fh_context *ctx, const fh_direction *directions, uint32_t directionCount,
                          uint32_t rasterWidth, uint32_t rasterHeight, uint32_t maxFragmentsPerPixel,
                          fh_graph_diagnostics *out);

/* computeSunlitFraction(direction) (vkh:98, vk:1064-1157) by the raster
 * method: per side, covered pixels where it is the nearest surface over
 * covered pixels (0 where it covers none). */
int fh_compute_raster_sunlit_fraction(fh_context *ctx, fh_direction sun, uint32_t rasterWidth,
                                      uint32_t rasterHeight, uint32_t maxFragmentsPerPixel,
                                      float *fraction);

/* facetCount() / surfaceCount() (vkh:103-104) */
uint32_t fh_facet_count(const fh_context *ctx);
uint32_t fh_surface_count(const fh_context *ctx);

/* ---- multi-band extension (fuses the band loop of core:1690-1763) ------ */
/* Arrays are [surfaceCount][nb] (band fastest); sky is [nb]; init may be
 * NULL.  Runs until, for every band, the largest relative change over the
 * sides in an FP32 sweep is <= tolRel (from 16 bands on, a band that meets
 * it stops iterating while the others continue), or maxIter sweeps.
 * FH_PRECISION_BF16_DC from 16 bands on gathers BF16 increments of the
 * radiosity with FP32 weights and accumulation, FH_PRECISION_FP8_INC FP8
 * increments (half the gathered bytes); both reach the FP32 fixed point,
 * and their stopping test uses max(tolRel, 3e-7).  From 16 bands on,
 * facets that are each other's strongest partner (at least 0.2 of a side's
 * row weight) are updated together with the entries between the two
 * solved exactly (block Jacobi sweep): same fixed point and stopping test
 * on the Jacobi change, fewer sweeps where leaves face each other closely. */
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

/* ---- band optics on the device -------------------------------------- */
/* The per-side, per-band optics the reference's host loops build before its
 * band solves: FH_OPTICS_SCENE is makeSpectralOptics (core:1057-1126: bands
 * above 2500 nm thermal, emission eps (f Planck(T_sunlit) + (1 - f)
 * Planck(T_shaded)) with f = clamp(sunlit, 0, 1), direct sunlit cos
 * directNormal), FH_OPTICS_STATE is makeStateSpectralOptics
 * (core:1135-1199: isOpticalWavelength, emission eps Planck(clamp(T_side,
 * 1, 2000)), direct clamp(sunlit, 0, 1) cos directNormal).  Both: r =
 * clamp(r_material factor, 0, 1) on optical bands (factor 1 on thermal
 * ones), t = clamp(t_material, 0, 1 - r); ground sides (facets from
 * leafFacetCount) take the Hapke reflectance (optical) or emissivity
 * (thermal) at backgroundAngularStrength. */
enum { FH_OPTICS_SCENE = 0, FH_OPTICS_STATE = 1 };
typedef struct {
    uint32_t mode;                  /* FH_OPTICS_SCENE or FH_OPTICS_STATE                */
    uint32_t bands;                 /* nb, 1..4096                                        */
    uint32_t materialCount;
    uint32_t leafFacetCount;
    const uint32_t *facetMaterial;  /* [facetCount]                                       */
    const float *reflectance;       /* [materialCount][bands]: thermal bands ref_TIR      */
    const float *transmittance;     /* [materialCount][bands]: thermal bands tau_TIR      */
    const float *wavelength;        /* [bands], nm (micrometres when <= 2.5)              */
    const float *reflectanceFactor; /* [facetCount] texture factors, or NULL              */
    float backgroundAngularStrength;
    fh_direction sun;               /* unit vector towards the sun                        */
    const float *sunlit;            /* [surfaceCount]                                     */
    const float *directNormal;      /* [bands] direct normal irradiance (optical bands)   */
    const float *temperature;       /* STATE: [surfaceCount] side temperatures, K         */
    const float *sunlitTemperature; /* SCENE: [materialCount], K                          */
    const float *shadedTemperature; /* SCENE: [materialCount], K                          */
} fh_spectral_optics;

/* The optics as host arrays [surfaceCount][bands] (any output may be NULL). */
int fh_spectral_optics_build(fh_context *ctx, const fh_spectral_optics *optics, float *reflectance,
                             float *transmittance, float *emission, float *direct);

/* The optics built on the device and solved as fh_solve_bands does, with no
 * per-band arrays crossing the bus: the reference's band loop with its host
 * optics (core:1690-1763, 3875-3907) in one call.  radiosity is
 * [surfaceCount][bands] on the host. */
int fh_solve_spectral(fh_context *ctx, const fh_spectral_optics *optics, const float *sky,
                      const float *init, uint32_t maxIter, float tolRel, float relaxation,
                      uint32_t precision, float *radiosity, fh_solve_stats *stats);

/* H = max(0, in(B) + D) per side and band (core:1531-1559, multi-band) */
int fh_incident_bands(fh_context *ctx, uint32_t nb, const float *radiosity, const float *direct,
                      const float *sky, float *irradiance);

/* One Jacobi sweep on device pointers, launched as a single kernel with no
 * convergence bookkeeping and no synchronisation (benchmark / roofline entry).
 * d_c, d_r, d_t, d_Bin, d_Bout are [surfaceCount][ldb] with ldb =
 * fh_band_stride(nb): nb rounded up to a multiple of 4 (1 when nb == 1).
 * precision selects the F values read (FP32 or BF16). */
int fh_sweep_device(fh_context *ctx, uint32_t nb, const float *d_c, const float *d_r,
                    const float *d_t, const float *d_Bin, float *d_Bout, uint32_t precision);
uint32_t fh_band_stride(uint32_t nb);

/* fh_sweep_device for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * the same sweep and results, with the per-block residual partials in a
 * fixed 2 MiB scratch array of the library instead of a context buffer that
 * grows on demand; the call makes no device allocation.  No counterpart in
 * the reference.  Limit: a sweep needing more than 262144 block partials
 * (about 2 M facets at the usual team sizes) fails with FH_INVALID_ARGUMENT
 * before any launch.  Every context on a device shares that device's
 * scratch, which keeps the last call's contents: calls that overlap on one
 * device (other streams or threads) race on it. */
int fh_sweep_device_noalloc(fh_context *ctx, uint32_t nb, const float *d_c, const float *d_r,
                            const float *d_t, const float *d_Bin, float *d_Bout, uint32_t precision);

/* The context's HIP stream (hipStream_t), for event timing by the caller. */
void *fh_stream(const fh_context *ctx);

/* Average duration (microseconds, HIP events on the context's stream) of
 * `reps` back-to-back SpMM sweeps over the current graph with nb bands, on
 * engine-owned HBM inputs, as the solves run them: the sweep of
 * fh_sweep_device, except that at band-parallel strides (nb >= 16) the
 * coupled facet pairs are updated as blocks by a second kernel on a second
 * stream beside the first (joined before the next sweep), and BF16
 * precision times the increment sweep the BF16 solves run (BF16 increments
 * of B gathered with FP32 weights, FP32 accumulated gathers and carries),
 * FH_PRECISION_FP8_INC the FP8 increment sweep and FH_PRECISION_BF16_F the
 * sweep of its BF16 phase (BF16 F, BF16 copy of B gathered).
 * materialTables = 1 (band-parallel strides): the sweep in the form the
 * energy balance's shortwave solves run it -- r and t from per-material
 * tables, c and D from per-side direct factors, Chebyshev weights (the
 * previous iterate read) -- instead of streamed [S][ldb] c, r, t arrays. */
int fh_benchmark_sweep(fh_context *ctx, uint32_t nb, uint32_t precision, uint32_t materialTables,
                       uint32_t reps, float *microsecondsPerSweep);
/* Average duration (microseconds) of a plain device-to-device copy kernel of
 * `bytes` bytes (reads `bytes`, writes `bytes`) on the context's stream: the
 * calibration point for hardware traffic counters and achievable bandwidth. */
int fh_benchmark_copy(fh_context *ctx, uint64_t bytes, uint32_t reps, float *microsecondsPerCopy);
/* Waits for all work queued on the context's stream. */
int fh_synchronize(fh_context *ctx);


/* ---- FacetEB: coupled shortwave / longwave radiosity + energy balance ---- */

/* Energy-balance fields of SurfaceParameters (core:113-143). */
typedef struct {
    float tirReflectance;
    float tirTransmittance;
    float surfaceResistance;   /* s m-1 */
    float specificHeat;        /* soil c, J kg-1 K-1 */
    float density;             /* soil rho, kg m-3 */
    float thermalConductivity; /* soil k, W m-1 K-1 */
    float soilMoisture;
    float heatCapacityPerArea; /* leaf C_A, J m-2 K-1 */
    float convectiveScale;
    int32_t soilTemperatureMethod; /* 0: G = 0.35 Rn, 1: k dT, 2: 1.5 k dT */
    int32_t vegetation;            /* 1: leaf transpires (LE on) */
    int32_t photovoltaic;          /* 1: object facets generate electric power (core:3655-3663) */
    float pvEta25;                 /* efficiency at 25 C */
    float pvGamma;                 /* efficiency temperature coefficient, K-1 */
    float pvBifaciality;           /* weight of the back side's incident shortwave */
} fh_eb_material;

/* Scene description for runFacetEBCore (core:3164-4154): facets
 * [0, leafFacetCount) are leaves (one temperature for both sides), the rest
 * ground (one temperature per side).  swReflectance / swTransmittance are
 * [materialCount][swBands]: the window-centre optics of each shortwave band
 * (makeEnergyShortwaveOptics, core:952-1003). */
/* scene.backgroundAngularStrength (core:2904-2910, Hapke background,
 * core:546-605): strength in [0, 1] (clamped) of the ground's directional
 * shortwave reflectance and angular longwave emissivity in the energy
 * balance; 0 (the default) keeps the ground Lambertian.  Kept across
 * fh_eb_set_scene calls. */
int fh_eb_set_background_angular_strength(fh_context *ctx, float strength);

int fh_eb_set_scene(fh_context *ctx, uint32_t leafFacetCount, const uint32_t *facetMaterial,
                    uint32_t materialCount, const fh_eb_material *materials, uint32_t swBands,
                    const float *swReflectance, const float *swTransmittance);

typedef enum {
    FH_WARM_PREVIOUS = 0, /* start each node from the previous node's state */
    FH_WARM_SURROGATE = 1 /* start the coupled fixed point from the MLP guess */
} fh_warm_start;

/* One meteorological node (core:3462-3846). */
typedef struct {
    fh_direction sun;          /* unit vector toward the sun              */
    float airTemperatureK;     /* clamp(Tair + 273.15, 220, 340)           */
    float vaporPressure;       /* hPa                                      */
    float pressure;            /* hPa                                      */
    float windSpeed;           /* m s-1                                    */
    float skyLongwave;         /* W m-2                                    */
    float dTime;               /* s                                        */
    float temperatureTolerance;/* K (also the residual-equivalent bound)  */
    float temperatureRelaxation;
    uint32_t maxCoupling;      /* coupling iteration budget                */
    uint32_t vegetationMethod; /* 0 | 1                                    */
    float radiosityTolerance;  /* relative change, SW and LW solves        */
    uint32_t maxSweeps;        /* per radiosity solve                      */
    uint32_t precision;        /* fh_precision of the SW solves from 33 windows on (BF16
                                  or FP8 increments; FP32 rows then span more than one
                                  128-byte line); up to 32 windows and for the LW: FP32 */
    uint32_t warmStart;        /* fh_warm_start                            */
    uint32_t firstNode;        /* 1: T = airTemperatureK (core:3470-3479)  */
    uint32_t zeroStart;        /* 1: every solve starts from B = 0         */
    uint32_t recordFeatures;   /* 1: keep the surrogate features (training) */
    uint32_t adaptiveRelaxation; /* 1: full Newton steps; temperatureRelaxation only for a
                                    unit whose step reverses direction (same fixed point;
                                    the stop test then bounds the full step) */
} fh_eb_node_params;

typedef struct {
    uint32_t couplingIterations;
    uint32_t converged;
    uint32_t swSweeps;
    uint32_t lwSweeps;
    float maxTemperatureDelta;
    float maxResidualEquivalent;
    uint32_t swRefreshes;
    float maxEnergyResidual;  /* max |energy residual| of the final check, W m-2 */
    double swInput;      /* sum A (D + fsky S) over sides and bands, W     */
    double swAbsorbed;   /* sum A alpha H                                  */
    double swEscaped;    /* sum A fsky B                                   */
    double netRadiation; /* sum A Rn over energy-balance sides             */
    float swMaxDelta;       /* max |dB| of the shortwave solve's last sweep   */
    uint32_t lwFinalSweeps; /* sweeps of the final LW solve (core:3834-3846)  */
    float lwMaxDelta;       /* max |dB| of its last sweep                     */
} fh_eb_node_stats;

/* Runs one node fully on the device: sunlit rays, SW optics, multi-band SW
 * solve, absorbed SW, then the coupling loop (LW source from T in FP64, LW
 * solve, incident LW, FP64 relaxed Newton per facet) until
 * max dT <= tol and max |residual|/max(1, response) <= tol, and the final
 * LW solve + residual (core:3834-3846).  swDirectNormal / swDiffuse hold
 * one value per SW band (W m-2). */
int fh_eb_run_node(fh_context *ctx, const fh_eb_node_params *params, const float *swDirectNormal,
                   const float *swDiffuse, fh_eb_node_stats *stats);

/* Downloads the node state; any pointer may be NULL.  Per side unless noted:
 * T, absorbed SW, LW radiosity, sunlit fraction, net radiation, sensible,
 * latent and storage heat; swRadiosity is [surfaceCount][swBands]. */
int fh_eb_get_state(fh_context *ctx, float *temperature, float *absorbedShortwave,
                    float *longwaveRadiosity, float *sunlit, float *netRadiation, float *sensibleHeat,
                    float *latentHeat, float *storageHeat, float *swRadiosity);

/* Per side, after fh_eb_run_node: the shortwave incident (sum of the windows'
 * incident irradiance, core:3544), the absorbed shortwave and the electric
 * power of photovoltaic object facets (W m-2, the same on both sides of a
 * facet, 0 elsewhere; the reference's photovoltaicPower, core:3657-3667).
 * Any pointer may be NULL.  No counterpart in the reference's backend API:
 * the reference keeps these arrays in its host loop (core:3485). */
int fh_eb_get_photovoltaic(fh_context *ctx, float *incidentShortwave, float *absorbedShortwave,
                           float *power);
/* Replaces the temperature state (surfaceCount values). */
int fh_eb_set_temperature(fh_context *ctx, const float *temperature);

/* Per-facet factors on the shortwave reflectance of the energy balance
 * (reflectanceTextureFactor, core:848-851, applied in
 * makeEnergyShortwaveOptics, core:965-972): a side's window reflectance is
 * clamp(r_material * factor, 0, 1), the transmittance then limited to
 * 1 - r.  factors: facetCount values, finite and >= 0; NULL removes them.
 * Call after fh_eb_set_scene, which clears them.  No counterpart in the
 * reference's backend API: its host loop builds the optics with the
 * factors (core:952-1003). */
int fh_eb_set_reflectance_factors(fh_context *ctx, const float *factors);

/* The last node's shortwave windows solved again with the direct
 * irradiance removed: the reference's per-window diffuse-only solve,
 * solveRadiation(withoutDirectLight(windowOptics), windowDiffuse, ...)
 * (core:3533-3536, core:1561-1569), whose sum over the windows gives the
 * result's light enhancement max(0, B_SW - B_diffuse) (core:4046-4047).
 * The node's optics and sky radiance, FP32 sweeps from zero to its
 * radiosity tolerance; the node's state is not changed.  radiosity
 * [surfaceCount]: the sum over the windows; bandRadiosity
 * [surfaceCount][swBands]: per window (either may be NULL).  Zero when the
 * last node had no shortwave; FH_LOGIC_ERROR before the first node. */
int fh_eb_diffuse_radiosity(fh_context *ctx, float *radiosity, float *bandRadiosity, fh_solve_stats *stats);


/* ---- learned warm start (no counterpart in the reference) -------------- */
/* MLP 9 -> hidden -> hidden -> 1 with tanh-form GELU.  Parameter layout:
 * W1[hidden][9], b1[hidden], W2[hidden][hidden], b2[hidden], W3[hidden], b3.
 * Inputs per energy-balance unit (leaf facet, or ground side): absorbed SW,
 * sky-view fraction, ground-view fraction, LW emissivity, material index,
 * air temperature, wind speed, previous temperature, sky LW; normalised as
 * (x - featureMean) / featureStd.  Output y: T = Tair + targetMean +
 * targetStd * y. */
uint32_t fh_mlp_param_count(uint32_t hidden);
int fh_mlp_set_params(fh_context *ctx, uint32_t hidden, const float *params, const float *featureMean,
                      const float *featureStd, float targetMean, float targetStd);
int fh_mlp_get_params(fh_context *ctx, float *params);
/* Raw network output y for n normalised input rows (host arrays). */
int fh_mlp_forward(fh_context *ctx, const float *x, uint32_t n, float *y);
/* Mean-squared error over n samples and its gradient (host arrays). */
int fh_mlp_loss_grad(fh_context *ctx, const float *x, const float *target, uint32_t n, float *grad,
                     float *loss);
/* Adam training on n normalised samples x [n][9], target [n] (host arrays,
 * uploaded once): `epochs` shuffled passes of `batch`-sample steps, learning
 * rate learningRate * lrDecay^epoch. */
int fh_mlp_train(fh_context *ctx, const float *x, const float *target, uint32_t n, uint32_t batch,
                 uint32_t epochs, float learningRate, float lrDecay, uint32_t seed, float *lastLoss);
/* Energy-balance units of the scene and, after a node run with
 * recordFeatures or the surrogate, their raw features [units][9] and current
 * temperatures [units]. */
uint32_t fh_eb_unit_count(const fh_context *ctx);
int fh_eb_get_features(fh_context *ctx, float *features, float *temperature);

#ifdef __cplusplus
}
#endif
#endif
