/*
 * facet_oracle.h -- CPU oracle for the facet radiosity / energy-balance path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product (streamsim_amd) never links, imports or calls it.
 *
 * It restates, in plain C with FP64 accumulation, the reference's retained
 * CPU backend (models/histream/src/facetrt/facetrtcore.cpp, "core" below):
 *   - golden-spiral hemisphere directions          core:475-489
 *   - sun direction                                core:491-498
 *   - orthographic A-buffer rasterizer             core:1212-1335
 *   - visibility graph (depth sort, den/sky/edges) core:1348-1436
 *   - raster sunlit fraction                       core:1438-1474
 *   - two-sided Jacobi (FP64 gather)               core:1476-1529
 *   - incident irradiance                          core:1531-1559
 * plus the host replay of the seeded cosine-weighted ray graph that the HIP
 * path computes (same counter RNG, same watertight triangle test, closest
 * hit ordered by (t, facet)), a dense FP64 direct solve of the radiosity
 * system, the FacetEB flux / Newton loop in FP64 (core:854-915,
 * core:3567-3832) and the MLP forward / backward in FP64.
 *
 * Every function that has no pin (golden vector, closed form, invariant or
 * brute force) in tests/ says "parity unpinned" next to its declaration.
 */
#ifndef FACET_ORACLE_H
#define FACET_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    float position[3];
    float normal[3];
} or_vertex;

typedef struct {
    float reflectance;
    float transmittance;
    float emission;
    float directIrradiance;
} or_optics;

/* Integer-count visibility graph over sides (side 2f = normal side).
 * Rows are sorted by neighbour index.  F_sn = cnt / den[s]. */
typedef struct {
    uint32_t S;
    uint32_t nnz;
    uint32_t *rowptr;  /* S+1 */
    uint32_t *col;     /* nnz */
    uint32_t *cnt;     /* nnz */
    uint32_t *den;     /* S   */
    uint32_t *sky;     /* S   */
    uint32_t overflow;
    float closure;
} or_graph;

void or_graph_free(or_graph *g);
/* Copy a graph out (arrays sized from S and nnz). */
void or_graph_copy(const or_graph *g, uint32_t *rowptr, uint32_t *col,
                   uint32_t *cnt, uint32_t *den, uint32_t *sky);
/* Build a graph from caller arrays (deep copy). */
int or_graph_from_arrays(or_graph *g, uint32_t S, uint32_t nnz,
                         const uint32_t *rowptr, const uint32_t *col,
                         const uint32_t *cnt, const uint32_t *den,
                         const uint32_t *sky);

/* core:475-489 */
void or_hemisphere_directions(uint32_t count, float *xyz);
/* core:491-498 */
void or_sun_direction(float zenith_deg, float azimuth_deg, float xyz[3]);

/* core:1348-1436.  Returns 0, or -1 on A-buffer overflow (the reference
 * throws), -2 on allocation failure. */
int or_raster_graph(const or_vertex *v, uint32_t F, const float bmin[3],
                    const float bmax[3], uint32_t raster, uint32_t layers,
                    uint32_t periodic, float periodicX, float periodicZ,
                    const float *dirs, uint32_t ndirs, or_graph *out);

/* core:1438-1474 */
int or_raster_sunlit(const or_vertex *v, uint32_t F, const float bmin[3],
                     const float bmax[3], uint32_t raster, uint32_t layers,
                     uint32_t periodic, float periodicX, float periodicZ,
                     const float sun[3], float *fraction);

/* Host replay of the seeded ray graph.  brute != 0 tests every triangle for
 * every ray (the definition); otherwise an independent median-split BVH with
 * conservative box culling is used.  hits (optional, S*R) receives the hit
 * side of every ray (0xffffffff = sky). */
int or_ray_graph(const or_vertex *v, uint32_t F, uint32_t seed, uint32_t R,
                 float tmin, int brute, or_graph *out, uint32_t *hits);

/* Host replay of the seeded sun-ray sunlit fraction (K rays per side). */
int or_ray_sunlit(const or_vertex *v, uint32_t F, const float sun[3],
                  uint32_t seed, uint32_t K, float tmin, int brute,
                  float *fraction);

/* The same with periodic copies of the scene (sh/visibility.vert:15-35):
 * tiles 1..periodic of the neighbour list, offsets (dx px, 0, dz pz); every
 * ray is traced against each tile in order with its origin moved by minus
 * the offset, keeping the best (t, facet); copies of the origin facet are
 * ordinary surfaces. */
int or_ray_graph_periodic(const or_vertex *v, uint32_t F, uint32_t seed, uint32_t R, float tmin, int brute,
                          uint32_t periodic, float px, float pz, or_graph *out, uint32_t *hits);
int or_ray_sunlit_periodic(const or_vertex *v, uint32_t F, const float sun[3], uint32_t seed, uint32_t K,
                           float tmin, int brute, uint32_t periodic, float px, float pz, float *fraction);

/* Single ray query (debug / tests): closest hit side or 0xffffffff. */
uint32_t or_trace_one(const or_vertex *v, uint32_t F, const float o[3],
                      const float d[3], uint32_t skipFacet, float tmin,
                      float *tHit);
/* The ray generator of the graph build (origin, direction). */
void or_ray_generate(const or_vertex *v, uint32_t side, uint32_t ray,
                     uint32_t seed, float o[3], float d[3]);

/* core:1476-1529: fixed iterations from zero, FP64 gather, FP32 storage. */
void or_solve(const or_graph *g, const or_optics *optics, float sky,
              uint32_t iterations, float relaxation, float *B,
              float *maxDelta);

/* Multi-band [S][nb] Jacobi with the same per-band update, run until the
 * largest relative change |dB|/max(|B|,1e-30) <= tolRel (checked after every
 * sweep) or maxIter sweeps.  init may be NULL (zero start).  Returns the
 * number of sweeps. */
uint32_t or_solve_bands(const or_graph *g, uint32_t nb, const float *r,
                        const float *t, const float *e, const float *D,
                        const float *sky, const float *init,
                        uint32_t maxIter, double tolRel, float relaxation,
                        float *B, double *lastRel);

/* Threads of the multi-band solve (OpenMP over facets; the results do not
 * depend on the count).  The single-threaded run is the restatement of the
 * reference's single-threaded CPU route; all cores are its parallel variant. */
void or_set_threads(int n);
int or_max_threads(void);

/* core:1531-1559 */
void or_incident(const or_graph *g, const float *B, const or_optics *optics,
                 float sky, float *H);
void or_incident_bands(const or_graph *g, uint32_t nb, const float *B,
                       const float *D, const float *sky, float *H);

/* Dense FP64 LU solve of the converged two-sided system, one band
 * (S <= 6000).  Returns 0 or -1. */
int or_solve_direct(const or_graph *g, const or_optics *optics, float sky,
                    double *B);

/* ---- FacetEB ---------------------------------------------------------- */

/* Per-material energy-balance parameters (SurfaceParameters, core:113-143). */
typedef struct {
    float tirReflectance;
    float tirTransmittance;
    float surfaceResistance;
    float specificHeat;
    float density;
    float thermalConductivity;
    float soilMoisture;
    float heatCapacityPerArea;
    float convectiveScale;
    int32_t soilTemperatureMethod;
    int32_t vegetation;
    int32_t photovoltaic;    /* object facets generate electric power */
    float pvEta25;
    float pvGamma;
    float pvBifaciality;
} or_eb_material;

typedef struct {
    float airTemperatureK;   /* clamped Tair */
    float vaporPressure;     /* hPa */
    float pressure;          /* hPa */
    float windSpeed;         /* m/s */
    float skyLongwave;       /* W m-2 */
    float dTime;             /* s */
    float tolerance;         /* K */
    float relaxation;        /* temperature relaxation */
    uint32_t maxCoupling;
    uint32_t vegetationMethod;
    uint32_t lwMaxSweeps;    /* inner LW sweeps cap */
    float lwTolRel;          /* inner LW relative tolerance (<0: fixed sweeps) */
    float backgroundAngularStrength; /* Hapke strength of the ground (0: Lambert) */
} or_eb_params;

/* Hapke background (core:546-605).  Shortwave reflectance of a ground side
 * (normal n) lit from `incoming` and seen from `outgoing`, and the ground's
 * longwave emissivity at cosine |n_y|; both blended with the Lambertian
 * value by strength in [0, 1]. */
float or_hapke_reflectance(float reflectance, const float incoming[3], const float outgoing[3],
                           const float normal[3], float strength);
float or_hapke_emissivity(float emissivity, float cosine, float strength);

/* One FacetEB node (core:3567-3846) in FP64: coupling loop of
 * LW optics -> LW solve -> incident -> fluxes -> relaxed Newton, then the
 * final LW solve and residual.  T (S) is in/out; Tprev (S) is the previous
 * node's state.  absorbedSW (S) is given; incidentSW (S, the shortwave
 * incident per side, may be NULL) feeds the power of photovoltaic object
 * facets (core:3655-3663).  Outputs per side may be NULL.  Returns the
 * number of coupling iterations; *converged set. */
/* FacetEB flux helpers in FP64 (core:854-915): saturation vapour pressure
 * (hPa) and its slope at T (K); latent heat flux and its T-derivative for air
 * density rho, pressure P (hPa), vapour pressure ea (hPa), aerodynamic and
 * surface resistances ra, rs (s m-1); soil storage coefficient
 * 2 sqrt(c rho lambda) / sqrt(pi dT) (W m-2 K-1). */
double or_sat_vp(double T);
double or_sat_vp_slope(double T);
double or_latent(double T, double rho, double P, double ea, double ra, double rs);
double or_latent_slope(double T, double rho, double P, double ea, double ra, double rs);
double or_soil_storage(const or_eb_material *m, double dT);

uint32_t or_eb_node(const or_graph *g, uint32_t F, uint32_t leafCount,
                    const uint32_t *material, const or_eb_material *mats,
                    const float *normalY, const float *absorbedSW,
                    const float *incidentSW,
                    const or_eb_params *p, float *T, const float *Tprev,
                    float *Blw, int *converged, float *maxDeltaT,
                    float *maxResidualEq);

/* ---- MLP 9 -> H -> H -> 1 (GELU, tanh form) in FP64 --------------------- */
/* Parameter layout (float): W1[H][IN], b1[H], W2[H][H], b2[H], W3[H], b3. */
void or_mlp_forward(const float *params, uint32_t in, uint32_t hidden,
                    const float *x, uint32_t n, double *y);
/* Mean-squared-error loss over n samples and its gradient (same layout). */
double or_mlp_loss_grad(const float *params, uint32_t in, uint32_t hidden,
                        const float *x, const float *target, uint32_t n,
                        double *grad);

#ifdef __cplusplus
}
#endif
#endif
