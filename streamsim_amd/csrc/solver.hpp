// solver.hpp -- host-side orchestration shared by the C-ABI and the FacetEB
// driver: kernel launchers and the device-converged Jacobi loop.
#pragma once

#include "engine.hpp"

namespace fh {

struct SweepArgs;

// raytrace.hip
void launch_trace_graph(fh_context* c, uint32_t R, uint32_t seed, float tmin);
void launch_rows(fh_context* c, uint32_t R);
// out[0..n] = exclusive scan of min(in[i], clampMax); out[n] = the total
void launch_scan(fh_context* c, const uint32_t* in, uint32_t n, uint32_t* out, uint32_t clampMax = 0xffffffffu);
void launch_compact(fh_context* c, uint32_t R);
// weights cnt / den, fsky and the closure; packed entries when packable
// (every count <= its row's den <= 65535) and column and count (<= maxCount)
// fit in 32 bits
void launch_weights(fh_context* c, uint32_t* closureBits, uint32_t maxCount, bool packable);
void launch_sun(fh_context* c, float sx, float sy, float sz, uint32_t* lit, float* frac);

// raster.hip: the reference's raster visibility graph (CSR, den, sky and the
// weights' inputs in c; returns the fragment overflow count, the graph is
// complete only at 0) and raster sunlit fraction (device frac [S]; returns
// the overflow count)
uint32_t launch_raster_graph(fh_context* c, const float* dirs, uint32_t ndirs, uint32_t W, uint32_t H,
                             uint32_t layers);
uint32_t launch_raster_sunlit(fh_context* c, const float sun[3], uint32_t W, uint32_t H, uint32_t layers,
                              float* frac);

// abi.hip: CSR entry arrays for nnz entries plus the zero pad
void alloc_csr_device(fh_context* c, uint32_t nnz);

// spmm.hip
void launch_fold(fh_context* c, const float* e, const float* r, const float* t, const float* D, uint32_t ldb,
                 float* out);
void launch_pad(fh_context* c, const float* src, uint32_t nb, uint32_t ldb, float* dst);
void launch_unpad(fh_context* c, const float* src, uint32_t nb, uint32_t ldb, float* dst);
void launch_split_optics(fh_context* c, const float4* o, float* r, float* t, float* e, float* D);

// optics.hip: per-side, per-band optics [S][nb] from device inputs
struct SpectralOpticsArgs {
    uint32_t mode, nb, S, leafCount;
    const uint32_t* mat;                          // [F]
    const float *reflectance, *transmittance;     // [materials][nb]
    const float *wavelength, *dni;                // [nb]
    const float* factor;                          // [F] or null
    const float *sunlit, *T;                      // [S] (T: state mode)
    const float *Tsun, *Tshade;                   // [materials] (scene mode)
    const float4* nrm;                            // [F]
    float strength, sx, sy, sz;
    float *r, *t, *e, *D;                         // [S][nb]
};
void launch_spectral_optics(fh_context* c, const SpectralOpticsArgs& a);
void launch_ctrl_reset(fh_context* c, uint32_t maxIter, float tolRel, float tolAbs, uint32_t checkEvery,
                       uint32_t nAct = 1u, float innerTol = -1.0f);
void launch_fill_u32(fh_context* c, uint32_t* p, size_t n, uint32_t v);
void launch_fill_f32(fh_context* c, float* p, size_t n, float v);

// One solve of (I - M) B = c on the current graph: bufs[0] holds the start
// value; c/r/t are padded [S][ldb] device arrays.  Returns the index (0/1)
// of the buffer holding the result and fills stats.
struct JacobiOptions {
    uint32_t ldb = 1;
    uint32_t precision = FH_PRECISION_FP32;
    float relax = 1.0f;
    uint32_t maxIter = 64;
    float tolRel = -1.0f;
    float tolAbs = -1.0f;
    uint32_t checkEvery = 1;
    uint32_t innerBf16 = 12;  // most BF16 sweeps between FP32 refreshes / before the FP32 phase
    // BF16-gather solves (band-parallel layouts): the BF16 phase ends at
    // relative change firstInner (the BF16 operator's fixed point is only
    // about that close), then FP32 sweeps finish the solve
    float firstInner = 2.0e-2f;
    bool retire = true;       // per-band retirement (tolRel >= 0, band-parallel)
    // optional incident H = max(0, fsky S + G + D) written by the judge
    // sweeps (band-parallel with retirement); *incidentDone reports whether
    // every band converged, i.e. H is complete
    float* H = nullptr;
    const float* D = nullptr;
    // squared spectral radius per band (device, ldb): plain FP32 band-parallel
    // solves then run the Chebyshev semi-iteration for [-rho_b, rho_b]
    const float* rho2 = nullptr;
    // energy-balance shortwave solves at band-parallel strides: r, t and c
    // from material tables and per-side direct factors (cc, rr, tt are then
    // read by entry-parallel sweeps and the spectral estimate only)
    const uint32_t* fmat = nullptr;  // material per facet
    const float* rtab = nullptr;     // [materials][ldb]
    const float* ttab = nullptr;     // [materials][ldb]
    const float* dirf = nullptr;     // [S] clamp(sunlit, 0, 1) max(0, cos)
    const float* dniv = nullptr;     // [ldb]
    // increment solves: *incSerial (when given and equal to the context's
    // intact increment state, whose final B the caller passes as bufs[0]
    // under the same operator) lets the solve continue from that state
    // instead of a base sweep; on return it holds this solve's serial (0:
    // no increment state left)
    uint64_t* incSerial = nullptr;
};

// Per-band spectral radius squared of the SW operator (r, t [S][ldb]) by
// power iteration from B = 1: max(B_m+2) / max(B_m) per band into rho2
// (device, ldb).  Uses scratch arrays v0, v1, zero ([S][ldb]).
void estimate_rho2(fh_context* c, uint32_t ldb, uint32_t nb, const float* rr, const float* tt, float* v0, float* v1,
                   float* zero, float* rho2);
int run_jacobi(fh_context* c, const JacobiOptions& o, const float* cc, const float* rr, const float* tt,
               float* bufs[2], fh_solve_stats* stats, bool* incidentDone = nullptr);

// H = max(0, fsky S + G(B) + D) for all bands ([S][ldb])
void run_incident(fh_context* c, uint32_t ldb, const float* B, const float* D, float* H);

void set_sky(fh_context* c, const float* sky, uint32_t nb, uint32_t ldb);

// true when sweeps of this band stride use the band-parallel kernels (BF16
// solves then gather BF16 copies of B with row stride bf16_stride(ldb))
bool uses_bf16_gather(uint32_t ldb);
// whether the energy balance's BF16 mode runs its shortwave solves with BF16
// increments at this band stride (FP32 rows longer than one 128-byte line)
bool bf16_pays(uint32_t ldb);
uint32_t bf16_stride(uint32_t ldb);
// bytes per row of FP8 increments: 16-byte chunks of 12 bands
uint32_t q8_stride(uint32_t ldb);

// whether FP32-weight gathers read the packed CSR entries (when present)
bool use_packed();

// Pair blocks of the band-parallel sweeps, built with the graph's weights:
// facets matched with their most strongly coupled facet, units of two
// facets in sweep order (c->units, c->unitW)
void build_pairs(fh_context* c);

// whether a solve of this band stride and precision runs increment sweeps
// (its accumulated gathers are then in c->gacc after the solve)
bool increment_solve(uint32_t ldb, uint32_t precision);

// facet processing order of the sweeps (null: natural order)
const uint32_t* sweep_schedule(const fh_context* c);

// lanes per facet team of the SpMM kernels for the current graph
int sweep_team(const fh_context* c);
// The single-band gather of the coupled longwave sweep through the staged
// layout (Iin = fsky L + F B for the facets the layout covers: all, or the
// singles when there are pair blocks); false: no staged layout.
bool launch_gather1_staged(fh_context* c, const float* B, float L, const SolveCtrl* ctrl, float* Iin);

// Energy-balance form of a band-parallel sweep for launch_single_sweep: r, t
// from material tables and c, D from per-side direct factors (as the
// shortwave solves of a node run them), and the Chebyshev weights per band
// (null: plain Jacobi).
struct SweepTables {
    const uint32_t* fmat = nullptr;  // material per facet
    const float* rtab = nullptr;     // [materials][ldb]
    const float* ttab = nullptr;     // [materials][ldb]
    const float* dirf = nullptr;     // [S]
    const float* dniv = nullptr;     // [ldb]
    const float* omega = nullptr;    // [ldb]
};

// One sweep, no host synchronisation (timing entry).  With w16 and BF16
// copies given (band-parallel strides), the BF16-gather sweep h16in -> h16out;
// with gacc and carry ([S][ldb]) also given, the increment sweep of BF16
// solves (h16in / h16out hold BF16 increments); with q8in / q8out (rows of
// q8_stride bytes), gacc and carry instead, the FP8 increment sweep.  pairs: the sweep of the
// solves (pair blocks where the graph has them); else one Jacobi sweep.
// partials: device buffer of partialCap per-block residual partials the
// sweep writes instead of the context's own (grown on demand) buffer; a
// sweep that needs more fails with FH_INVALID_ARGUMENT before any launch.
void launch_single_sweep(fh_context* c, uint32_t ldb, const float* cc, const float* rr, const float* tt,
                         const float* Bin, float* Bout, bool w16, uint16_t* h16in = nullptr,
                         uint16_t* h16out = nullptr, float* gacc = nullptr, float* carry = nullptr,
                         bool pairs = false, float2* partials = nullptr, uint32_t partialCap = 0u,
                         const SweepTables* tables = nullptr, uint8_t* q8in = nullptr, uint8_t* q8out = nullptr);

}  // namespace fh
