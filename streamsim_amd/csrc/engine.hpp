// engine.hpp -- internal types of the MI355X facet engine (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/facet_hip.h"

namespace fh {

// Error carried across the C-ABI as an fh_status code + message.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error(code, msg); }

#define FH_HIP(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            ::fh::fail(FH_RUNTIME_ERROR, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Device buffer that only grows.
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    size_t cap = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void ensure(size_t count) {
        if (count > cap) {
            release();
            const size_t bytes = (count ? count : 1) * sizeof(T);
            FH_HIP(hipMalloc(reinterpret_cast<void**>(&p), bytes));
            cap = count ? count : 1;
        }
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = cap = 0;
    }
    void swap(DBuf& o) noexcept {
        std::swap(p, o.p);
        std::swap(n, o.n);
        std::swap(cap, o.cap);
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Child-pair BVH node (64 B): the boxes of both children, so one node fetch
// (4 x 16 B) tests two boxes.
//   x = (lo0.x, hi0.x, lo1.x, hi1.x), y = (...y...), z = (...z...)
//   meta = (child0, child1, count0, count1): count > 0 -> leaf with
//   primitives [child, child + count); count == 0 && child >= 0 -> internal
//   node index; child < 0 -> empty slot (its box never hits).
struct alignas(16) BvhNode {
    float4 x, y, z;
    int4 meta;
};

struct BvhHost {
    std::vector<BvhNode> nodes;
    std::vector<float4> tris;        // 3 per primitive slot; v0.w = facet id bits
    std::vector<uint32_t> primFacet;  // slot -> facet
    int depth = 0;
};

// Binned-SAH build over facet triangles tri9 (9 floats per facet).
void build_bvh(const float* tri9, uint32_t F, BvhHost& out);

constexpr int kBvhMaxDepth = 48;  // stack depth bound of the traversal kernels

// Device-side solve control block (one per context).
struct SolveCtrl {
    uint32_t done;
    uint32_t iter;
    uint32_t maxIter;
    uint32_t checkEvery;  // absolute tolerance checked when iter % checkEvery == 0
    float tolRel;         // < 0: off
    float tolAbs;         // < 0: off
    float lastAbs;
    float lastRel;
    uint32_t refreshes;
    uint32_t innerDone;   // BF16 inner sweeps reached innerTol: skip until the next refresh
    uint32_t nAct;        // band retirement: bands [0, nAct) are still iterated
    uint32_t nActPrev;    // bands [nAct, nActPrev) retired by the last judge: copied once
    float innerTol;       // relative change that ends a group of BF16 sweeps
    uint32_t chebStart;   // sweep index that started the current Chebyshev sequence
    uint32_t pad[2];
};
static_assert(sizeof(SolveCtrl) == 64, "control block is one 64-byte line");

uint32_t band_stride(uint32_t nb);

// CSR column / weight arrays are allocated kCsrPad entries beyond nnz (zero)
// so that the sweep kernels load aligned groups of 8 entries with vector
// loads; the row offsets need no alignment.
constexpr uint32_t kCsrPad = 16;

// FacetEB device state (one per context).
struct EbState {
    bool ready = false;
    uint32_t facets = 0;  // facet count of the scene fh_eb_set_scene was given
    uint32_t leafCount = 0;
    uint32_t nMat = 0;
    uint32_t swNb = 0, swLdb = 0;
    bool groundClamp = false;  // a ground material's SW reflectance exceeds 0.999 (per-side arrays)
    bool textured = false;     // reflFactor scales the SW reflectance per facet (per-side arrays)
    DBuf<float> reflFactor;    // F
    bool swValid = false;  // Bsw holds the previous node's SW solution
    bool lwValid = false;  // Blw holds a LW solution
    bool tValid = false;   // T holds a state
    DBuf<uint32_t> material;       // F
    DBuf<fh_eb_material> mats;     // nMat
    DBuf<float> swRefl, swTrans;   // nMat x swNb
    DBuf<float> swRtab, swTtab;    // nMat x swLdb, clamped, internal band order
    DBuf<float> swDir;             // [S] clamp(sunlit, 0, 1) max(0, cos(normal, sun)) of the current node
    DBuf<double> area;             // F
    // per side
    DBuf<float> T, Tprev, absorbed, incident, sunlit, rlw, tlw, eps, elw, clw, Hlw, zero;
    DBuf<float> Blw0, Blw1;
    DBuf<float> Rn, Hs, LE, G;
    DBuf<float> pvPower;  // electric power of photovoltaic object facets, W m-2 (both sides)
    // SW [S][swLdb]
    DBuf<float> Bsw0, Bsw1, rsw, tsw, Dsw, csw, Hsw, dni;
    // internal SW band j is window swOrder[j]: windows sorted by the largest
    // r + t over the materials (slowest to converge first), so the bands
    // still iterating form a prefix
    std::vector<uint32_t> swOrder;
    DBuf<uint32_t> swOrderD;
    DBuf<float> swRho2;            // [swLdb] spectral radius^2 per internal band
    bool swRhoValid = false;
    // serial of the increment state (accumulated gathers, carries, pending
    // increments) the last shortwave solve left, 0: none; the next node's
    // solve continues from it while it is the context's latest
    uint64_t swIncSerial = 0;
    DBuf<double> dpart;            // per-block FP64 partial sums
    DBuf<double> dsum;             // reduced FP64 sums
    DBuf<float2> fout;             // reduced (max dT, max residual-equivalent)
    DBuf<float4> part4;            // coupled-sweep per-block partials
    DBuf<float> Iin;               // coupled sweep: incident before clamping, per side
    DBuf<float> stepPrev;          // coupled sweep: previous full Newton step per side
    DBuf<float> last4;             // last coupled-sweep maxima
    DBuf<uint32_t> lit;
    // surrogate inputs / outputs per energy-balance unit (leaf facet or
    // ground side): unit u < leafCount is facet u, else ground side
    // 2 leafCount + (u - leafCount)
    DBuf<float> gview;             // per side: view fraction of ground sides
    DBuf<float> feat;              // units x 9 raw features
    DBuf<float> featN;             // units x 9 normalised
    DBuf<float> guess;             // units
    DBuf<float> guessPrev;         // units: the previous node's network output
    float guessPrevTa = 0.0f;      // air temperature of that node
    bool guessPrevValid = false;   // the previous node was warm-started by the network
    bool featValid = false;
    // the last node's shortwave inputs, for its diffuse-only solve: sun,
    // direct normal (internal band order, padded to swLdb) and diffuse per
    // internal band, radiosity tolerance and sweep cap
    float lastSun[3] = {0.0f, 0.0f, 0.0f};
    std::vector<float> lastDni, lastDiffuse;
    float lastTolerance = 1.0e-5f;
    uint32_t lastMaxSweeps = 0;
    bool lastNodeValid = false;    // a node ran on the current graph and scene
    DBuf<float> Bdf0, Bdf1;        // [S][swLdb] diffuse-only solve
    DBuf<double> dfSum;            // [S] its sum over the windows
};

// Surrogate MLP (9 -> H -> H -> 1, GELU) parameters and normalisation.
struct MlpState {
    uint32_t hidden = 0;
    bool ready = false;
    DBuf<float> params;            // W1[H][9], b1[H], W2[H][H], b2[H], W3[H], b3
    float inMean[9] = {}, inStd[9] = {};
    float outMean = 0.0f, outStd = 1.0f;
    uint64_t adamStep = 0;         // Adam steps taken since the parameters were set
    // training workspace
    DBuf<float> A1, H1, A2, H2, y, dA2, dA1, grad, m, v, red;
    DBuf<float> W2T;               // layer-2 weights transposed (matrix-core inference)
};

}  // namespace fh

struct fh_context {
    fh_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    uint32_t F = 0, S = 0;
    bool hasGeometry = false;
    bool hasGraph = false;

    // geometry
    fh::DBuf<float> facetVerts;      // F x 9
    fh::DBuf<float4> facetNormal;    // F
    fh::DBuf<fh::BvhNode> nodes;
    fh::DBuf<float4> tris;
    int bvhDepth = 0;
    uint32_t bvhNodes = 0;

    // graph (CSR over sides)
    uint32_t nnz = 0;
    // packed CSR entries col | cnt << pkBits (4 bytes; the weight cnt / den
    // of the row is recomputed in the kernels), when column and count fit in
    // 32 bits; pkBits = 0: no packed array
    fh::DBuf<uint32_t> pk;
    uint32_t pkBits = 0;
    float closure = 0.0f;
    float hapkeStrength = 0.0f;  // Hapke strength of the ground (energy balance optics)
    float boundsMin[3] = {0.0f, 0.0f, 0.0f};  // vertex bounds (raster projection)
    float boundsMax[3] = {0.0f, 0.0f, 0.0f};
    fh::DBuf<uint32_t> rowptr, col, cnt, den, sky;
    fh::DBuf<uint32_t> sched;        // processing order of facets in the sweeps (Morton order)
    // pair blocks of the band-parallel sweeps: units of two strongly
    // coupled facets in sweep order, per unit the weights of the CSR
    // entries from each facet's sides to the other's (w[f0->g0], w[f0->g1],
    // w[f1->g0], w[f1->g1]; then g -> f), and the other facets in sweep
    // order; unitCount = 0: no pair blocks
    fh::DBuf<uint2> units;
    fh::DBuf<float4> unitW;
    fh::DBuf<uint32_t> singles;
    fh::DBuf<uint32_t> pairFacets;  // the facets of the units, two per unit (single-band gathers)
    uint32_t unitCount = 0;
    uint32_t singleCount = 0;
    // the pair launch of a sweep runs beside the singles launch: second
    // stream (default priority) and the fork / join events
    hipStream_t pairStream = nullptr;
    hipEvent_t pairFork = nullptr, pairJoin = nullptr;
    // staged gathers of the increment sweeps and the coupled longwave
    // gathers (single facets): per block of 128 facets of the singles
    // launch's order its sorted distinct columns (stDcol from stDoff[b]),
    // per chunk and team a run descriptor (stMeta from stMoff[b]) and the
    // padded entries it points to, twice: stKey (16 x slot in the chunk |
    // BF16 count << 16) and stKey16 (slot | count << 10), 0 on padding;
    // stBlocks = 0: not built
    fh::DBuf<uint32_t> stDcol, stDoff, stMoff, stKey;
    fh::DBuf<uint16_t> stKey16;
    fh::DBuf<uint2> stMeta;
    uint32_t stBlocks = 0;
    bool stOnSingles = false;  // built on the singles list (pair blocks on), else on the schedule
    fh::DBuf<float> val32, fsky;
    fh::DBuf<uint16_t> val16;

    // scratch for the ray graph build
    fh::DBuf<uint32_t> hits, rowlen, scanTmp;
    fh::DBuf<uint16_t> cnt16;
    // raster graph build: fragments of the largest direction and directed
    // exchanges of all directions measured by the last build (the first
    // capacities of the next one)
    size_t rasterFragHint = 0, rasterPairHint = 0;

    // solve workspace
    fh::DBuf<float> B0, B1, cc, rr, tt, DD, c2, HH, skyv;
    fh::DBuf<fh::SolveCtrl> ctrl;
    fh::DBuf<float2> partial;        // per-block residual partials
    fh::DBuf<uint32_t> partAct;      // per-block 1 + highest unconverged band
    fh::DBuf<uint16_t> X16a, X16b;   // BF16 copies of the radiosity (or its increments) gathered by BF16 sweeps
    fh::DBuf<uint8_t> X8a, X8b;      // FP8 increments of the radiosity gathered by FP8 increment sweeps
    fh::DBuf<float> gacc, carry;     // increment sweeps: accumulated gathers, B minus the gathered state
    // the increment state in gacc, carry and X16a (the increments written
    // last, not yet gathered) belongs to the increment solve of this serial
    // (its final B, ldb incLdb); 0: no intact state.  solveSerial counts them.
    uint64_t incSerial = 0, solveSerial = 0;
    uint32_t incLdb = 0;
    fh::DBuf<float> omega;           // Chebyshev weights per band
    fh::DBuf<uint32_t> bandMax;      // per-band maxima (float bits) of the rho estimate
    fh::SolveCtrl* hctrl = nullptr;  // pinned mirror
    fh::DBuf<uint32_t> misc;         // small counters
    fh::DBuf<float> fbuf;            // small float scratch

    fh::EbState eb;
    fh::MlpState mlp;

    // device address of this device's fixed scratch array (abi.hip), the
    // per-block partials of fh_sweep_device_noalloc
    float2* fixedPartials = nullptr;
};

namespace fh {

// A new graph invalidates what the energy balance derived from the old one:
// the spectral radius of the shortwave operator, the warm-start radiosities
// and the ground-view fractions of the surrogate features.
inline void eb_graph_changed(fh_context* c) {
    EbState& e = c->eb;
    e.swRhoValid = false;
    e.swIncSerial = 0;
    e.swValid = false;
    e.lwValid = false;
    e.featValid = false;
    e.guessPrevValid = false;
    e.lastNodeValid = false;
    e.gview.release();
}

// New geometry invalidates the whole energy-balance scene (its per-facet
// and per-side buffers are sized for the old facet count):
// fh_eb_set_scene must be called again.
inline void eb_geometry_changed(fh_context* c) {
    eb_graph_changed(c);
    c->eb.ready = false;
    c->eb.tValid = false;
}

}  // namespace fh
