// raster.hip -- the reference's raster visibility graph and sunlit fraction
// on the device: per direction an orthographic A-buffer of the scene (every
// triangle of every periodic copy appended to the lists of the pixel centres
// it covers, at most `layers` fragments per pixel), then per pixel the
// fragments in depth order give the denominators, sky counts and the
// exchanges between consecutive surfaces (sh/visibility.{vert,frag},
// sh/visibility_resolve.comp, sh/direct_resolve.comp; the CPU twins
// CpuRasterizer core:1212-1335, buildCpuGraph core:1337-1435 and
// computeCpuSunlit core:1438-1474).
//
// Floating point follows the CPU twin expression by expression (this file
// is compiled without FP contraction, divisions correctly rounded), and the
// fragments of a pixel are ordered by (depth, submission order): tile-major,
// facet-minor, which is the order the CPU twin appends them in, so ties in
// depth resolve as there.  Pixels whose list overflows keep a subset that
// depends on the atomic order (the reference's GPU path has the same
// property); the overflow count is reported.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "solver.hpp"

namespace fh {
namespace {

// periodic copies: the neighbour list of the periodic traversal (the same
// order as the ray tracer's tiles: +-X, +-Z, corners, then the ring at 2)
__constant__ int kRasterTiles[21][2] = {{0, 0},   {-1, 0},  {1, 0},  {0, -1}, {0, 1},   {-1, -1}, {-1, 1},
                                        {1, -1},  {1, 1},   {-2, 0}, {2, 0},  {0, -2},  {0, 2},   {-2, -1},
                                        {-2, 1},  {2, -1},  {2, 1},  {-1, -2}, {1, -2}, {-1, 2},  {1, 2}};

struct RasterView {
    float right[3], up[3], view[3], center[3];
    float radius;
    uint32_t W, H, L;
    uint32_t tiles;  // copies rasterised: 1 + periodic neighbours
    float px, pz;
};

constexpr uint32_t kTileShift = 27;  // order key: tile << 27 | side (sides < 2^27)

__device__ __forceinline__ float dot3(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

__device__ __forceinline__ float edgef(float ax, float ay, float bx, float by, float px, float py) {
    return (px - ax) * (by - ay) - (py - ay) * (bx - ax);
}

// Per direction the A-buffer is built in two passes over the same pixel
// coverage: the first counts each pixel's fragments, an exclusive scan of the
// counts (each clamped to `layers`) gives every pixel a contiguous fragment
// list, the second writes the fragments.  The write pass counts each pixel
// down from its total, so the slot a fragment takes depends on the atomic
// order; slots at or beyond `layers` are the overflow (the reference's
// A-buffer keeps the first `layers` arrivals, an order-dependent subset too).
enum class Pass { Count, Graph, Sun };

// adds 1 to arr[idx] for every active lane, one atomic per distinct index of
// the wave (neighbouring pixels mostly share their nearest facet)
__device__ __forceinline__ void wave_count(uint32_t* __restrict__ arr, uint32_t idx, bool active) {
    unsigned long long pending = __ballot(active);
    const uint32_t lane = threadIdx.x & 63u;
    while (pending) {
        const uint32_t leader = uint32_t(__builtin_ctzll(pending));
        const uint32_t lidx = __shfl(idx, int(leader), 64);
        const unsigned long long same = __ballot(active && idx == lidx) & pending;
        if (lane == leader) atomicAdd(&arr[lidx], uint32_t(__popcll(same)));
        pending &= ~same;
    }
}

__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v) {
    const uint32_t lane = threadIdx.x & 63u;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(v, d, 64);
        if (lane >= uint32_t(d)) v += t;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// one wave per (tile, facet) in submission order, its lanes striding over
// the pixels of the triangle's clamped bounding box (a building wall covers
// 10^5 pixels, a leaf a few: per-thread triangles left one lane looping over
// the wall); fragments are (depth bits, tile << 27 | camera side).
// Pass::Count adds to the pixel counts; the write passes place the fragment
// at off[pixel] + slot (below fragCap: a short capacity is measured by the
// caller and the build runs again) and count the facet's kept fragments:
// Pass::Graph for the denominators of both sides (every kept fragment adds to
// both, core:1368-1372), Pass::Sun for the camera side's pixel total
// (core:1456-1460).
template <Pass P>
__global__ __launch_bounds__(256) void k_raster(const float* __restrict__ fverts, const float4* __restrict__ fnorm,
                                                uint32_t F, RasterView rv, uint32_t* __restrict__ counts,
                                                const uint32_t* __restrict__ off, uint2* __restrict__ frags,
                                                size_t fragCap, uint32_t* __restrict__ kept,
                                                uint32_t* __restrict__ overflow) {
    const uint64_t i = uint64_t(blockIdx.x) * (blockDim.x / 64u) + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    if (i >= uint64_t(F) * rv.tiles) return;  // whole waves
    const uint32_t t = uint32_t(i / F), f = uint32_t(i - uint64_t(t) * F);
    const float tx = float(kRasterTiles[t][0]) * rv.px, tz = float(kRasterTiles[t][1]) * rv.pz;
    float x[3], y[3], d[3];
    for (int c = 0; c < 3; ++c) {
        const float* v = fverts + size_t(f) * 9u + 3u * c;
        const float pos[3] = {v[0] + tx, v[1], v[2] + tz};
        const float rel[3] = {pos[0] - rv.center[0], pos[1] - rv.center[1], pos[2] - rv.center[2]};
        x[c] = (0.5f + 0.5f * dot3(rel, rv.right) / rv.radius) * float(rv.W);
        y[c] = (0.5f + 0.5f * dot3(rel, rv.up) / rv.radius) * float(rv.H);
        d[c] = 0.5f - dot3(rel, rv.view) / (2.0f * rv.radius);
    }
    const float4 nn = fnorm[f];
    const float n[3] = {nn.x, nn.y, nn.z};
    const uint32_t side = f * 2u + (dot3(n, rv.view) >= 0.0f ? 0u : 1u);
    const float area = edgef(x[0], y[0], x[1], y[1], x[2], y[2]);
    if (fabsf(area) < 1.0e-8f) return;
    const float minx = fminf(fminf(x[0], x[1]), x[2]), maxx = fmaxf(fmaxf(x[0], x[1]), x[2]);
    const float miny = fminf(fminf(y[0], y[1]), y[2]), maxy = fmaxf(fmaxf(y[0], y[1]), y[2]);
    int x0 = int(floorf(minx));
    if (x0 < 0) x0 = 0;
    int x1 = int(ceilf(maxx));
    if (x1 > int(rv.W) - 1) x1 = int(rv.W) - 1;
    int y0 = int(floorf(miny));
    if (y0 < 0) y0 = 0;
    int y1 = int(ceilf(maxy));
    if (y1 > int(rv.H) - 1) y1 = int(rv.H) - 1;
    if (x1 < x0 || y1 < y0) return;
    const uint32_t bw = uint32_t(x1 - x0 + 1);
    const uint32_t npx = bw * uint32_t(y1 - y0 + 1);  // <= W H <= 2^28
    const uint32_t key = (t << kTileShift) | side;
    const float invArea = 1.0f / area;
    uint32_t keptN = 0, over = 0;
    for (uint32_t q = lane; q < npx; q += 64u) {
        const uint32_t row = q / bw;
        const int py = y0 + int(row), pxl = x0 + int(q - row * bw);
        const float cx = float(pxl) + 0.5f, cy = float(py) + 0.5f;
        const float w0 = edgef(x[1], y[1], x[2], y[2], cx, cy);
        const float w1 = edgef(x[2], y[2], x[0], y[0], cx, cy);
        const float w2 = edgef(x[0], y[0], x[1], y[1], cx, cy);
        const bool inside = area > 0.0f ? (w0 >= 0.0f && w1 >= 0.0f && w2 >= 0.0f)
                                        : (w0 <= 0.0f && w1 <= 0.0f && w2 <= 0.0f);
        if (!inside) continue;
        const size_t pixel = size_t(py) * rv.W + size_t(pxl);
        if constexpr (P == Pass::Count) {
            atomicAdd(&counts[pixel], 1u);
        } else {
            const uint32_t slot = atomicSub(&counts[pixel], 1u) - 1u;
            if (slot >= rv.L) {
                ++over;
                continue;
            }
            const size_t at = size_t(off[pixel]) + slot;
            if (at >= fragCap) continue;
            const float depth = (w0 * d[0] + w1 * d[1] + w2 * d[2]) * invArea;
            frags[at] = make_uint2(__float_as_uint(depth), key);
            ++keptN;
        }
    }
    if constexpr (P != Pass::Count) {
        keptN = wave_sum(keptN);
        over = wave_sum(over);
        if (lane == 0u) {
            if (keptN) {
                if constexpr (P == Pass::Graph) {
                    atomicAdd(&kept[f * 2u], keptN);
                    atomicAdd(&kept[f * 2u + 1u], keptN);
                } else {
                    atomicAdd(&kept[side], keptN);
                }
            }
            if (over) atomicAdd(overflow, over);
        }
    }
}

__device__ __forceinline__ bool frag_less(uint2 a, uint2 b) {
    const float da = __uint_as_float(a.x), db = __uint_as_float(b.x);
    return da < db || (da == db && a.y < b.y);
}

__device__ __forceinline__ uint32_t frag_side(uint2 v) { return v.y & ((1u << kTileShift) - 1u); }

// the largest fragment total of a direction (the scan's last offset)
__global__ void k_note_max(const uint32_t* __restrict__ v, uint32_t* __restrict__ mx) { atomicMax(mx, *v); }

// graph resolve (buildCpuGraph's pixel loop, core:1361-1396), one thread per
// pixel: its fragments sorted by (depth, submission order) in place
// (insertion sort; lists are short), the first camera side and the last back
// side see the sky, and consecutive surfaces of different facets exchange
// (back side of the nearer, camera side of the farther), emitted as both
// directed pairs a << sbits | b.  A wave reserves its pairs with one atomic
// on the cursor (a prefix sum over the lanes); pairs beyond cap are counted,
// not written, and the caller runs the build again with the measured total.
// Pixels whose list lies beyond fragCap are skipped (that build is repeated).
// The write pass counts every pixel back to zero when both passes covered the
// same pixels; a pixel left nonzero is flagged in *mismatch and skipped, so
// no unwritten slot is read.
template <class K>
__global__ __launch_bounds__(256) void k_graph_resolve(const uint32_t* __restrict__ counts,
                                                       const uint32_t* __restrict__ off, uint2* __restrict__ frags,
                                                       size_t pixels, size_t fragCap, uint32_t sbits,
                                                       uint32_t* __restrict__ sky, K* __restrict__ pairs,
                                                       unsigned long long* __restrict__ cursor,
                                                       unsigned long long cap, uint32_t* __restrict__ mismatch) {
    const size_t pixel = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    uint32_t cnt = 0;
    size_t b = 0;
    if (pixel < pixels) {
        b = off[pixel];
        const size_t e = off[pixel + 1u];
        if (counts[pixel] != 0u)
            atomicOr(mismatch, 1u);
        else if (e <= fragCap)
            cnt = uint32_t(e - b);
    }
    uint2* fr = frags + b;
    for (uint32_t i = 1; i < cnt; ++i) {
        const uint2 v = fr[i];
        uint32_t j = i;
        while (j > 0u && frag_less(v, fr[j - 1u])) {
            fr[j] = fr[j - 1u];
            --j;
        }
        fr[j] = v;
    }
    uint32_t first = 0, last = 0, np = 0;
    if (cnt) {
        first = frag_side(fr[0]);
        last = frag_side(fr[cnt - 1u]) ^ 1u;
        for (uint32_t i = 0; i + 1u < cnt; ++i)
            if ((frag_side(fr[i]) >> 1) != (frag_side(fr[i + 1u]) >> 1)) np += 2u;
    }
    wave_count(sky, first, cnt != 0u);
    wave_count(sky, last, cnt != 0u);
    const uint32_t incl = wave_inclusive_sum(np);
    const uint32_t waveTotal = __shfl(incl, 63, 64);
    unsigned long long base = 0;
    if ((threadIdx.x & 63u) == 63u && waveTotal) base = atomicAdd(cursor, static_cast<unsigned long long>(waveTotal));
    base = __shfl(base, 63, 64);
    unsigned long long slot = base + (incl - np);
    for (uint32_t i = 0; np && i + 1u < cnt; ++i) {
        const uint32_t a = frag_side(fr[i]) ^ 1u, c = frag_side(fr[i + 1u]);
        if ((a >> 1) == (c >> 1)) continue;
        if (slot + 2ull <= cap) {
            pairs[slot] = (K(a) << sbits) | K(c);
            pairs[slot + 1ull] = (K(c) << sbits) | K(a);
        }
        slot += 2ull;
    }
}

// sunlit resolve (computeCpuSunlit, core:1438-1474): the nearest fragment of
// each pixel (first in submission order among equal depths) is visible; the
// per-side totals come from the write pass
__global__ __launch_bounds__(256) void k_sun_resolve(const uint32_t* __restrict__ counts,
                                                     const uint32_t* __restrict__ off, const uint2* __restrict__ frags,
                                                     size_t pixels, uint32_t* __restrict__ visible,
                                                     uint32_t* __restrict__ mismatch) {
    const size_t pixel = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    uint32_t cnt = 0;
    size_t b = 0;
    if (pixel < pixels) {
        b = off[pixel];
        if (counts[pixel] != 0u)
            atomicOr(mismatch, 1u);
        else
            cnt = uint32_t(off[pixel + 1u] - b);
    }
    uint2 best = make_uint2(0u, 0u);
    if (cnt) best = frags[b];
    for (uint32_t i = 1; i < cnt; ++i) {
        const uint2 v = frags[b + i];
        if (frag_less(v, best)) best = v;
    }
    wave_count(visible, frag_side(best), cnt != 0u);
}

__global__ void k_sun_fraction_raster(const uint32_t* __restrict__ total, const uint32_t* __restrict__ visible,
                                      uint32_t S, float* __restrict__ frac) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S) frac[s] = total[s] ? float(visible[s]) / float(total[s]) : 0.0f;
}

// ---- pair sort: LSD radix sort, 8-bit digits, tiles of 2048 keys ----------
// Per pass: each tile's digit histogram (digit-major over the tiles), an
// exclusive scan of it gives every (digit, tile) its output start, and the
// scatter places each key after the equal digits of earlier tiles, earlier
// rounds of its tile, earlier waves of its round and earlier lanes of its
// wave (ranks from eight ballots over the digit bits): stable, so the passes
// from the low digit up sort the keys.
constexpr uint32_t kSortRounds = 8u, kSortTile = 256u * kSortRounds;

template <class K>
__global__ __launch_bounds__(256) void k_digit_hist(const K* __restrict__ keys, uint32_t n, uint32_t shift,
                                                    uint32_t nTiles, uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const size_t base = size_t(blockIdx.x) * kSortTile;
    for (uint32_t r = 0; r < kSortRounds; ++r) {
        const size_t i = base + r * 256u + threadIdx.x;
        if (i < n) atomicAdd(&h[uint32_t(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[size_t(threadIdx.x) * nTiles + blockIdx.x] = h[threadIdx.x];
}

template <class K>
__global__ __launch_bounds__(256) void k_digit_scatter(const K* __restrict__ in, K* __restrict__ out, uint32_t n,
                                                       uint32_t shift, uint32_t nTiles,
                                                       const uint32_t* __restrict__ start) {
    __shared__ uint32_t running[256];
    __shared__ uint32_t waveCnt[4][256];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    running[tid] = start[size_t(tid) * nTiles + blockIdx.x];
    for (uint32_t v = 0; v < 4u; ++v) waveCnt[v][tid] = 0u;
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1ull;
    const size_t base = size_t(blockIdx.x) * kSortTile;
    for (uint32_t r = 0; r < kSortRounds; ++r) {
        const size_t i = base + r * 256u + tid;
        const bool valid = i < n;
        const K key = valid ? in[i] : K(0);
        const uint32_t dg = uint32_t(key >> shift) & 255u;
        unsigned long long same = __ballot(valid);
        for (uint32_t bit = 0; bit < 8u; ++bit) {
            const bool set = (dg >> bit) & 1u;
            const unsigned long long bb = __ballot(set);
            same &= set ? bb : ~bb;
        }
        const uint32_t rank = uint32_t(__popcll(same & below));
        if (valid && rank == 0u) waveCnt[w][dg] = uint32_t(__popcll(same));
        __syncthreads();
        if (valid) {
            uint32_t at = running[dg];
            for (uint32_t v = 0; v < w; ++v) at += waveCnt[v][dg];
            out[at + rank] = key;
        }
        __syncthreads();
        running[tid] += waveCnt[0][tid] + waveCnt[1][tid] + waveCnt[2][tid] + waveCnt[3][tid];
        for (uint32_t v = 0; v < 4u; ++v) waveCnt[v][tid] = 0u;
        __syncthreads();
    }
}

// run-length encoding of the sorted pairs: head flags, their scan numbers
// the distinct pairs, each head records its key and position
template <class K>
__global__ void k_run_heads(const K* __restrict__ keys, uint32_t n, uint32_t* __restrict__ head) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) head[i] = (i == 0u || keys[i] != keys[i - 1u]) ? 1u : 0u;
}

template <class K>
__global__ void k_run_starts(const K* __restrict__ keys, uint32_t n, const uint32_t* __restrict__ headIdx,
                             K* __restrict__ uniq, uint32_t* __restrict__ pos) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i == 0u || keys[i] != keys[i - 1u]) {
        uniq[headIdx[i]] = keys[i];
        pos[headIdx[i]] = i;
    }
    if (i == n - 1u) pos[headIdx[n]] = n;
}

// CSR entries from the distinct pairs: column and multiplicity (the run
// length)
template <class K>
__global__ void k_csr_entries(const K* __restrict__ keys, const uint32_t* __restrict__ pos, uint32_t n,
                              uint32_t sbits, uint32_t* __restrict__ col, uint32_t* __restrict__ cnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    col[i] = uint32_t(keys[i] & ((K(1) << sbits) - K(1)));
    cnt[i] = pos[i + 1u] - pos[i];
}

// row offsets: the first sorted pair of each row (binary search)
template <class K>
__global__ void k_csr_rows(const K* __restrict__ keys, uint32_t n, uint32_t sbits, uint32_t S,
                           uint32_t* __restrict__ rowptr) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > S) return;
    if (r == S) {
        rowptr[S] = n;
        return;
    }
    const K target = K(r) << sbits;
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2u;
        if (keys[mid] < target)
            lo = mid + 1u;
        else
            hi = mid;
    }
    rowptr[r] = lo;
}

// host twins of the CPU rasterizer's vector helpers (same float operations)
float hdot(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
void hcross(const float* a, const float* b, float* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
void hnormalize(const float* a, float* o) {
    const float inv = 1.0f / std::sqrt(hdot(a, a));
    o[0] = a[0] * inv;
    o[1] = a[1] * inv;
    o[2] = a[2] * inv;
}

const int kHostTiles[21][2] = {{0, 0},   {-1, 0},  {1, 0},  {0, -1}, {0, 1},   {-1, -1}, {-1, 1},
                               {1, -1},  {1, 1},   {-2, 0}, {2, 0},  {0, -2},  {0, 2},   {-2, -1},
                               {-2, 1},  {2, -1},  {2, 1},  {-1, -2}, {1, -2}, {-1, 2},  {1, 2}};

// the bounding sphere of the periodically extended scene and the camera
// basis of one direction (makeProjection, facetrt_vulkan.cpp:472-510; the
// CPU twin's constructor core:1212-1240 and core:1248-1260)
RasterView make_view(const fh_context* c, uint32_t W, uint32_t H, uint32_t layers, const float dir[3]) {
    RasterView rv{};
    const float* bmin = c->boundsMin;
    const float* bmax = c->boundsMax;
    const uint32_t periodic = c->cfg.periodicNeighborCount;
    rv.tiles = 1u + periodic;
    rv.px = c->cfg.periodicSizeX;
    rv.pz = c->cfg.periodicSizeZ;
    float emin[3] = {bmin[0], bmin[1], bmin[2]}, emax[3] = {bmax[0], bmax[1], bmax[2]};
    for (uint32_t t = 1; t <= periodic; ++t) {
        const float ox = float(kHostTiles[t][0]) * rv.px, oz = float(kHostTiles[t][1]) * rv.pz;
        emin[0] = std::fmin(emin[0], bmin[0] + ox);
        emax[0] = std::fmax(emax[0], bmax[0] + ox);
        emin[2] = std::fmin(emin[2], bmin[2] + oz);
        emax[2] = std::fmax(emax[2], bmax[2] + oz);
    }
    const float diag[3] = {emax[0] - emin[0], emax[1] - emin[1], emax[2] - emin[2]};
    for (int k = 0; k < 3; ++k) rv.center[k] = (emin[k] + emax[k]) * 0.5f;
    rv.radius = std::fmax(0.5f * std::sqrt(hdot(diag, diag)) * 1.001f, 1.0e-4f);
    rv.W = W;
    rv.H = H;
    rv.L = layers;
    hnormalize(dir, rv.view);
    const float refz[3] = {0.0f, 0.0f, 1.0f}, refy[3] = {0.0f, 1.0f, 0.0f};
    const float* ref = std::fabs(rv.view[2]) < 0.9f ? refz : refy;
    float rc[3];
    hcross(ref, rv.view, rc);
    hnormalize(rc, rv.right);
    hcross(rv.view, rv.right, rv.up);
    return rv;
}

// the reference's limits: a 64-bit A-buffer of 8-byte fragments
// (createVisibilityResources, facetrt_vulkan.cpp:526-531); here also the
// device-side bounds: fragments of one pass addressable in 64 bits, sides
// below 2^27 for the order key
void check_raster(const fh_context* c, uint32_t W, uint32_t H, uint32_t layers, const float* dir) {
    if (W == 0u || H == 0u || layers == 0u) fail(FH_INVALID_ARGUMENT, "raster size and layers must be positive");
    if (W > 16384u || H > 16384u || layers > 1024u) fail(FH_INVALID_ARGUMENT, "raster size or layers too large");
    if (uint64_t(W) * H * layers > (uint64_t(1) << 31))
        fail(FH_OVERFLOW_ERROR, "Visibility A-buffer size overflow");
    if (c->S >= (1u << kTileShift)) fail(FH_OVERFLOW_ERROR, "too many facets for the raster graph");
    if (c->cfg.periodicNeighborCount > 20u) fail(FH_INVALID_ARGUMENT, "periodicNeighborCount must be 0 .. 20");
    if (dir) {
        const float len2 = dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2];
        if (!std::isfinite(len2) || !(len2 > 0.0f)) fail(FH_INVALID_ARGUMENT, "Direction must be finite and non-zero");
    }
}

// bits of a side index: pair keys are a << bits | b
uint32_t side_bits(uint32_t S) {
    uint32_t b = 1;
    while ((uint64_t(1) << b) < S) ++b;
    return b;
}

template <Pass P>
void launch_raster(fh_context* c, const RasterView& rv, uint32_t* counts, const uint32_t* off, uint2* frags,
                   size_t fragCap, uint32_t* kept, uint32_t* overflow) {
    const uint64_t n = uint64_t(c->F) * rv.tiles;  // one wave each, 4 per block
    hipLaunchKernelGGL(k_raster<P>, dim3(uint32_t((n + 3u) / 4u)), dim3(256), 0, c->stream, c->facetVerts.p,
                       c->facetNormal.p, c->F, rv, counts, off, frags, fragCap, kept, overflow);
    FH_HIP(hipGetLastError());
}

// The directions run back to back without a host round trip: the fragment
// lists and the directed pairs of all directions go to buffers of a guessed
// capacity (the last build's measurements, else 4 fragments per pixel and
// one pair per 2 pixels and direction), the kernels count what did not fit,
// and the build runs again at the measured sizes (without overflow the
// counts are deterministic, so the second run fits unless the first was
// short of fragment space and the pairs were undercounted: a third run).
// The pairs are then radix sorted once (2 x side bits) and run-length
// encoded into the CSR: the multiplicity of a pair is its ray count.
template <class K>
uint32_t raster_graph(fh_context* c, const float* dirs, uint32_t ndirs, uint32_t W, uint32_t H, uint32_t layers,
                      uint32_t sbits) {
    const uint32_t S = c->S;
    const size_t pixels = size_t(W) * H;
    std::vector<RasterView> views(ndirs);
    for (uint32_t d = 0; d < ndirs; ++d) views[d] = make_view(c, W, H, layers, dirs + 3u * d);
    DBuf<uint32_t> counts, off, st;  // st: largest fragment total, overflow, pass mismatch
    DBuf<uint2> frags;
    DBuf<K> pairs, alt;
    DBuf<unsigned long long> cursor;
    counts.ensure(pixels);
    off.ensure(pixels + 1u);
    st.ensure(3);
    cursor.ensure(1);
    c->den.ensure(S);
    c->sky.ensure(S);
    size_t fragCap = c->rasterFragHint ? c->rasterFragHint : std::min(4u * pixels, pixels * layers);
    size_t pairCap = c->rasterPairHint ? c->rasterPairHint : std::max<size_t>(pixels * ndirs / 2u, 65536u);
    unsigned long long total = 0;
    uint32_t hs[3] = {0u, 0u, 0u};
    for (int attempt = 0;; ++attempt) {
        frags.ensure(std::max<size_t>(fragCap, 1u));
        pairs.ensure(std::max<size_t>(pairCap, 1u));
        launch_fill_u32(c, c->den.p, S, 0u);
        launch_fill_u32(c, c->sky.p, S, 0u);
        launch_fill_u32(c, st.p, 3, 0u);
        launch_fill_u32(c, reinterpret_cast<uint32_t*>(cursor.p), 2, 0u);
        for (uint32_t d = 0; d < ndirs; ++d) {
            const RasterView& rv = views[d];
            launch_fill_u32(c, counts.p, pixels, 0u);
            launch_raster<Pass::Count>(c, rv, counts.p, nullptr, nullptr, 0, nullptr, nullptr);
            launch_scan(c, counts.p, uint32_t(pixels), off.p, layers);
            hipLaunchKernelGGL(k_note_max, dim3(1), dim3(1), 0, c->stream, off.p + pixels, st.p);
            launch_raster<Pass::Graph>(c, rv, counts.p, off.p, frags.p, fragCap, c->den.p, st.p + 1);
            hipLaunchKernelGGL(k_graph_resolve<K>, dim3(uint32_t((pixels + 255u) / 256u)), dim3(256), 0, c->stream,
                               counts.p, off.p, frags.p, pixels, fragCap, sbits, c->sky.p, pairs.p, cursor.p,
                               static_cast<unsigned long long>(pairCap), st.p + 2);
            FH_HIP(hipGetLastError());
        }
        FH_HIP(hipMemcpyAsync(hs, st.p, sizeof(hs), hipMemcpyDeviceToHost, c->stream));
        FH_HIP(hipMemcpyAsync(&total, cursor.p, sizeof(total), hipMemcpyDeviceToHost, c->stream));
        FH_HIP(hipStreamSynchronize(c->stream));
        if (hs[2]) fail(FH_RUNTIME_ERROR, "raster graph: the count and write passes covered different pixels");
        if (hs[1]) return hs[1];  // overflow: no graph
        const bool fragOk = hs[0] <= fragCap, pairOk = total <= pairCap;
        if (fragOk && pairOk) break;
        if (attempt == 2) fail(FH_RUNTIME_ERROR, "raster graph: buffer sizes did not settle");
        if (!fragOk) fragCap = hs[0];
        if (!pairOk) pairCap = size_t(total);
    }
    c->rasterFragHint = hs[0];
    c->rasterPairHint = size_t(total);
    counts.release();
    frags.release();
    uint32_t nnz = 0;
    DBuf<K> uniq;
    DBuf<uint32_t> hist, start, head, headIdx, pos;
    const K* sortedKeys = nullptr;
    if (total) {
        if (total > 0x7fffffffull) fail(FH_OVERFLOW_ERROR, "raster graph: too many exchanges");
        const uint32_t n = uint32_t(total);
        const uint32_t nTiles = (n + kSortTile - 1u) / kSortTile;
        alt.ensure(total);
        hist.ensure(size_t(nTiles) * 256u);
        start.ensure(size_t(nTiles) * 256u + 1u);
        K* src = pairs.p;
        K* dst = alt.p;
        for (uint32_t shift = 0; shift < 2u * sbits; shift += 8u) {
            hipLaunchKernelGGL(k_digit_hist<K>, dim3(nTiles), dim3(256), 0, c->stream, src, n, shift, nTiles, hist.p);
            launch_scan(c, hist.p, nTiles * 256u, start.p);
            hipLaunchKernelGGL(k_digit_scatter<K>, dim3(nTiles), dim3(256), 0, c->stream, src, dst, n, shift, nTiles,
                               start.p);
            std::swap(src, dst);
        }
        head.ensure(n);
        headIdx.ensure(size_t(n) + 1u);
        const uint32_t g = (n + 255u) / 256u;
        hipLaunchKernelGGL(k_run_heads<K>, dim3(g), dim3(256), 0, c->stream, src, n, head.p);
        launch_scan(c, head.p, n, headIdx.p);
        FH_HIP(hipMemcpyAsync(&nnz, headIdx.p + n, sizeof(nnz), hipMemcpyDeviceToHost, c->stream));
        FH_HIP(hipStreamSynchronize(c->stream));
        uniq.ensure(nnz);
        pos.ensure(size_t(nnz) + 1u);
        hipLaunchKernelGGL(k_run_starts<K>, dim3(g), dim3(256), 0, c->stream, src, n, headIdx.p, uniq.p, pos.p);
        FH_HIP(hipGetLastError());
        sortedKeys = uniq.p;
    }
    // CSR: rows and the columns within a row come out sorted
    c->nnz = nnz;
    alloc_csr_device(c, nnz);
    if (nnz)
        hipLaunchKernelGGL(k_csr_entries<K>, dim3((nnz + 255u) / 256u), dim3(256), 0, c->stream, sortedKeys, pos.p,
                           nnz, sbits, c->col.p, c->cnt.p);
    c->rowptr.ensure(size_t(S) + 1u);
    hipLaunchKernelGGL(k_csr_rows<K>, dim3((S + 1u + 255u) / 256u), dim3(256), 0, c->stream, sortedKeys, nnz, sbits,
                       S, c->rowptr.p);
    FH_HIP(hipGetLastError());
    FH_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

}  // namespace

uint32_t launch_raster_graph(fh_context* c, const float* dirs, uint32_t ndirs, uint32_t W, uint32_t H,
                             uint32_t layers) {
    check_raster(c, W, H, layers, nullptr);
    for (uint32_t d = 0; d < ndirs; ++d) check_raster(c, W, H, layers, dirs + 3u * d);
    c->nnz = 0;
    const uint32_t sbits = side_bits(c->S);
    if (2u * sbits <= 32u) return raster_graph<uint32_t>(c, dirs, ndirs, W, H, layers, sbits);
    return raster_graph<unsigned long long>(c, dirs, ndirs, W, H, layers, sbits);
}

// one direction: the fragment total is read back before the write pass, so
// the lists get exactly their size
uint32_t launch_raster_sunlit(fh_context* c, const float sun[3], uint32_t W, uint32_t H, uint32_t layers,
                              float* frac) {
    check_raster(c, W, H, layers, sun);
    const uint32_t S = c->S;
    const RasterView rv = make_view(c, W, H, layers, sun);
    const size_t pixels = size_t(rv.W) * rv.H;
    DBuf<uint32_t> counts, off, ovf, total, visible;
    DBuf<uint2> frags;
    counts.ensure(pixels);
    off.ensure(pixels + 1u);
    ovf.ensure(2);  // overflow, pass mismatch
    total.ensure(S);
    visible.ensure(S);
    launch_fill_u32(c, counts.p, pixels, 0u);
    launch_raster<Pass::Count>(c, rv, counts.p, nullptr, nullptr, 0, nullptr, nullptr);
    launch_scan(c, counts.p, uint32_t(pixels), off.p, layers);
    uint32_t nfrag = 0;
    FH_HIP(hipMemcpyAsync(&nfrag, off.p + pixels, sizeof(nfrag), hipMemcpyDeviceToHost, c->stream));
    FH_HIP(hipStreamSynchronize(c->stream));
    frags.ensure(std::max<uint32_t>(nfrag, 1u));
    launch_fill_u32(c, total.p, S, 0u);
    launch_fill_u32(c, visible.p, S, 0u);
    launch_fill_u32(c, ovf.p, 2, 0u);
    launch_raster<Pass::Sun>(c, rv, counts.p, off.p, frags.p, nfrag, total.p, ovf.p);
    hipLaunchKernelGGL(k_sun_resolve, dim3(uint32_t((pixels + 255u) / 256u)), dim3(256), 0, c->stream, counts.p,
                       off.p, frags.p, pixels, visible.p, ovf.p + 1);
    hipLaunchKernelGGL(k_sun_fraction_raster, dim3((S + 255u) / 256u), dim3(256), 0, c->stream, total.p, visible.p,
                       S, frac);
    FH_HIP(hipGetLastError());
    uint32_t res[2] = {0u, 0u};
    FH_HIP(hipMemcpyAsync(res, ovf.p, sizeof(res), hipMemcpyDeviceToHost, c->stream));
    FH_HIP(hipStreamSynchronize(c->stream));
    if (res[1]) fail(FH_RUNTIME_ERROR, "raster sunlit: the count and write passes covered different pixels");
    return res[0];
}

}  // namespace fh
