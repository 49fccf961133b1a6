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
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>

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

// one wave per (tile, facet) in submission order, its lanes striding over
// the pixels of the triangle's clamped bounding box (a building wall covers
// 10^5 pixels, a leaf a few: per-thread triangles left one lane looping over
// the wall); fragments are (depth bits, tile << 27 | camera side)
__global__ __launch_bounds__(256) void k_raster(const float* __restrict__ fverts, const float4* __restrict__ fnorm,
                                                uint32_t F, RasterView rv, uint32_t* __restrict__ counts,
                                                uint2* __restrict__ frags, uint32_t* __restrict__ overflow) {
    const uint64_t i = uint64_t(blockIdx.x) * (blockDim.x / 64u) + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    if (i >= uint64_t(F) * rv.tiles) return;
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
        const uint32_t layer = atomicAdd(&counts[pixel], 1u);
        if (layer >= rv.L) {
            atomicAdd(overflow, 1u);
            continue;
        }
        const float depth = (w0 * d[0] + w1 * d[1] + w2 * d[2]) * invArea;
        frags[size_t(layer) * (size_t(rv.W) * rv.H) + pixel] = make_uint2(__float_as_uint(depth), key);
    }
}

__device__ __forceinline__ bool frag_less(uint2 a, uint2 b) {
    const float da = __uint_as_float(a.x), db = __uint_as_float(b.x);
    return da < db || (da == db && a.y < b.y);
}

// sorts a pixel's fragments in place (insertion sort; lists are short)
// (fragment i of a pixel at frags[i * pixels + pixel]: layer-major, so the
// lanes of a wave, neighbouring pixels, read and write neighbouring entries)
__device__ __forceinline__ uint32_t pixel_sorted(const uint32_t* __restrict__ counts, uint2* __restrict__ frags,
                                                 size_t pixel, size_t pixels, uint32_t L) {
    const uint32_t cnt = min(counts[pixel], L);
    uint2* fr = frags + pixel;
    for (uint32_t i = 1; i < cnt; ++i) {
        const uint2 v = fr[i * pixels];
        uint32_t j = i;
        while (j > 0u && frag_less(v, fr[(j - 1u) * pixels])) {
            fr[j * pixels] = fr[(j - 1u) * pixels];
            --j;
        }
        fr[j * pixels] = v;
    }
    return cnt;
}

__device__ __forceinline__ uint32_t frag_side(uint2 v) { return v.y & ((1u << kTileShift) - 1u); }

// upper bound of this direction's directed exchange pairs: 2 (count - 1)
// summed over the pixels
__global__ __launch_bounds__(256) void k_pair_bound(const uint32_t* __restrict__ counts, size_t pixels, uint32_t L,
                                                    unsigned long long* __restrict__ total) {
    unsigned long long s = 0;
    for (size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x; p < pixels; p += size_t(gridDim.x) * blockDim.x) {
        const uint32_t c = min(counts[p], L);
        s += c > 1u ? 2ull * (c - 1u) : 0ull;
    }
    for (int m = 32; m > 0; m >>= 1) s += __shfl_xor(s, m, 64);
    if ((threadIdx.x & 63u) == 0u) atomicAdd(total, s);
}

// graph resolve (buildCpuGraph's pixel loop, core:1361-1396): every
// fragment adds to the denominators of both sides of its facet, the first
// camera side and the last back side see the sky, and consecutive surfaces
// of different facets exchange (back side of the nearer, camera side of the
// farther), emitted as both directed pairs
__global__ __launch_bounds__(256) void k_graph_resolve(const uint32_t* __restrict__ counts, uint2* __restrict__ frags,
                                                       size_t pixels, uint32_t L, uint32_t* __restrict__ den,
                                                       uint32_t* __restrict__ sky,
                                                       unsigned long long* __restrict__ pairs,
                                                       unsigned long long* __restrict__ cursor,
                                                       unsigned long long cap) {
    const size_t pixel = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (pixel >= pixels) return;
    const uint32_t cnt = pixel_sorted(counts, frags, pixel, pixels, L);
    if (cnt == 0u) return;
    const uint2* fr = frags + pixel;
    for (uint32_t i = 0; i < cnt; ++i) {
        const uint32_t s = frag_side(fr[i * pixels]);
        atomicAdd(&den[s], 1u);
        atomicAdd(&den[s ^ 1u], 1u);
    }
    atomicAdd(&sky[frag_side(fr[0])], 1u);
    atomicAdd(&sky[frag_side(fr[(cnt - 1u) * pixels]) ^ 1u], 1u);
    for (uint32_t i = 0; i + 1u < cnt; ++i) {
        const uint32_t a = frag_side(fr[i * pixels]) ^ 1u, b = frag_side(fr[(i + 1u) * pixels]);
        if ((a >> 1) == (b >> 1)) continue;
        const unsigned long long slot = atomicAdd(cursor, 2ull);
        if (slot + 2ull <= cap) {  // cap is the exact bound: always true
            pairs[slot] = (static_cast<unsigned long long>(a) << 32) | b;
            pairs[slot + 1ull] = (static_cast<unsigned long long>(b) << 32) | a;
        }
    }
}

// sunlit resolve (computeCpuSunlit, core:1438-1474): every fragment counts
// for its side, the nearest one (first in submission order among equal
// depths) is visible
__global__ __launch_bounds__(256) void k_sun_resolve(const uint32_t* __restrict__ counts,
                                                     const uint2* __restrict__ frags, size_t pixels, uint32_t L,
                                                     uint32_t* __restrict__ total, uint32_t* __restrict__ visible) {
    const size_t pixel = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (pixel >= pixels) return;
    const uint32_t cnt = min(counts[pixel], L);
    if (cnt == 0u) return;
    const uint2* fr = frags + pixel;
    uint2 best = fr[0];
    for (uint32_t i = 0; i < cnt; ++i) {
        const uint2 v = fr[i * pixels];
        atomicAdd(&total[frag_side(v)], 1u);
        if (frag_less(v, best)) best = v;
    }
    atomicAdd(&visible[frag_side(best)], 1u);
}

__global__ void k_sun_fraction_raster(const uint32_t* __restrict__ total, const uint32_t* __restrict__ visible,
                                      uint32_t S, float* __restrict__ frac) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S) frac[s] = total[s] ? float(visible[s]) / float(total[s]) : 0.0f;
}

__global__ void k_split_keys(const unsigned long long* __restrict__ keys, const uint32_t* __restrict__ vals,
                             uint32_t n, uint32_t* __restrict__ rowlen, uint32_t* __restrict__ col,
                             uint32_t* __restrict__ cnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long k = keys[i];
    col[i] = uint32_t(k & 0xffffffffull);
    cnt[i] = vals[i];
    atomicAdd(&rowlen[uint32_t(k >> 32)], 1u);
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

// rasterises one direction into counts / frags; returns the overflow count
uint32_t raster_pass(fh_context* c, const RasterView& rv, DBuf<uint32_t>& counts, DBuf<uint2>& frags,
                     DBuf<uint32_t>& ovf) {
    const size_t pixels = size_t(rv.W) * rv.H;
    counts.ensure(pixels);
    frags.ensure(pixels * rv.L);
    ovf.ensure(1);
    launch_fill_u32(c, counts.p, pixels, 0u);
    launch_fill_u32(c, ovf.p, 1, 0u);
    const uint64_t n = uint64_t(c->F) * rv.tiles;  // one wave each, 4 per block
    hipLaunchKernelGGL(k_raster, dim3(uint32_t((n + 3u) / 4u)), dim3(256), 0, c->stream, c->facetVerts.p,
                       c->facetNormal.p, c->F, rv, counts.p, frags.p, ovf.p);
    FH_HIP(hipGetLastError());
    uint32_t o = 0;
    FH_HIP(hipMemcpyAsync(&o, ovf.p, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    FH_HIP(hipStreamSynchronize(c->stream));
    return o;
}

}  // namespace

uint32_t launch_raster_graph(fh_context* c, const float* dirs, uint32_t ndirs, uint32_t W, uint32_t H,
                             uint32_t layers) {
    check_raster(c, W, H, layers, nullptr);
    for (uint32_t d = 0; d < ndirs; ++d) check_raster(c, W, H, layers, dirs + 3u * d);
    const uint32_t S = c->S;
    DBuf<uint32_t> counts, ovf;
    DBuf<uint2> frags;
    DBuf<unsigned long long> pairs, sorted, uniq, accKeys, accKeys2, total;
    DBuf<uint32_t> runs, accVals, accVals2, nRuns;
    DBuf<uint8_t> temp;
    c->den.ensure(S);
    c->sky.ensure(S);
    launch_fill_u32(c, c->den.p, S, 0u);
    launch_fill_u32(c, c->sky.p, S, 0u);
    total.ensure(1);
    nRuns.ensure(1);
    size_t accN = 0;
    uint32_t overflow = 0;
    const auto ensure_temp = [&](size_t bytes) {
        if (bytes > temp.cap) temp.ensure(bytes);
    };
    for (uint32_t d = 0; d < ndirs; ++d) {
        const RasterView rv = make_view(c, W, H, layers, dirs + 3u * d);
        overflow += raster_pass(c, rv, counts, frags, ovf);
        const size_t pixels = size_t(rv.W) * rv.H;
        // exact capacity of this direction's pairs
        FH_HIP(hipMemsetAsync(total.p, 0, sizeof(unsigned long long), c->stream));
        hipLaunchKernelGGL(k_pair_bound, dim3(1024), dim3(256), 0, c->stream, counts.p, pixels, rv.L, total.p);
        unsigned long long cap = 0;
        FH_HIP(hipMemcpyAsync(&cap, total.p, sizeof(cap), hipMemcpyDeviceToHost, c->stream));
        FH_HIP(hipStreamSynchronize(c->stream));
        pairs.ensure(cap ? cap : 1);
        FH_HIP(hipMemsetAsync(total.p, 0, sizeof(unsigned long long), c->stream));
        hipLaunchKernelGGL(k_graph_resolve, dim3(uint32_t((pixels + 255u) / 256u)), dim3(256), 0, c->stream,
                           counts.p, frags.p, pixels, rv.L, c->den.p, c->sky.p, pairs.p, total.p, cap);
        FH_HIP(hipGetLastError());
        unsigned long long np = 0;
        FH_HIP(hipMemcpyAsync(&np, total.p, sizeof(np), hipMemcpyDeviceToHost, c->stream));
        FH_HIP(hipStreamSynchronize(c->stream));
        if (np == 0) continue;
        if (np > 0x7fffffffull) fail(FH_OVERFLOW_ERROR, "raster graph: too many exchanges in one direction");
        const int n = int(np);
        // sort this direction's pairs and count each distinct pair
        sorted.ensure(np);
        size_t tb = 0;
        FH_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, pairs.p, sorted.p, n, 0, 64, c->stream));
        ensure_temp(tb);
        FH_HIP(hipcub::DeviceRadixSort::SortKeys(temp.p, tb, pairs.p, sorted.p, n, 0, 64, c->stream));
        uniq.ensure(np);
        runs.ensure(np);
        tb = 0;
        FH_HIP(hipcub::DeviceRunLengthEncode::Encode(nullptr, tb, sorted.p, uniq.p, runs.p, nRuns.p, n, c->stream));
        ensure_temp(tb);
        FH_HIP(hipcub::DeviceRunLengthEncode::Encode(temp.p, tb, sorted.p, uniq.p, runs.p, nRuns.p, n, c->stream));
        uint32_t nu = 0;
        FH_HIP(hipMemcpyAsync(&nu, nRuns.p, sizeof(nu), hipMemcpyDeviceToHost, c->stream));
        FH_HIP(hipStreamSynchronize(c->stream));
        // append to the running (pair, count) list
        if (accN + nu > accKeys.cap) {
            const size_t cap2 = std::max(accN + nu, 2 * accKeys.cap);
            accKeys2.ensure(cap2);
            accVals2.ensure(cap2);
            if (accN) {
                FH_HIP(hipMemcpyAsync(accKeys2.p, accKeys.p, accN * sizeof(unsigned long long),
                                      hipMemcpyDeviceToDevice, c->stream));
                FH_HIP(hipMemcpyAsync(accVals2.p, accVals.p, accN * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                                      c->stream));
            }
            accKeys.swap(accKeys2);
            accVals.swap(accVals2);
        }
        FH_HIP(hipMemcpyAsync(accKeys.p + accN, uniq.p, nu * sizeof(unsigned long long), hipMemcpyDeviceToDevice,
                              c->stream));
        FH_HIP(hipMemcpyAsync(accVals.p + accN, runs.p, nu * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
        accN += nu;
    }
    counts.release();
    frags.release();
    pairs.release();
    // merge the directions: sort by pair, sum the counts
    uint32_t nnz = 0;
    DBuf<unsigned long long> keysOut, keysFinal;
    DBuf<uint32_t> valsOut, valsFinal;
    if (accN) {
        if (accN > 0x7fffffffull) fail(FH_OVERFLOW_ERROR, "raster graph: too many exchanges");
        const int n = int(accN);
        keysOut.ensure(accN);
        valsOut.ensure(accN);
        size_t tb = 0;
        FH_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, accKeys.p, keysOut.p, accVals.p, valsOut.p, n, 0, 64,
                                                  c->stream));
        ensure_temp(tb);
        FH_HIP(hipcub::DeviceRadixSort::SortPairs(temp.p, tb, accKeys.p, keysOut.p, accVals.p, valsOut.p, n, 0, 64,
                                                  c->stream));
        keysFinal.ensure(accN);
        valsFinal.ensure(accN);
        tb = 0;
        FH_HIP(hipcub::DeviceReduce::ReduceByKey(nullptr, tb, keysOut.p, keysFinal.p, valsOut.p, valsFinal.p, nRuns.p,
                                                 hipcub::Sum(), n, c->stream));
        ensure_temp(tb);
        FH_HIP(hipcub::DeviceReduce::ReduceByKey(temp.p, tb, keysOut.p, keysFinal.p, valsOut.p, valsFinal.p, nRuns.p,
                                                 hipcub::Sum(), n, c->stream));
        FH_HIP(hipMemcpyAsync(&nnz, nRuns.p, sizeof(nnz), hipMemcpyDeviceToHost, c->stream));
        FH_HIP(hipStreamSynchronize(c->stream));
    }
    // CSR: rows from the pair keys (rows and the columns within a row come
    // out sorted)
    c->nnz = nnz;
    alloc_csr_device(c, nnz);
    c->rowlen.ensure(S);
    launch_fill_u32(c, c->rowlen.p, S, 0u);
    if (nnz)
        hipLaunchKernelGGL(k_split_keys, dim3((nnz + 255u) / 256u), dim3(256), 0, c->stream, keysFinal.p,
                           valsFinal.p, nnz, c->rowlen.p, c->col.p, c->cnt.p);
    FH_HIP(hipGetLastError());
    c->rowptr.ensure(size_t(S) + 1u);
    launch_scan(c, c->rowlen.p, S, c->rowptr.p);
    FH_HIP(hipStreamSynchronize(c->stream));
    return overflow;
}

uint32_t launch_raster_sunlit(fh_context* c, const float sun[3], uint32_t W, uint32_t H, uint32_t layers,
                              float* frac) {
    check_raster(c, W, H, layers, sun);
    const uint32_t S = c->S;
    DBuf<uint32_t> counts, ovf, total, visible;
    DBuf<uint2> frags;
    const RasterView rv = make_view(c, W, H, layers, sun);
    const uint32_t overflow = raster_pass(c, rv, counts, frags, ovf);
    total.ensure(S);
    visible.ensure(S);
    launch_fill_u32(c, total.p, S, 0u);
    launch_fill_u32(c, visible.p, S, 0u);
    const size_t pixels = size_t(rv.W) * rv.H;
    hipLaunchKernelGGL(k_sun_resolve, dim3(uint32_t((pixels + 255u) / 256u)), dim3(256), 0, c->stream, counts.p,
                       frags.p, pixels, rv.L, total.p, visible.p);
    hipLaunchKernelGGL(k_sun_fraction_raster, dim3((S + 255u) / 256u), dim3(256), 0, c->stream, total.p, visible.p,
                       S, frac);
    FH_HIP(hipGetLastError());
    FH_HIP(hipStreamSynchronize(c->stream));
    return overflow;
}

}  // namespace fh
