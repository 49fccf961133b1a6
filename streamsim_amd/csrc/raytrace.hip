// raytrace.hip -- form factors by seeded cosine-weighted ray casting and the
// sunlit fraction by seeded sun rays, on gfx950.
//
// Replaces the reference's A-buffer raster + depth-sort resolve + host CSR
// compaction (shader/facetrt/visibility.{vert,frag}, visibility_resolve.comp,
// direct_resolve.comp; facetrt_vulkan.cpp:866-1158) with:
//   k_trace_graph  one lane per ray: child-pair BVH traversal with the node
//                  stack in LDS, watertight triangle test, closest hit;
//   k_rows         one workgroup per side: bitonic sort of the R hit sides in
//                  LDS, run-length encode -> (neighbour, count), sky count;
//   scan + k_compact  exclusive scan of row lengths -> CSR; weights count/R.
// Everything here is compiled with FP contraction off so that every float
// expression rounds exactly as written; the host replay in the oracle
// evaluates the same expressions and must produce identical hit sides.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>

#include "engine.hpp"
#include "sweep_common.hpp"

namespace fh {

constexpr uint32_t kSky = 0xffffffffu;
constexpr int kTraceBlock = 128;

// ---- counter RNG (PCG RXS-M-XS output permutation as a hash) -------------
__device__ __forceinline__ uint32_t pcg_hash(uint32_t x) {
    const uint32_t state = x * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
__device__ __forceinline__ float u01(uint32_t h) { return float(h >> 8) * 5.9604644775390625e-8f; }
__device__ __forceinline__ uint32_t side_key(uint32_t seed, uint32_t side) { return pcg_hash(seed ^ pcg_hash(side)); }
__device__ __forceinline__ uint32_t ray_rand(uint32_t key, uint32_t ray, uint32_t k) {
    return pcg_hash(key ^ pcg_hash(ray * 64u + k));
}

// uniform point on the triangle of facet f (square folded onto the triangle)
__device__ __forceinline__ void triangle_point(const float* __restrict__ fv, uint32_t f, float u, float w,
                                               float o[3]) {
    if (u + w > 1.0f) {
        u = 1.0f - u;
        w = 1.0f - w;
    }
    const float* t = fv + 9 * size_t(f);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float e1 = t[3 + i] - t[i];
        const float e2 = t[6 + i] - t[i];
        o[i] = t[i] + (u * e1 + w * e2);
    }
}

struct Ray {
    float o[3], d[3];
    float inv[3];
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

__device__ __forceinline__ float comp(const float v[3], int k) { return k == 0 ? v[0] : (k == 1 ? v[1] : v[2]); }

__device__ __forceinline__ void ray_setup(Ray& r) {
    int kz = 0;
    if (fabsf(r.d[1]) > fabsf(r.d[kz])) kz = 1;
    if (fabsf(r.d[2]) > fabsf(comp(r.d, kz))) kz = 2;
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    if (comp(r.d, kz) < 0.0f) {
        const int tmp = kx;
        kx = ky;
        ky = tmp;
    }
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    const float dz = comp(r.d, kz);
    r.Sx = comp(r.d, kx) / dz;
    r.Sy = comp(r.d, ky) / dz;
    r.Sz = 1.0f / dz;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float di = r.d[i];
        if (fabsf(di) < 1.0e-20f) di = copysignf(1.0e-20f, di);
        r.inv[i] = 1.0f / di;
    }
}

// Watertight ray/triangle test (Woop, Benthin, Wald 2013) with the
// double-precision fallback for zero edge functions.
__device__ __forceinline__ bool watertight(const Ray& r, float4 p0, float4 p1, float4 p2, float tmin, float& tout) {
    const float A[3] = {p0.x - r.o[0], p0.y - r.o[1], p0.z - r.o[2]};
    const float B[3] = {p1.x - r.o[0], p1.y - r.o[1], p1.z - r.o[2]};
    const float C[3] = {p2.x - r.o[0], p2.y - r.o[1], p2.z - r.o[2]};
    const float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
    const float Ax = comp(A, r.kx) - r.Sx * Akz;
    const float Ay = comp(A, r.ky) - r.Sy * Akz;
    const float Bx = comp(B, r.kx) - r.Sx * Bkz;
    const float By = comp(B, r.ky) - r.Sy * Bkz;
    const float Cx = comp(C, r.kx) - r.Sx * Ckz;
    const float Cy = comp(C, r.ky) - r.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = float(double(Cx) * double(By) - double(Cy) * double(Bx));
        V = float(double(Ax) * double(Cy) - double(Ay) * double(Cx));
        W = float(double(Bx) * double(Ay) - double(By) * double(Ax));
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = r.Sz * Akz;
    const float Bz = r.Sz * Bkz;
    const float Cz = r.Sz * Ckz;
    const float T = U * Az + V * Bz + W * Cz;
    const float t = T / det;
    if (!(t > tmin)) return false;
    tout = t;
    return true;
}

__device__ __forceinline__ bool slab(const Ray& r, float lx, float hx, float ly, float hy, float lz, float hz,
                                     float tlimit, float& tnear) {
    const float ax = (lx - r.o[0]) * r.inv[0], bx = (hx - r.o[0]) * r.inv[0];
    const float ay = (ly - r.o[1]) * r.inv[1], by = (hy - r.o[1]) * r.inv[1];
    const float az = (lz - r.o[2]) * r.inv[2], bz = (hz - r.o[2]) * r.inv[2];
    const float tn = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), 0.0f));
    const float tf = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), tlimit));
    tnear = tn;
    return tn <= tf;
}

// Closest hit (ANY == false) or any hit (ANY == true).  Returns the facet or
// kSky.  stk points at this lane's column of the LDS stack (stride kTraceBlock).
template <bool ANY>
__device__ uint32_t traverse(const BvhNode* __restrict__ nodes, const float4* __restrict__ tris, const Ray& r,
                             uint32_t skip, float tmin, int* stk, float& tbest, uint32_t bestF = kSky) {
    int sp = 0;
    int cur = 0;
    while (true) {
        const float4 nx = nodes[cur].x, ny = nodes[cur].y, nz = nodes[cur].z;
        const int4 meta = nodes[cur].meta;
        const float limit = tbest * 1.0001f + 1.0e-6f;
        float tn0, tn1;
        bool h0 = (meta.z > 0 || meta.x >= 0) && slab(r, nx.x, nx.y, ny.x, ny.y, nz.x, nz.y, limit, tn0);
        bool h1 = (meta.w > 0 || meta.y >= 0) && slab(r, nx.z, nx.w, ny.z, ny.w, nz.z, nz.w, limit, tn1);
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const bool hc = c == 0 ? h0 : h1;
            const int cnt = c == 0 ? meta.z : meta.w;
            if (hc && cnt > 0) {
                const int first = c == 0 ? meta.x : meta.y;
                for (int i = first; i < first + cnt; ++i) {
                    const float4 p0 = tris[3 * size_t(i)];
                    const uint32_t f = __float_as_uint(p0.w);
                    if (f == skip) continue;
                    float t;
                    if (watertight(r, p0, tris[3 * size_t(i) + 1], tris[3 * size_t(i) + 2], tmin, t)) {
                        if (t < tbest || (t == tbest && f < bestF)) {
                            tbest = t;
                            bestF = f;
                        }
                    }
                }
                if (ANY && bestF != kSky) return bestF;
                if (c == 0) h0 = false; else h1 = false;
            }
        }
        if (h0 && h1) {
            const bool nearFirst = tn0 <= tn1;
            stk[sp * kTraceBlock] = nearFirst ? meta.y : meta.x;
            ++sp;
            cur = nearFirst ? meta.x : meta.y;
        } else if (h0) {
            cur = meta.x;
        } else if (h1) {
            cur = meta.y;
        } else {
            if (sp == 0) break;
            --sp;
            cur = stk[sp * kTraceBlock];
        }
    }
    return bestF;
}

// Periodic tiles (sh/visibility.vert:15-26): tile 0 is the scene, then the
// neighbour list; tile t holds a copy shifted by (dx sizeX, 0, dz sizeZ).
__constant__ int kTileOffset[21][2] = {{0, 0},   {-1, 0},  {1, 0},   {0, -1}, {0, 1},  {-1, -1}, {-1, 1},
                                       {1, -1},  {1, 1},   {-2, 0},  {2, 0},  {0, -2}, {0, 2},   {-2, -1},
                                       {-2, 1},  {2, -1},  {2, 1},   {-1, -2}, {1, -2}, {-1, 2},  {1, 2}};

struct Periodic {
    uint32_t tiles;  // 1 + neighbour count
    float sx, sz;
};

// Closest (or any) hit over the periodic copies: the ray is traced against
// the scene once per tile with its origin moved by minus the tile offset
// (the copies of the origin facet are ordinary surfaces), keeping the
// running best (t, facet) across tiles; the root boxes cull tiles the ray
// does not reach before the best hit.
template <bool ANY>
__device__ uint32_t trace_tiles(const BvhNode* __restrict__ nodes, const float4* __restrict__ tris, const Ray& r,
                                uint32_t facet, float tmin, int* stk, const Periodic& p, float& tbest) {
    tbest = FLT_MAX;
    uint32_t best = traverse<ANY>(nodes, tris, r, facet, tmin, stk, tbest);
    for (uint32_t t = 1; t < p.tiles; ++t) {
        if (ANY && best != kSky) break;
        Ray rt = r;
        rt.o[0] = r.o[0] - float(kTileOffset[t][0]) * p.sx;
        rt.o[2] = r.o[2] - float(kTileOffset[t][1]) * p.sz;
        best = traverse<ANY>(nodes, tris, rt, kSky, tmin, stk, tbest, best);
    }
    return best;
}

// ---- form-factor rays: one lane per (side, ray) ---------------------------
__global__ __launch_bounds__(kTraceBlock) void k_trace_graph(const BvhNode* __restrict__ nodes,
                                                             const float4* __restrict__ tris,
                                                             const float* __restrict__ fverts,
                                                             const float4* __restrict__ fnorm, uint32_t S,
                                                             uint32_t R, uint32_t seed, float tmin, Periodic per,
                                                             uint32_t* __restrict__ hits) {
    // traversal stack in LDS, (BVH depth + 1) x kTraceBlock ints sized at
    // launch: a shallower tree leaves room for more resident waves
    extern __shared__ int stack[];
    const uint64_t gid = uint64_t(blockIdx.x) * kTraceBlock + threadIdx.x;
    if (gid >= uint64_t(S) * R) return;
    const uint32_t side = uint32_t(gid / R);
    const uint32_t ray = uint32_t(gid - uint64_t(side) * R);
    const uint32_t f = side >> 1;
    const uint32_t key = side_key(seed, side);
    Ray r;
    triangle_point(fverts, f, u01(ray_rand(key, ray, 0u)), u01(ray_rand(key, ray, 1u)), r.o);
    float x = 0.0f, y = 0.0f, z = 1.0f;
    for (uint32_t k = 2u; k < 62u; k += 2u) {
        const float a = 2.0f * u01(ray_rand(key, ray, k)) - 1.0f;
        const float b = 2.0f * u01(ray_rand(key, ray, k + 1u)) - 1.0f;
        const float r2 = a * a + b * b;
        if (r2 < 1.0f) {
            x = a;
            y = b;
            z = sqrtf(1.0f - r2);
            break;
        }
    }
    const float4 nf = fnorm[f];
    const float sgnSide = (side & 1u) ? -1.0f : 1.0f;
    const float n[3] = {sgnSide * nf.x, sgnSide * nf.y, sgnSide * nf.z};
    const float sgn = copysignf(1.0f, n[2]);
    const float a = -1.0f / (sgn + n[2]);
    const float b = n[0] * n[1] * a;
    const float t1[3] = {1.0f + sgn * n[0] * n[0] * a, sgn * b, -sgn * n[0]};
    const float t2[3] = {b, sgn + n[1] * n[1] * a, -n[1]};
#pragma unroll
    for (int i = 0; i < 3; ++i) r.d[i] = (x * t1[i] + y * t2[i]) + z * n[i];
    ray_setup(r);
    float tbest;
    const uint32_t hf = trace_tiles<false>(nodes, tris, r, f, tmin, stack + threadIdx.x, per, tbest);
    uint32_t out = kSky;
    if (hf != kSky) {
        const float4 hn = fnorm[hf];
        const float dd = r.d[0] * hn.x + r.d[1] * hn.y + r.d[2] * hn.z;
        out = hf * 2u + (dd > 0.0f ? 1u : 0u);
    }
    hits[gid] = out;
}

// ---- sun rays: one lane per (side, sample) ---------------------------------
__global__ __launch_bounds__(kTraceBlock) void k_trace_sun(const BvhNode* __restrict__ nodes,
                                                           const float4* __restrict__ tris,
                                                           const float* __restrict__ fverts,
                                                           const float4* __restrict__ fnorm, uint32_t S,
                                                           uint32_t K, uint32_t seed, float sx, float sy, float sz,
                                                           float tmin, Periodic per, uint32_t* __restrict__ lit) {
    // traversal stack in LDS, (BVH depth + 1) x kTraceBlock ints sized at
    // launch: a shallower tree leaves room for more resident waves
    extern __shared__ int stack[];
    const uint64_t gid = uint64_t(blockIdx.x) * kTraceBlock + threadIdx.x;
    if (gid >= uint64_t(S) * K) return;
    const uint32_t side = uint32_t(gid / K);
    const uint32_t j = uint32_t(gid - uint64_t(side) * K);
    const uint32_t f = side >> 1;
    const float4 nf = fnorm[f];
    const float sg = (side & 1u) ? -1.0f : 1.0f;
    const float ns[3] = {sg * nf.x, sg * nf.y, sg * nf.z};
    if (!(ns[0] * sx + ns[1] * sy + ns[2] * sz > 0.0f)) return;
    const uint32_t key = side_key(seed ^ 0x5bd1e995u, side);
    Ray r;
    triangle_point(fverts, f, u01(ray_rand(key, j, 0u)), u01(ray_rand(key, j, 1u)), r.o);
    r.d[0] = sx;
    r.d[1] = sy;
    r.d[2] = sz;
    ray_setup(r);
    float tbest;
    const uint32_t hf = trace_tiles<true>(nodes, tris, r, f, tmin, stack + threadIdx.x, per, tbest);
    const bool sky = hf == kSky;
    if (64u % K == 0u) {
        // the K rays of a side are K consecutive lanes of one wave (all
        // active: the facing test is per side): one add per side
        const uint64_t m = __ballot(sky);
        const uint32_t lane = threadIdx.x & 63u, first = lane - lane % K;
        if (lane == first) {
            const uint64_t bits = K == 64u ? m : (m >> first) & ((uint64_t(1) << K) - 1u);
            const uint32_t n = uint32_t(__popcll(bits));
            if (n) atomicAdd(&lit[side], n);
        }
    } else if (sky) {
        atomicAdd(&lit[side], 1u);
    }
}

// ---- sun rays, one packet per wave -----------------------------------------
// The K rays of a side are parallel and start on one facet, and the 64 / K
// sides of a wave are neighbours: the wave walks the BVH as one packet (one
// node stream, its stack in LDS, a child visited when the box test of any
// searching lane passes, triangles tested by the lanes whose own box test
// passed; a lane stops searching at its first hit).  Boxes are conservative
// and the triangle test exact, so each ray's answer (some triangle beyond
// tmin, own facet excluded) is the one of its own traversal.
constexpr int kSunBlock = 256;

__device__ void packet_any(const BvhNode* __restrict__ nodes, const float4* __restrict__ tris, const Ray& r,
                           uint32_t skip, float tmin, int* stk, bool& on, bool& hit) {
    const float limit = FLT_MAX * 1.0001f + 1.0e-6f;  // no hit yet (any-hit traversal)
    int sp = 0;
    int cur = 0;
    while (true) {
        cur = __builtin_amdgcn_readfirstlane(cur);
        const float4 nx = nodes[cur].x, ny = nodes[cur].y, nz = nodes[cur].z;
        const int4 meta = nodes[cur].meta;
        float tn0, tn1;
        bool h0 = on && (meta.z > 0 || meta.x >= 0) && slab(r, nx.x, nx.y, ny.x, ny.y, nz.x, nz.y, limit, tn0);
        bool h1 = on && (meta.w > 0 || meta.y >= 0) && slab(r, nx.z, nx.w, ny.z, ny.w, nz.z, nz.w, limit, tn1);
        bool any0 = __ballot(h0) != 0u, any1 = __ballot(h1) != 0u;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const bool anyc = c == 0 ? any0 : any1;
            const int cnt = c == 0 ? meta.z : meta.w;
            if (anyc && cnt > 0) {
                const bool hc = c == 0 ? h0 : h1;
                const int first = c == 0 ? meta.x : meta.y;
                for (int i = first; i < first + cnt; ++i) {
                    const float4 p0 = tris[3 * size_t(i)];
                    const uint32_t f = __float_as_uint(p0.w);
                    float t;
                    if (hc && on && f != skip &&
                        watertight(r, p0, tris[3 * size_t(i) + 1], tris[3 * size_t(i) + 2], tmin, t)) {
                        hit = true;
                        on = false;
                    }
                }
                if (__ballot(on) == 0u) return;
                if (c == 0) any0 = false; else any1 = false;
            }
        }
        if (any0 && any1) {
            stk[sp] = meta.y;
            ++sp;
            cur = meta.x;
        } else if (any0) {
            cur = meta.x;
        } else if (any1) {
            cur = meta.y;
        } else {
            if (sp == 0) break;
            --sp;
            cur = stk[sp];
        }
    }
}

__global__ __launch_bounds__(kSunBlock) void k_trace_sun_packet(const BvhNode* __restrict__ nodes,
                                                                const float4* __restrict__ tris,
                                                                const float* __restrict__ fverts,
                                                                const float4* __restrict__ fnorm, uint32_t S,
                                                                uint32_t K, uint32_t seed, float sx, float sy,
                                                                float sz, float tmin, Periodic per,
                                                                uint32_t* __restrict__ lit) {
    __shared__ int stacks[kSunBlock / 64][kBvhMaxDepth];
    int* stk = stacks[threadIdx.x >> 6];
    // lanes run over (facet, sample): at most one side of a facet faces the
    // sun (n.s > 0: side 2f, -n.s > 0: side 2f + 1), so every lane of a wave
    // traces a ray of a facing side instead of half of them idling on the
    // sides that face away (their fraction stays 0)
    const uint32_t F = S >> 1;
    const uint64_t gid = uint64_t(blockIdx.x) * kSunBlock + threadIdx.x;
    bool on = gid < uint64_t(F) * K;
    const uint32_t f = on ? uint32_t(gid / K) : F - 1u;
    const uint32_t j = uint32_t(gid % K);
    const float4 nf = fnorm[f];
    const float d0 = nf.x * sx + nf.y * sy + nf.z * sz;
    const float d1 = (-nf.x) * sx + (-nf.y) * sy + (-nf.z) * sz;
    const uint32_t side = 2u * f + (d0 > 0.0f ? 0u : 1u);
    on = on && (d0 > 0.0f || d1 > 0.0f);
    const uint32_t key = side_key(seed ^ 0x5bd1e995u, side);
    Ray r;
    triangle_point(fverts, f, u01(ray_rand(key, j, 0u)), u01(ray_rand(key, j, 1u)), r.o);
    r.d[0] = sx;
    r.d[1] = sy;
    r.d[2] = sz;
    ray_setup(r);
    bool hit = false;
    const bool lane = on;  // sides facing the sun
    for (uint32_t t = 0; t < per.tiles; ++t) {
        if (__ballot(on) == 0u) break;
        Ray rt = r;
        if (t) {
            rt.o[0] = r.o[0] - float(kTileOffset[t][0]) * per.sx;
            rt.o[2] = r.o[2] - float(kTileOffset[t][1]) * per.sz;
        }
        packet_any(nodes, tris, rt, t == 0 ? f : kSky, tmin, stk, on, hit);
    }
    const bool sky = lane && !hit;
    // K consecutive lanes hold the rays of one side: one add per side
    const uint64_t m = __ballot(sky);
    const uint32_t l = threadIdx.x & 63u, firstLane = l - l % K;
    if (l == firstLane && gid < uint64_t(F) * K) {
        const uint64_t bits = K == 64u ? m : (m >> firstLane) & ((uint64_t(1) << K) - 1u);
        const uint32_t n = uint32_t(__popcll(bits));
        if (n) atomicAdd(&lit[side], n);
    }
}

__global__ void k_sun_fraction(const uint32_t* __restrict__ lit, uint32_t S, uint32_t K, float* __restrict__ frac) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S) frac[s] = float(lit[s]) / float(K);
}

// ---- per-side sort + run-length encode -------------------------------------
// Block scan (exclusive) of NP flags held in LDS with 256 threads.
template <int NP>
__device__ void block_exclusive_scan(uint32_t* v, uint32_t* part) {
    constexpr int IPT = NP >= 256 ? NP / 256 : 1;
    const int t = threadIdx.x;
    const int base = t * IPT;
    uint32_t sum = 0;
    if (base < NP)
        for (int i = 0; i < IPT; ++i) sum += v[base + i];
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const uint32_t add = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;  // exclusive prefix of this thread
    if (base < NP)
        for (int i = 0; i < IPT; ++i) {
            const uint32_t x = v[base + i];
            v[base + i] = run;
            run += x;
        }
    __syncthreads();
}

template <int NP>
__global__ __launch_bounds__(256) void k_rows(uint32_t* __restrict__ hits, uint16_t* __restrict__ cnt16, uint32_t S,
                                               uint32_t R, uint32_t* __restrict__ rowlen, uint32_t* __restrict__ skyc) {
    __shared__ uint32_t key[NP];
    __shared__ uint32_t rank[NP];
    __shared__ uint32_t start[NP];
    __shared__ uint32_t part[256];
    const uint32_t side = blockIdx.x;
    const size_t base = size_t(side) * R;
    for (int i = threadIdx.x; i < NP; i += 256) key[i] = uint32_t(i) < R ? hits[base + i] : kSky;
    __syncthreads();
    for (int k = 2; k <= NP; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < NP; i += 256) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const uint32_t a = key[i], b = key[ixj];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        key[i] = b;
                        key[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < NP; i += 256)
        rank[i] = (key[i] != kSky && (i == 0 || key[i] != key[i - 1])) ? 1u : 0u;
    __syncthreads();
    // total unique before the scan overwrites the flags
    __shared__ uint32_t nUnique, nValid;
    if (threadIdx.x == 0) {
        nUnique = 0;
        nValid = 0;
    }
    __syncthreads();
    uint32_t lu = 0, lv = 0;
    for (int i = threadIdx.x; i < NP; i += 256) {
        lu += rank[i];
        lv += key[i] != kSky ? 1u : 0u;
    }
    atomicAdd(&nUnique, lu);
    atomicAdd(&nValid, lv);
    __syncthreads();
    block_exclusive_scan<NP>(rank, part);
    for (int i = threadIdx.x; i < NP; i += 256)
        if (key[i] != kSky && (i == 0 || key[i] != key[i - 1])) start[rank[i]] = uint32_t(i);
    __syncthreads();
    // A tail at i belongs to the run whose head is the last head at or before
    // i; heads are rank-numbered, so the run index is (#heads in [0, i]) - 1.
    for (int i = threadIdx.x; i < NP; i += 256) {
        if (key[i] != kSky && (i == NP - 1 || key[i + 1] != key[i])) {
            const bool isHead = (i == 0 || key[i] != key[i - 1]);
            const uint32_t run = isHead ? rank[i] : rank[i] - 1u;
            hits[base + run] = key[i];
            cnt16[base + run] = uint16_t(uint32_t(i) + 1u - start[run]);
        }
    }
    if (threadIdx.x == 0) {
        rowlen[side] = nUnique;
        skyc[side] = R - nValid;
    }
}

// ---- exclusive scan of n lengths (each clamped to clampMax) into n+1 offsets
__global__ __launch_bounds__(1024) void k_scan_blocks(const uint32_t* __restrict__ in, uint32_t n, uint32_t clampMax,
                                                      uint32_t* __restrict__ out, uint32_t* __restrict__ blockSums) {
    __shared__ uint32_t s[1024];
    const uint32_t i = blockIdx.x * 1024u + threadIdx.x;
    const uint32_t x = i < n ? min(in[i], clampMax) : 0u;
    s[threadIdx.x] = x;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const uint32_t add = threadIdx.x >= uint32_t(off) ? s[threadIdx.x - off] : 0u;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    if (i < n) out[i] = s[threadIdx.x] - x;
    if (threadIdx.x == 1023) blockSums[blockIdx.x] = s[1023];
}

__global__ __launch_bounds__(1024) void k_scan_sums(uint32_t* __restrict__ sums, uint32_t nb, uint32_t* __restrict__ total) {
    __shared__ uint32_t s[1024];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t b = 0; b < nb; b += 1024) {
        const uint32_t i = b + threadIdx.x;
        const uint32_t x = i < nb ? sums[i] : 0u;
        s[threadIdx.x] = x;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const uint32_t add = threadIdx.x >= uint32_t(off) ? s[threadIdx.x - off] : 0u;
            __syncthreads();
            s[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < nb) sums[i] = carry + s[threadIdx.x] - x;
        __syncthreads();
        if (threadIdx.x == 1023) carry += s[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void k_scan_add(uint32_t* __restrict__ out, uint32_t n, const uint32_t* __restrict__ sums,
                           const uint32_t* __restrict__ total) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += sums[i / 1024u];
    if (i == n) out[n] = *total;
}

// ---- CSR fill: one wavefront per side --------------------------------------
__global__ __launch_bounds__(256) void k_compact(const uint32_t* __restrict__ hits, const uint16_t* __restrict__ cnt16,
                                                 const uint32_t* __restrict__ rowptr, uint32_t S, uint32_t R,
                                                 uint32_t* __restrict__ col, uint32_t* __restrict__ cnt,
                                                 uint32_t* __restrict__ den) {
    const uint32_t side = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (side >= S) return;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t b = rowptr[side], e = rowptr[side + 1];
    const size_t src = size_t(side) * R;
    for (uint32_t k = b + lane; k < e; k += 64u) {
        col[k] = hits[src + (k - b)];
        cnt[k] = cnt16[src + (k - b)];
    }
    if (lane == 0) den[side] = R;
}

// weights for an uploaded integer graph (den per side)
__global__ __launch_bounds__(256) void k_weights(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ cnt,
                                                 const uint32_t* __restrict__ den, const uint32_t* __restrict__ sky,
                                                 uint32_t S, float* __restrict__ val32, uint16_t* __restrict__ val16,
                                                 float* __restrict__ fsky, uint32_t* __restrict__ closureBits) {
    const uint32_t side = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (side >= S) return;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t b = rowptr[side], e = rowptr[side + 1];
    const uint32_t d = den[side];
    uint32_t sum = 0;
    for (uint32_t k = b + lane; k < e; k += 64u) {
        const uint32_t c = cnt[k];
        sum += c;
        const float w = d ? float(c) / float(d) : 0.0f;
        val32[k] = w;
        val16[k] = f32_to_bf16(w);
    }
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
    if (lane == 0) {
        fsky[side] = d ? float(sky[side]) / float(d) : 1.0f;
        if (d) {
            const float closure = fabsf(float(uint64_t(sky[side]) + sum) / float(d) - 1.0f);
            atomicMax(closureBits, __float_as_uint(closure));
        }
    }
}

// ---- host launchers ---------------------------------------------------------
static Periodic periodic(const fh_context* c) {
    Periodic p{1u + c->cfg.periodicNeighborCount, c->cfg.periodicSizeX, c->cfg.periodicSizeZ};
    return p;
}

// LDS bytes of the traversal stack: one entry per internal level of the tree
static size_t stack_bytes(const fh_context* c) {
    const int levels = std::min(std::max(c->bvhDepth + 1, 1), kBvhMaxDepth);
    return size_t(levels) * kTraceBlock * sizeof(int);
}

void launch_trace_graph(fh_context* c, uint32_t R, uint32_t seed, float tmin) {
    const uint64_t n = uint64_t(c->S) * R;
    const uint64_t blocks = (n + kTraceBlock - 1) / kTraceBlock;
    hipLaunchKernelGGL(k_trace_graph, dim3(uint32_t(blocks)), dim3(kTraceBlock), stack_bytes(c), c->stream, c->nodes.p,
                       c->tris.p,
                       c->facetVerts.p, c->facetNormal.p, c->S, R, seed, tmin, periodic(c), c->hits.p);
    FH_HIP(hipGetLastError());
}

void launch_rows(fh_context* c, uint32_t R) {
    const dim3 grid(c->S), block(256);
#define FH_ROWS(NP) hipLaunchKernelGGL(k_rows<NP>, grid, block, 0, c->stream, c->hits.p, c->cnt16.p, c->S, R, c->rowlen.p, c->sky.p)
    if (R <= 64) FH_ROWS(64);
    else if (R <= 128) FH_ROWS(128);
    else if (R <= 256) FH_ROWS(256);
    else if (R <= 512) FH_ROWS(512);
    else if (R <= 1024) FH_ROWS(1024);
    else FH_ROWS(2048);
#undef FH_ROWS
    FH_HIP(hipGetLastError());
}

void launch_scan(fh_context* c, const uint32_t* in, uint32_t n, uint32_t* out, uint32_t clampMax) {
    const uint32_t nblocks = (n + 1023u) / 1024u;
    c->scanTmp.ensure(size_t(nblocks) + 1u);
    hipLaunchKernelGGL(k_scan_blocks, dim3(nblocks ? nblocks : 1), dim3(1024), 0, c->stream, in, n, clampMax, out,
                       c->scanTmp.p);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, c->stream, c->scanTmp.p, nblocks, c->scanTmp.p + nblocks);
    hipLaunchKernelGGL(k_scan_add, dim3((n + 1u + 255u) / 256u), dim3(256), 0, c->stream, out, n, c->scanTmp.p,
                       c->scanTmp.p + nblocks);
    FH_HIP(hipGetLastError());
}

void launch_compact(fh_context* c, uint32_t R) {
    hipLaunchKernelGGL(k_compact, dim3((c->S + 3u) / 4u), dim3(256), 0, c->stream, c->hits.p, c->cnt16.p, c->rowptr.p,
                       c->S, R, c->col.p, c->cnt.p, c->den.p);
    FH_HIP(hipGetLastError());
}

// packed entries col | cnt << bits
__global__ void k_pack(const uint32_t* __restrict__ col, const uint32_t* __restrict__ cnt, size_t n, uint32_t bits,
                       uint32_t* __restrict__ pk) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) pk[i] = col[i] | (cnt[i] << bits);
}

static uint32_t bits_for(uint64_t v) {  // bits to hold 0 .. v
    uint32_t b = 1;
    while (b < 32 && (uint64_t(1) << b) <= v) ++b;
    return b;
}

void launch_weights(fh_context* c, uint32_t* closureBits, uint32_t maxCount, bool packable) {
    hipLaunchKernelGGL(k_weights, dim3((c->S + 3u) / 4u), dim3(256), 0, c->stream, c->rowptr.p, c->cnt.p, c->den.p,
                       c->sky.p, c->S, c->val32.p, c->val16.p, c->fsky.p, closureBits);
    FH_HIP(hipGetLastError());
    const uint32_t colBits = bits_for(c->S ? c->S - 1u : 0u), cntBits = bits_for(maxCount);
    c->pkBits = 0;
    if (packable && colBits + cntBits <= 32u) {
        const size_t n = size_t(c->nnz);
        c->pk.ensure(n + kCsrPad);
        FH_HIP(hipMemsetAsync(c->pk.p + n, 0, kCsrPad * sizeof(uint32_t), c->stream));
        if (n) {
            hipLaunchKernelGGL(k_pack, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, c->stream, c->col.p, c->cnt.p,
                               n, colBits, c->pk.p);
            FH_HIP(hipGetLastError());
        }
        c->pkBits = colBits;
    } else {
        c->pk.release();
    }
}

void launch_sun(fh_context* c, float sx, float sy, float sz, uint32_t* lit, float* frac) {
    const uint32_t K = c->cfg.sunRaysPerSide;
    const uint64_t n = uint64_t(c->S) * K;
    static const bool packets = [] {
        const char* e = std::getenv("FH_SUN_PACKETS");  // 0: one traversal per ray (tuning knob)
        return e ? std::atoi(e) != 0 : true;
    }();
    if (packets && 64u % K == 0u && c->bvhDepth < kBvhMaxDepth) {
        const uint64_t blocks = (uint64_t(c->F) * K + kSunBlock - 1) / kSunBlock;  // lanes per (facet, sample)
        hipLaunchKernelGGL(k_trace_sun_packet, dim3(uint32_t(blocks)), dim3(kSunBlock), 0, c->stream, c->nodes.p,
                           c->tris.p, c->facetVerts.p, c->facetNormal.p, c->S, K, c->cfg.seed, sx, sy, sz,
                           c->cfg.rayEpsilon, periodic(c), lit);
    } else {
        const uint64_t blocks = (n + kTraceBlock - 1) / kTraceBlock;
        hipLaunchKernelGGL(k_trace_sun, dim3(uint32_t(blocks)), dim3(kTraceBlock), stack_bytes(c), c->stream,
                           c->nodes.p, c->tris.p, c->facetVerts.p, c->facetNormal.p, c->S, K, c->cfg.seed, sx, sy, sz,
                           c->cfg.rayEpsilon, periodic(c), lit);
    }
    hipLaunchKernelGGL(k_sun_fraction, dim3((c->S + 255u) / 256u), dim3(256), 0, c->stream, lit, c->S, K, frac);
    FH_HIP(hipGetLastError());
}

}  // namespace fh
