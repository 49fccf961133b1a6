// abi.hip -- extern "C" entry points of include/facet_hip.h.
//
// Each entry point mirrors one facetvk::FacetrtVulkan operation
// (models/histream/src/facetrt/facetrt_vulkan.{h,cpp}): same argument
// meaning, same validation and the same error classes, returned as
// fh_status codes instead of exceptions.
#include <hip/hip_runtime.h>

#include <cfloat>

#include <cmath>
#include <cstdio>
#include <memory>
#include <mutex>
#include <vector>

#include "solver.hpp"

using fh::fail;

// Fixed scratch of fh_sweep_device_noalloc: the per-block residual
// partials of one sweep, 2 MiB on each device the library runs on (held
// until the process ends).  Host code never passes `scratch` itself (its
// name is a host-side stand-in): fh_create looks up the device address once
// per device and every context keeps it in fixedPartials.
constexpr uint32_t kFixedPartials = 262144;  // float2 entries
static __device__ __align__(256) float2 scratch[kFixedPartials];

namespace {

thread_local std::string g_create_error;

std::mutex g_scratch_mu;
std::vector<float2*> g_scratch_addr;  // per device ordinal

// Device address of this device's scratch array; the device is current.
float2* fixed_scratch(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_addr.size() <= size_t(device)) g_scratch_addr.resize(size_t(device) + 1u, nullptr);
    if (!g_scratch_addr[size_t(device)]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            fail(FH_RUNTIME_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_addr[size_t(device)] = static_cast<float2*>(ptr);
    }
    return g_scratch_addr[size_t(device)];
}

template <class F>
int guarded(fh_context* c, F&& body) {
    try {
        if (c) FH_HIP(hipSetDevice(c->device));
        body();
        if (c) c->err.clear();
        return FH_OK;
    } catch (const fh::Error& e) {
        if (c) c->err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        if (c) c->err = "host allocation failed";
        return FH_RUNTIME_ERROR;
    } catch (const std::exception& e) {
        if (c) c->err = e.what();
        return FH_RUNTIME_ERROR;
    }
}

void require_geometry(const fh_context* c, const char* op) {
    if (!c->hasGeometry) fail(FH_LOGIC_ERROR, std::string("setGeometry() must be called before ") + op);
}

void require_graph(const fh_context* c, const char* op) {
    if (!c->hasGraph) fail(FH_LOGIC_ERROR, std::string("buildVisibilityGraph() must be called before ") + op);
}

// facetrt_vulkan.cpp:1160-1175
void validate_optics(const fh_context* c, const fh_surface_optics* o) {
    if (!o) fail(FH_INVALID_ARGUMENT, "Optics array must contain exactly two sides per facet");
    for (uint32_t s = 0; s < c->S; ++s) {
        const fh_surface_optics& v = o[s];
        if (!std::isfinite(v.reflectance) || !std::isfinite(v.transmittance) || !std::isfinite(v.emission) ||
            !std::isfinite(v.directIrradiance) || v.reflectance < 0.0f || v.transmittance < 0.0f ||
            v.emission < 0.0f || v.directIrradiance < 0.0f || v.reflectance + v.transmittance > 1.0001f)
            fail(FH_INVALID_ARGUMENT,
                 "Optical values must be finite, non-negative, and reflectance+transmittance<=1");
    }
}

void validate_sky(float sky) {
    if (!std::isfinite(sky) || sky < 0.0f) fail(FH_INVALID_ARGUMENT, "skyRadiosity must be finite and non-negative");
}

void validate_relax(float relaxation) {
    if (!(relaxation > 0.0f && relaxation <= 1.0f)) fail(FH_INVALID_ARGUMENT, "GPU Jacobi relaxation must be in (0,1]");
}

void validate_bands(const fh_context* c, uint32_t nb, const float* r, const float* t, const float* e, const float* D,
                    const float* sky) {
    if (nb == 0u || nb > 4096u) fail(FH_INVALID_ARGUMENT, "band count must be in [1, 4096]");
    if (!r || !t || !e || !D || !sky) fail(FH_INVALID_ARGUMENT, "band arrays must not be null");
    const size_t n = size_t(c->S) * nb;
    for (size_t i = 0; i < n; ++i) {
        if (!std::isfinite(r[i]) || !std::isfinite(t[i]) || !std::isfinite(e[i]) || !std::isfinite(D[i]) ||
            r[i] < 0.0f || t[i] < 0.0f || e[i] < 0.0f || D[i] < 0.0f || r[i] + t[i] > 1.0001f)
            fail(FH_INVALID_ARGUMENT,
                 "Optical values must be finite, non-negative, and reflectance+transmittance<=1");
    }
    for (uint32_t b = 0; b < nb; ++b) validate_sky(sky[b]);
}

template <class T>
void upload(fh_context* c, T* dst, const T* src, size_t n) {
    if (n) FH_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
}

template <class T>
void download(fh_context* c, T* dst, const T* src, size_t n) {
    if (n) FH_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, c->stream));
    FH_HIP(hipStreamSynchronize(c->stream));
}

// CSR entry arrays with kCsrPad zero entries past nnz (aligned group loads)
void alloc_csr(fh_context* c, uint32_t nnz) {
    const size_t n = size_t(nnz) + fh::kCsrPad;
    c->col.ensure(n);
    c->cnt.ensure(n);
    c->val32.ensure(n);
    c->val16.ensure(n);
    FH_HIP(hipMemsetAsync(c->col.p + nnz, 0, fh::kCsrPad * sizeof(uint32_t), c->stream));
    FH_HIP(hipMemsetAsync(c->cnt.p + nnz, 0, fh::kCsrPad * sizeof(uint32_t), c->stream));
    FH_HIP(hipMemsetAsync(c->val32.p + nnz, 0, fh::kCsrPad * sizeof(float), c->stream));
    FH_HIP(hipMemsetAsync(c->val16.p + nnz, 0, fh::kCsrPad * sizeof(uint16_t), c->stream));
}

// Shared tail of fh_solve / fh_solve_accelerated: optics AoS -> r,t,e,D;
// fold; Jacobi; download.
void solve_single(fh_context* c, const fh_surface_optics* optics, float sky, const float* init,
                  fh::JacobiOptions opt, float* radiosity, fh_solve_stats* stats) {
    const uint32_t S = c->S;
    c->fbuf.ensure(size_t(S) * 4u);
    upload(c, c->fbuf.p, reinterpret_cast<const float*>(optics), size_t(S) * 4u);
    c->rr.ensure(S);
    c->tt.ensure(S);
    c->cc.ensure(S);
    c->DD.ensure(S);
    c->HH.ensure(S);  // emission staging
    fh::launch_split_optics(c, reinterpret_cast<const float4*>(c->fbuf.p), c->rr.p, c->tt.p, c->HH.p, c->DD.p);
    fh::set_sky(c, &sky, 1u, 1u);
    fh::launch_fold(c, c->HH.p, c->rr.p, c->tt.p, c->DD.p, 1u, c->cc.p);
    c->B0.ensure(S);
    c->B1.ensure(S);
    if (init) upload(c, c->B0.p, init, S);
    else fh::launch_fill_f32(c, c->B0.p, S, 0.0f);
    float* bufs[2] = {c->B0.p, c->B1.p};
    opt.ldb = 1u;
    const int fin = fh::run_jacobi(c, opt, c->cc.p, c->rr.p, c->tt.p, bufs, stats);
    if (radiosity) download(c, radiosity, bufs[fin], S);
}

void solve_bands_impl(fh_context* c, uint32_t nb, const float* dr, const float* dt, const float* de, const float* dD,
                      const float* sky, const float* dinit, uint32_t maxIter, float tolRel, float relax,
                      uint32_t precision, float* dout, fh_solve_stats* stats) {
    const uint32_t S = c->S;
    const uint32_t ldb = fh::band_stride(nb);
    const size_t n = size_t(S) * ldb;
    c->rr.ensure(n);
    c->tt.ensure(n);
    c->cc.ensure(n);
    c->DD.ensure(n);
    c->HH.ensure(n);
    c->B0.ensure(n);
    c->B1.ensure(n);
    fh::launch_pad(c, dr, nb, ldb, c->rr.p);
    fh::launch_pad(c, dt, nb, ldb, c->tt.p);
    fh::launch_pad(c, de, nb, ldb, c->HH.p);
    fh::launch_pad(c, dD, nb, ldb, c->DD.p);
    fh::set_sky(c, sky, nb, ldb);
    fh::launch_fold(c, c->HH.p, c->rr.p, c->tt.p, c->DD.p, ldb, c->cc.p);
    if (dinit) fh::launch_pad(c, dinit, nb, ldb, c->B0.p);
    else fh::launch_fill_f32(c, c->B0.p, n, 0.0f);
    float* bufs[2] = {c->B0.p, c->B1.p};
    fh::JacobiOptions opt;
    opt.ldb = ldb;
    opt.precision = precision;
    opt.relax = relax;
    opt.maxIter = maxIter;
    opt.tolRel = tolRel;
    const int fin = fh::run_jacobi(c, opt, c->cc.p, c->rr.p, c->tt.p, bufs, stats);
    fh::launch_unpad(c, bufs[fin], nb, ldb, dout);
    FH_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace

void fh::alloc_csr_device(fh_context* c, uint32_t nnz) { alloc_csr(c, nnz); }

extern "C" {

const char* fh_status_string(int status) {
    switch (status) {
        case FH_OK: return "ok";
        case FH_INVALID_ARGUMENT: return "invalid argument";
        case FH_LOGIC_ERROR: return "logic error";
        case FH_RUNTIME_ERROR: return "runtime error";
        case FH_OVERFLOW_ERROR: return "overflow error";
        default: return "unknown status";
    }
}

const char* fh_last_error(const fh_context* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int fh_create(const fh_config* config, fh_context** out) {
    g_create_error.clear();
    if (!out) return FH_INVALID_ARGUMENT;
    *out = nullptr;
    try {
        if (!config) fail(FH_INVALID_ARGUMENT, "config must not be null");
        if (config->raysPerSide == 0u || config->raysPerSide > 2048u)
            fail(FH_INVALID_ARGUMENT, "raysPerSide must be in [1, 2048]");
        if (config->sunRaysPerSide == 0u || config->sunRaysPerSide > 65536u)
            fail(FH_INVALID_ARGUMENT, "sunRaysPerSide must be in [1, 65536]");
        if (!std::isfinite(config->rayEpsilon) || config->rayEpsilon < 0.0f)
            fail(FH_INVALID_ARGUMENT, "rayEpsilon must be finite and non-negative");
        if (config->periodicNeighborCount > 20u)
            fail(FH_INVALID_ARGUMENT, "periodicNeighborCount must be in [0, 20]");
        if (config->periodicNeighborCount &&
            !(std::isfinite(config->periodicSizeX) && std::isfinite(config->periodicSizeZ) &&
              config->periodicSizeX > 0.0f && config->periodicSizeZ > 0.0f))
            fail(FH_INVALID_ARGUMENT, "periodic tile sizes must be finite and positive");
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) fail(FH_RUNTIME_ERROR, "no HIP device available");
        if (int(config->gpuIndex) >= count) fail(FH_RUNTIME_ERROR, "gpuIndex out of range");
        auto c = std::make_unique<fh_context>();
        c->cfg = *config;
        c->device = int(config->gpuIndex);
        FH_HIP(hipSetDevice(c->device));
        c->fixedPartials = fixed_scratch(c->device);
        FH_HIP(hipStreamCreate(&c->stream));
        FH_HIP(hipHostMalloc(reinterpret_cast<void**>(&c->hctrl), sizeof(fh::SolveCtrl), hipHostMallocDefault));
        c->ctrl.ensure(1);
        *out = c.release();
        return FH_OK;
    } catch (const fh::Error& e) {
        g_create_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        return FH_RUNTIME_ERROR;
    }
}

void fh_destroy(fh_context* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->pairStream) (void)hipStreamSynchronize(c->pairStream);
    if (c->hctrl) (void)hipHostFree(c->hctrl);
    if (c->pairFork) (void)hipEventDestroy(c->pairFork);
    if (c->pairJoin) (void)hipEventDestroy(c->pairJoin);
    if (c->pairStream) (void)hipStreamDestroy(c->pairStream);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

uint32_t fh_facet_count(const fh_context* c) { return c ? c->F : 0u; }
uint32_t fh_surface_count(const fh_context* c) { return c ? c->S : 0u; }
uint32_t fh_band_stride(uint32_t nb) { return fh::band_stride(nb); }
void* fh_stream(const fh_context* c) { return c ? static_cast<void*>(c->stream) : nullptr; }
uint32_t fh_directed_edge_count(const fh_context* c) { return c && c->hasGraph ? c->nnz : 0u; }

// setGeometry / validateGeometry (facetrt_vulkan.cpp:425-471)
int fh_set_geometry(fh_context* c, const fh_vertex* v, uint32_t vertexCount) {
    if (!c) return FH_INVALID_ARGUMENT;
    return guarded(c, [&] {
        if (!v || vertexCount == 0u || vertexCount % 3u != 0u)
            fail(FH_INVALID_ARGUMENT, "Geometry must be a non-empty triangle list");
        if (vertexCount / 3u > 0x7fffffffu / 2u) fail(FH_OVERFLOW_ERROR, "too many facets");
        for (uint32_t i = 0; i < vertexCount; ++i) {
            for (int k = 0; k < 3; ++k)
                if (!std::isfinite(v[i].position[k])) fail(FH_INVALID_ARGUMENT, "Geometry contains a non-finite position");
            const float* n = v[i].normal;
            const float len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            if (!(len > 0.99f && len < 1.01f)) fail(FH_INVALID_ARGUMENT, "Facet normals must be normalized");
        }
        c->hasGeometry = false;
        c->hasGraph = false;
        fh::eb_geometry_changed(c);
        const uint32_t F = vertexCount / 3u;
        std::vector<float> tri9(size_t(F) * 9u);
        std::vector<float4> nrm(F);
        for (uint32_t f = 0; f < F; ++f) {
            for (int k = 0; k < 3; ++k)
                for (int q = 0; q < 3; ++q) tri9[size_t(f) * 9u + 3u * k + q] = v[3u * f + k].position[q];
            nrm[f] = make_float4(v[3u * f].normal[0], v[3u * f].normal[1], v[3u * f].normal[2], 0.0f);
        }
        fh::BvhHost bvh;
        fh::build_bvh(tri9.data(), F, bvh);
        // sweep schedule: facets in the leaf order of the BVH (spatially
        // compact runs of facets), so the rows a sweep processes together see
        // overlapping columns; Morton or Hilbert order of the centroids by
        // FH_SCHED_KIND = 0 / 1 (measured: BVH order 1% faster than Morton,
        // Hilbert 0.5%, on the terrain forest)
        {
            float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
            std::vector<float> cen(size_t(F) * 3u);
            for (uint32_t f = 0; f < F; ++f)
                for (int q = 0; q < 3; ++q) {
                    const float x = (tri9[size_t(f) * 9u + q] + tri9[size_t(f) * 9u + 3u + q] +
                                     tri9[size_t(f) * 9u + 6u + q]) / 3.0f;
                    cen[size_t(f) * 3u + q] = x;
                    lo[q] = std::min(lo[q], x);
                    hi[q] = std::max(hi[q], x);
                }
            const float ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1.0e-20f});
            // FH_SCHED_KIND: 0 Morton, 1 Hilbert, 2 BVH leaf order (default; tuning knob)
            static const int kind = [] {
                const char* e = std::getenv("FH_SCHED_KIND");
                return e ? std::atoi(e) : 2;
            }();
            std::vector<std::pair<uint64_t, uint32_t>> key(F);
            for (uint32_t f = 0; f < F; ++f) {
                uint64_t code = 0;
                uint32_t cell[3];
                for (int q = 0; q < 3; ++q)
                    cell[q] = uint32_t(std::min(1023.0f, std::max(0.0f, (cen[size_t(f) * 3u + q] - lo[q]) / ext *
                                                                             1023.0f)));
                if (kind == 1) {
                    // Hilbert index (Skilling's transpose form, 10 bits per axis)
                    uint32_t X[3] = {cell[0], cell[1], cell[2]};
                    for (uint32_t Q = 1u << 9; Q > 1; Q >>= 1) {
                        const uint32_t P = Q - 1;
                        for (int i = 0; i < 3; ++i) {
                            if (X[i] & Q) {
                                X[0] ^= P;
                            } else {
                                const uint32_t t = (X[0] ^ X[i]) & P;
                                X[0] ^= t;
                                X[i] ^= t;
                            }
                        }
                    }
                    for (int i = 1; i < 3; ++i) X[i] ^= X[i - 1];
                    uint32_t t = 0;
                    for (uint32_t Q = 1u << 9; Q > 1; Q >>= 1)
                        if (X[2] & Q) t ^= Q - 1;
                    for (int i = 0; i < 3; ++i) X[i] ^= t;
                    for (int b = 9; b >= 0; --b)
                        for (int q = 0; q < 3; ++q) code = (code << 1) | ((X[q] >> b) & 1u);
                } else {
                    for (int b = 0; b < 10; ++b)
                        for (int q = 0; q < 3; ++q) code |= uint64_t((cell[q] >> b) & 1u) << (3 * b + q);
                }
                key[f] = {code, f};
            }
            std::stable_sort(key.begin(), key.end(),
                             [](const auto& a, const auto& b) { return a.first < b.first; });
            std::vector<uint32_t> order(F);
            for (uint32_t i = 0; i < F; ++i) order[i] = key[i].second;
            if (kind == 2 && bvh.primFacet.size() == F)
                for (uint32_t i = 0; i < F; ++i) order[i] = bvh.primFacet[i];
            c->sched.ensure(F);
            upload(c, c->sched.p, order.data(), order.size());
        }
        c->F = F;
        c->S = 2u * F;
        // scene bounds of the raster projection (vk:456-464)
        for (int q = 0; q < 3; ++q) {
            c->boundsMin[q] = FLT_MAX;
            c->boundsMax[q] = -FLT_MAX;
        }
        for (uint32_t i = 0; i < vertexCount; ++i)
            for (int q = 0; q < 3; ++q) {
                c->boundsMin[q] = std::min(c->boundsMin[q], v[i].position[q]);
                c->boundsMax[q] = std::max(c->boundsMax[q], v[i].position[q]);
            }
        c->facetVerts.ensure(tri9.size());
        c->facetNormal.ensure(F);
        c->nodes.ensure(bvh.nodes.size());
        c->tris.ensure(bvh.tris.size());
        upload(c, c->facetVerts.p, tri9.data(), tri9.size());
        upload(c, c->facetNormal.p, nrm.data(), nrm.size());
        upload(c, c->nodes.p, bvh.nodes.data(), bvh.nodes.size());
        upload(c, c->tris.p, bvh.tris.data(), bvh.tris.size());
        FH_HIP(hipStreamSynchronize(c->stream));
        c->bvhDepth = bvh.depth;
        c->bvhNodes = uint32_t(bvh.nodes.size());
        c->hasGeometry = true;
    });
}

// buildVisibilityGraph (vkh:80) as ray-cast form factors
int fh_build_form_factors(fh_context* c, fh_graph_diagnostics* out) {
    if (!c) return FH_INVALID_ARGUMENT;
    return guarded(c, [&] {
        require_geometry(c, "buildVisibilityGraph()");
        c->hasGraph = false;
        fh::eb_graph_changed(c);
        const uint32_t S = c->S, R = c->cfg.raysPerSide;
        c->hits.ensure(size_t(S) * R);
        c->cnt16.ensure(size_t(S) * R);
        c->rowlen.ensure(S);
        c->sky.ensure(S);
        c->den.ensure(S);
        c->rowptr.ensure(size_t(S) + 1u);
        fh::launch_trace_graph(c, R, c->cfg.seed, c->cfg.rayEpsilon);
        fh::launch_rows(c, R);
        fh::launch_scan(c, c->rowlen.p, S, c->rowptr.p);
        uint32_t nnz = 0;
        download(c, &nnz, c->rowptr.p + S, 1);
        c->nnz = nnz;
        alloc_csr(c, nnz);
        c->fsky.ensure(S);
        fh::launch_compact(c, R);
        c->misc.ensure(4);
        fh::launch_fill_u32(c, c->misc.p, 4, 0u);
        // a count never exceeds its row's den, which never exceeds the rays of a side
        fh::launch_weights(c, c->misc.p, R, R <= 65535u);
        fh::build_pairs(c);
        uint32_t bits = 0;
        download(c, &bits, c->misc.p, 1);
        std::memcpy(&c->closure, &bits, sizeof(float));
        c->hits.release();
        c->cnt16.release();
        c->hasGraph = true;
        if (out) {
            out->facetCount = c->F;
            out->surfaceCount = S;
            out->directedEdgeCount = nnz;
            out->fragmentOverflow = 0u;
            out->hashOverflow = 0u;
            out->maxClosureError = c->closure;
        }
    });
}

int fh_set_graph(fh_context* c, uint32_t nnz, const uint32_t* rowOffsets, const uint32_t* neighbors,
                 const uint32_t* counts, const uint32_t* denominators, const uint32_t* skyCounts,
                 fh_graph_diagnostics* out) {
    if (!c) return FH_INVALID_ARGUMENT;
    return guarded(c, [&] {
        require_geometry(c, "setGraph()");
        const uint32_t S = c->S;
        if (!rowOffsets || !denominators || !skyCounts || (nnz && (!neighbors || !counts)))
            fail(FH_INVALID_ARGUMENT, "graph arrays must not be null");
        if (rowOffsets[0] != 0u || rowOffsets[S] != nnz) fail(FH_INVALID_ARGUMENT, "rowOffsets must run from 0 to nnz");
        for (uint32_t s = 0; s < S; ++s)
            if (rowOffsets[s + 1u] < rowOffsets[s]) fail(FH_INVALID_ARGUMENT, "rowOffsets must be non-decreasing");
        for (uint32_t e = 0; e < nnz; ++e)
            if (neighbors[e] >= S) fail(FH_RUNTIME_ERROR, "visibility graph contains an invalid surface index");
        bool any = false;
        for (uint32_t s = 0; s < S; ++s) any = any || denominators[s] > 0u;
        if (!any) fail(FH_RUNTIME_ERROR, "Visibility raster has no surface samples; radiation results are invalid");
        c->hasGraph = false;
        fh::eb_graph_changed(c);
        c->nnz = nnz;
        c->rowptr.ensure(size_t(S) + 1u);
        alloc_csr(c, nnz);
        c->den.ensure(S);
        c->sky.ensure(S);
        c->fsky.ensure(S);
        upload(c, c->rowptr.p, rowOffsets, size_t(S) + 1u);
        upload(c, c->col.p, neighbors, nnz);
        upload(c, c->cnt.p, counts, nnz);
        upload(c, c->den.p, denominators, S);
        upload(c, c->sky.p, skyCounts, S);
        c->misc.ensure(4);
        fh::launch_fill_u32(c, c->misc.p, 4, 0u);
        uint32_t maxCount = 0;
        bool packable = true;
        for (uint32_t s = 0; s < S; ++s) {
            packable = packable && denominators[s] <= 65535u;
            for (uint32_t e = rowOffsets[s]; e < rowOffsets[s + 1u]; ++e) {
                maxCount = std::max(maxCount, counts[e]);
                packable = packable && counts[e] <= denominators[s];
            }
        }
        fh::launch_weights(c, c->misc.p, maxCount, packable);
        fh::build_pairs(c);
        uint32_t bits = 0;
        download(c, &bits, c->misc.p, 1);
        std::memcpy(&c->closure, &bits, sizeof(float));
        c->hasGraph = true;
        if (out) {
            out->facetCount = c->F;
            out->surfaceCount = S;
            out->directedEdgeCount = nnz;
            out->fragmentOverflow = 0u;
            out->hashOverflow = 0u;
            out->maxClosureError = c->closure;
        }
    });
}

int fh_get_graph(fh_context* c, uint32_t* rowOffsets, uint32_t* neighbors, uint32_t* counts, uint32_t* denominators,
                 uint32_t* skyCounts) {
    if (!c) return FH_INVALID_ARGUMENT;
    return guarded(c, [&] {
        require_graph(c, "getGraph()");
        if (rowOffsets) download(c, rowOffsets, c->rowptr.p, size_t(c->S) + 1u);
        if (neighbors) download(c, neighbors, c->col.p, c->nnz);
        if (counts) download(c, counts, c->cnt.p, c->nnz);
        if (denominators) download(c, denominators, c->den.p, c->S);
        if (skyCounts) download(c, skyCounts, c->sky.p, c->S);
    });
}

// solve (facetrt_vulkan.cpp:1177-1247): zero start, fixed sweeps
int fh_solve(fh_context* c, const fh_surface_optics* optics, float sky, uint32_t iterations, float relaxation,
             float* radiosity, fh_solve_stats* stats) {
    if (!c) return FH_INVALID_ARGUMENT;
    return guarded(c, [&] {
        require_graph(c, "solve()");
        validate_optics(c, optics);
        validate_sky(sky);
        validate_relax(relaxation);
        fh::JacobiOptions opt;
        opt.relax = relaxation;
        opt.maxIter = iterations ? iterations : 1u;
        solve_single(c, optics, sky, nullptr, opt, radiosity, stats);
    });
}

// solveAccelerated (facetrt_vulkan.cpp:1249-1358)
int fh_solve_accelerated(fh_context* c, const fh_surface_optics* optics, float sky, const float* init,
                         uint32_t maximumIterations, uint32_t batchIterations, float tolerance, float relaxation,
                         float* radiosity, fh_solve_stats* stats) {
    if (!c) return FH_INVALID_ARGUMENT;
    return guarded(c, [&] {
        require_graph(c, "solveAccelerated()");
        validate_optics(c, optics);
        validate_sky(sky);
        validate_relax(relaxation);
        if (maximumIterations == 0u || batchIterations == 0u || !std::isfinite(tolerance) || tolerance < 0.0f)
            fail(FH_INVALID_ARGUMENT, "Accelerated solve iterations must be positive and tolerance non-negative");
        if (init)
            for (uint32_t s = 0; s < c->S; ++s)
                if (!std::isfinite(init[s]) || init[s] < 0.0f)
                    fail(FH_INVALID_ARGUMENT, "Initial radiosity must contain one finite non-negative value per surface");
        fh::JacobiOptions opt;
        opt.relax = relaxation;
        opt.maxIter = maximumIterations;
        opt.tolAbs = tolerance;
        opt.checkEvery = batchIterations;
        solve_single(c, optics, sky, init, opt, radiosity, stats);
    });
}

// incidentIrradiance (facetrt_vulkan.cpp:1360-1399), computed on the device
int fh_incident_irradiance(fh_context* c, const float* radiosity, const fh_surface_optics* optics, float sky,
                           float* irradiance) {
    if (!c) return FH_INVALID_ARGUMENT;
    return guarded(c, [&] {
        require_graph(c, "incidentIrradiance()");
        if (!radiosity || !optics || !irradiance)
            fail(FH_INVALID_ARGUMENT, "Incident irradiance requires a completed solve and visibility graph");
        validate_sky(sky);
        const uint32_t S = c->S;
        c->fbuf.ensure(size_t(S) * 4u);
        c->rr.ensure(S);
        c->tt.ensure(S);
        c->DD.ensure(S);
        c->HH.ensure(S);
        c->B0.ensure(S);
        upload(c, c->fbuf.p, reinterpret_cast<const float*>(optics), size_t(S) * 4u);
        fh::launch_split_optics(c, reinterpret_cast<const float4*>(c->fbuf.p), c->rr.p, c->tt.p, c->HH.p, c->DD.p);
        upload(c, c->B0.p, radiosity, S);
        fh::set_sky(c, &sky, 1u, 1u);
        fh::run_incident(c, 1u, c->B0.p, c->DD.p, c->HH.p);
        download(c, irradiance, c->HH.p, S);
    });
}

// computeSunlitFraction (facetrt_vulkan.cpp:1064-1158) by sun rays
int fh_compute_sunlit_fraction(fh_context* c, fh_direction sun, float* fraction) {
    if (!c) return FH_INVALID_ARGUMENT;
    return guarded(c, [&] {
        require_geometry(c, "computeSunlitFraction()");
        if (!fraction) fail(FH_INVALID_ARGUMENT, "fraction must not be null");
        if (!std::isfinite(sun.x) || !std::isfinite(sun.y) || !std::isfinite(sun.z) ||
            (sun.x == 0.0f && sun.y == 0.0f && sun.z == 0.0f))
            fail(FH_INVALID_ARGUMENT, "sun direction must be finite and non-zero");
        const uint32_t S = c->S;
        c->misc.ensure(S);
        c->HH.ensure(S);
        fh::launch_fill_u32(c, c->misc.p, S, 0u);
        fh::launch_sun(c, sun.x, sun.y, sun.z, c->misc.p, c->HH.p);
        download(c, fraction, c->HH.p, S);
    });
}

// buildVisibilityGraph(directions) (vk:890-1061) by the reference's own
// method: an orthographic A-buffer per direction, exchanges between
// consecutive surfaces of each pixel's depth-sorted fragments
int fh_build_raster_graph(fh_context* c, const fh_direction* directions, uint32_t directionCount,
                          uint32_t rasterWidth, uint32_t rasterHeight, uint32_t maxFragmentsPerPixel,
                          fh_graph_diagnostics* out) {
    if (!c) return FH_INVALID_ARGUMENT;
    return guarded(c, [&] {
        require_geometry(c, "buildVisibilityGraph()");
        if (!directions || directionCount == 0u) fail(FH_INVALID_ARGUMENT, "Visibility directions must not be empty");
        c->hasGraph = false;
        fh::eb_graph_changed(c);
        std::vector<float> dirs(size_t(directionCount) * 3u);
        for (uint32_t d = 0; d < directionCount; ++d) {
            dirs[3u * d] = directions[d].x;
            dirs[3u * d + 1u] = directions[d].y;
            dirs[3u * d + 2u] = directions[d].z;
        }
        const uint32_t overflow =
            fh::launch_raster_graph(c, dirs.data(), directionCount, rasterWidth, rasterHeight, maxFragmentsPerPixel);
        if (out) {
            out->facetCount = c->F;
            out->surfaceCount = c->S;
            out->directedEdgeCount = c->nnz;
            out->fragmentOverflow = overflow;
            out->hashOverflow = 0u;
            out->maxClosureError = 0.0f;
        }
        if (overflow) fail(FH_RUNTIME_ERROR, "Visibility A-buffer overflow; increase Config::maxFragmentsPerPixel");
        const uint32_t S = c->S;
        std::vector<uint32_t> den(S);
        download(c, den.data(), c->den.p, S);
        bool any = false;
        for (uint32_t s = 0; s < S; ++s) any = any || den[s] > 0u;
        if (!any) fail(FH_RUNTIME_ERROR, "Visibility raster has no surface samples; radiation results are invalid");
        c->fsky.ensure(S);
        c->misc.ensure(4);
        fh::launch_fill_u32(c, c->misc.p, 4, 0u);
        // counts never exceed their row's den (each exchange adds to both
        // sides' den); packable while every den fits 16 bits
        uint32_t maxDen = 0;
        for (uint32_t s = 0; s < S; ++s) maxDen = std::max(maxDen, den[s]);
        fh::launch_weights(c, c->misc.p, maxDen, maxDen <= 65535u);
        fh::build_pairs(c);
        uint32_t bits = 0;
        download(c, &bits, c->misc.p, 1);
        std::memcpy(&c->closure, &bits, sizeof(float));
        c->hasGraph = true;
        if (out) out->maxClosureError = c->closure;
    });
}

// computeSunlitFraction(direction) (vk:1064-1157) by the reference's raster
// method: per side, the pixels where it is the nearest surface over the
// pixels it covers
int fh_compute_raster_sunlit_fraction(fh_context* c, fh_direction sun, uint32_t rasterWidth, uint32_t rasterHeight,
                                      uint32_t maxFragmentsPerPixel, float* fraction) {
    if (!c) return FH_INVALID_ARGUMENT;
    return guarded(c, [&] {
        require_geometry(c, "computeSunlitFraction()");
        if (!fraction) fail(FH_INVALID_ARGUMENT, "fraction must not be null");
        const float dir[3] = {sun.x, sun.y, sun.z};
        c->HH.ensure(c->S);
        const uint32_t overflow =
            fh::launch_raster_sunlit(c, dir, rasterWidth, rasterHeight, maxFragmentsPerPixel, c->HH.p);
        if (overflow) fail(FH_RUNTIME_ERROR, "Direct-light A-buffer overflow; increase Config::maxFragmentsPerPixel");
        download(c, fraction, c->HH.p, c->S);
    });
}

int fh_solve_bands(fh_context* c, uint32_t nb, const float* r, const float* t, const float* e, const float* D,
                   const float* sky, const float* init, uint32_t maxIter, float tolRel, float relaxation,
                   uint32_t precision, float* radiosity, fh_solve_stats* stats) {
    if (!c) return FH_INVALID_ARGUMENT;
    return guarded(c, [&] {
        require_graph(c, "solveBands()");
        validate_bands(c, nb, r, t, e, D, sky);
        validate_relax(relaxation);
        if (precision > FH_PRECISION_BF16_F) fail(FH_INVALID_ARGUMENT, "unknown precision");
        if (!radiosity) fail(FH_INVALID_ARGUMENT, "radiosity must not be null");
        const size_t n = size_t(c->S) * nb;
        fh::DBuf<float> in;
        in.ensure(n * (init ? 6u : 5u));
        float* dr = in.p;
        float* dt = dr + n;
        float* de = dt + n;
        float* dD = de + n;
        float* dout = dD + n;
        float* dinit = init ? dout + n : nullptr;
        upload(c, dr, r, n);
        upload(c, dt, t, n);
        upload(c, de, e, n);
        upload(c, dD, D, n);
        if (init) upload(c, dinit, init, n);
        solve_bands_impl(c, nb, dr, dt, de, dD, sky, dinit, maxIter ? maxIter : 1u, tolRel, relaxation, precision,
                         dout, stats);
        download(c, radiosity, dout, n);
    });
}

int fh_solve_bands_device(fh_context* c, uint32_t nb, const float* dr, const float* dt, const float* de,
                          const float* dD, const float* sky, const float* dinit, uint32_t maxIter, float tolRel,
                          float relaxation, uint32_t precision, float* dout, fh_solve_stats* stats) {
    if (!c) return FH_INVALID_ARGUMENT;
    return guarded(c, [&] {
        require_graph(c, "solveBands()");
        if (nb == 0u || nb > 4096u) fail(FH_INVALID_ARGUMENT, "band count must be in [1, 4096]");
        if (!dr || !dt || !de || !dD || !sky || !dout) fail(FH_INVALID_ARGUMENT, "band arrays must not be null");
        for (uint32_t b = 0; b < nb; ++b) validate_sky(sky[b]);
        validate_relax(relaxation);
        if (precision > FH_PRECISION_BF16_F) fail(FH_INVALID_ARGUMENT, "unknown precision");
        solve_bands_impl(c, nb, dr, dt, de, dD, sky, dinit, maxIter ? maxIter : 1u, tolRel, relaxation, precision,
                         dout, stats);
    });
}

}  // extern "C"

namespace {

// validates the optics inputs and builds them on the device into dr, dt,
// de, dD ([S][nb] each); the host inputs travel in one staging buffer
void build_spectral(fh_context* c, const fh_spectral_optics* o, fh::DBuf<float>& stage, float* dr, float* dt,
                    float* de, float* dD) {
    if (!o) fail(FH_INVALID_ARGUMENT, "optics must not be null");
    const uint32_t nb = o->bands, nm = o->materialCount, F = c->F, S = c->S;
    if (o->mode > FH_OPTICS_STATE) fail(FH_INVALID_ARGUMENT, "unknown optics mode");
    if (nb == 0u || nb > 4096u) fail(FH_INVALID_ARGUMENT, "band count must be in [1, 4096]");
    if (nm == 0u || !o->facetMaterial || !o->reflectance || !o->transmittance || !o->wavelength || !o->sunlit ||
        !o->directNormal)
        fail(FH_INVALID_ARGUMENT, "materials, bands, sunlit fractions and direct irradiance must be given");
    if (o->leafFacetCount > F) fail(FH_INVALID_ARGUMENT, "leafFacetCount exceeds the facet count");
    const bool state = o->mode == FH_OPTICS_STATE;
    if (state ? !o->temperature : (!o->sunlitTemperature || !o->shadedTemperature))
        fail(FH_INVALID_ARGUMENT, "temperatures must be given");
    const float sunLen2 = o->sun.x * o->sun.x + o->sun.y * o->sun.y + o->sun.z * o->sun.z;
    if (!std::isfinite(sunLen2) || (o->backgroundAngularStrength > 0.0f && !(sunLen2 > 0.0f)))
        fail(FH_INVALID_ARGUMENT, "sun direction must be finite (and non-zero with a Hapke ground)");
    if (!std::isfinite(o->backgroundAngularStrength)) fail(FH_INVALID_ARGUMENT, "angular strength must be finite");
    for (uint32_t f = 0; f < F; ++f)
        if (o->facetMaterial[f] >= nm) fail(FH_INVALID_ARGUMENT, "facet material index out of range");
    const auto finite = [](const float* v, size_t n, bool nonneg) {
        for (size_t i = 0; i < n; ++i)
            if (!std::isfinite(v[i]) || (nonneg && v[i] < 0.0f)) return false;
        return true;
    };
    if (!finite(o->reflectance, size_t(nm) * nb, false) || !finite(o->transmittance, size_t(nm) * nb, false) ||
        !finite(o->wavelength, nb, false) || !finite(o->directNormal, nb, true) || !finite(o->sunlit, S, false) ||
        (o->reflectanceFactor && !finite(o->reflectanceFactor, F, true)) ||
        (state && !finite(o->temperature, S, false)) ||
        (!state && (!finite(o->sunlitTemperature, nm, false) || !finite(o->shadedTemperature, nm, false))))
        fail(FH_INVALID_ARGUMENT, "optics inputs must be finite (direct irradiance and factors non-negative)");
    // staging layout (floats): mat F | refl nm nb | trans nm nb | wl nb | dni nb | factor F | sunlit S | T S |
    // Tsun nm | Tshade nm
    const size_t total = size_t(F) * 2 + size_t(nm) * nb * 2 + size_t(nb) * 2 + size_t(S) * 2 + size_t(nm) * 2;
    stage.ensure(total);
    float* p = stage.p;
    fh::SpectralOpticsArgs a{};
    a.mode = o->mode;
    a.nb = nb;
    a.S = S;
    a.leafCount = o->leafFacetCount;
    const auto put = [&](const void* src, size_t count) {
        float* at = p;
        if (src) upload(c, reinterpret_cast<uint32_t*>(p), static_cast<const uint32_t*>(src), count);
        p += count;
        return at;
    };
    a.mat = reinterpret_cast<const uint32_t*>(put(o->facetMaterial, F));
    a.reflectance = put(o->reflectance, size_t(nm) * nb);
    a.transmittance = put(o->transmittance, size_t(nm) * nb);
    a.wavelength = put(o->wavelength, nb);
    a.dni = put(o->directNormal, nb);
    a.factor = o->reflectanceFactor ? put(o->reflectanceFactor, F) : (p += F, nullptr);
    a.sunlit = put(o->sunlit, S);
    a.T = state ? put(o->temperature, S) : (p += S, nullptr);
    a.Tsun = state ? (p += nm, nullptr) : put(o->sunlitTemperature, nm);
    a.Tshade = state ? (p += nm, nullptr) : put(o->shadedTemperature, nm);
    a.nrm = c->facetNormal.p;
    a.strength = std::min(std::max(o->backgroundAngularStrength, 0.0f), 1.0f);
    a.sx = o->sun.x;
    a.sy = o->sun.y;
    a.sz = o->sun.z;
    a.r = dr;
    a.t = dt;
    a.e = de;
    a.D = dD;
    fh::launch_spectral_optics(c, a);
}

}  // namespace

extern "C" {

int fh_spectral_optics_build(fh_context* c, const fh_spectral_optics* o, float* r, float* t, float* e, float* D) {
    if (!c) return FH_INVALID_ARGUMENT;
    return guarded(c, [&] {
        require_geometry(c, "spectral optics");
        const size_t n = size_t(c->S) * (o ? o->bands : 0u);
        fh::DBuf<float> out, stage;
        out.ensure(4 * (n ? n : 1));
        build_spectral(c, o, stage, out.p, out.p + n, out.p + 2 * n, out.p + 3 * n);
        float* host[4] = {r, t, e, D};
        for (int k = 0; k < 4; ++k)
            if (host[k]) download(c, host[k], out.p + size_t(k) * n, n);
        FH_HIP(hipStreamSynchronize(c->stream));
    });
}

int fh_solve_spectral(fh_context* c, const fh_spectral_optics* o, const float* sky, const float* init,
                      uint32_t maxIter, float tolRel, float relaxation, uint32_t precision, float* radiosity,
                      fh_solve_stats* stats) {
    if (!c) return FH_INVALID_ARGUMENT;
    return guarded(c, [&] {
        require_graph(c, "solveBands()");
        if (!o || !sky || !radiosity) fail(FH_INVALID_ARGUMENT, "optics, sky and radiosity must not be null");
        const uint32_t nb = o->bands;
        if (nb == 0u || nb > 4096u) fail(FH_INVALID_ARGUMENT, "band count must be in [1, 4096]");
        for (uint32_t b = 0; b < nb; ++b) validate_sky(sky[b]);
        validate_relax(relaxation);
        if (precision > FH_PRECISION_BF16_F) fail(FH_INVALID_ARGUMENT, "unknown precision");
        const size_t n = size_t(c->S) * nb;
        fh::DBuf<float> in, stage;
        in.ensure(n * (init ? 6u : 5u));
        float* dout = in.p + 4 * n;
        float* dinit = init ? dout + n : nullptr;
        build_spectral(c, o, stage, in.p, in.p + n, in.p + 2 * n, in.p + 3 * n);
        if (init) upload(c, dinit, init, n);
        solve_bands_impl(c, nb, in.p, in.p + n, in.p + 2 * n, in.p + 3 * n, sky, dinit, maxIter ? maxIter : 1u,
                         tolRel, relaxation, precision, dout, stats);
        download(c, radiosity, dout, n);
    });
}

int fh_incident_bands(fh_context* c, uint32_t nb, const float* radiosity, const float* direct, const float* sky,
                      float* irradiance) {
    if (!c) return FH_INVALID_ARGUMENT;
    return guarded(c, [&] {
        require_graph(c, "incidentIrradiance()");
        if (nb == 0u || nb > 4096u) fail(FH_INVALID_ARGUMENT, "band count must be in [1, 4096]");
        if (!radiosity || !direct || !sky || !irradiance) fail(FH_INVALID_ARGUMENT, "arrays must not be null");
        for (uint32_t b = 0; b < nb; ++b) validate_sky(sky[b]);
        const uint32_t ldb = fh::band_stride(nb);
        const size_t n = size_t(c->S) * nb, np = size_t(c->S) * ldb;
        fh::DBuf<float> tmp;
        tmp.ensure(2 * n + 3 * np);
        float* dB = tmp.p;
        float* dD = dB + n;
        float* pB = dD + n;
        float* pD = pB + np;
        float* pH = pD + np;
        upload(c, dB, radiosity, n);
        upload(c, dD, direct, n);
        fh::launch_pad(c, dB, nb, ldb, pB);
        fh::launch_pad(c, dD, nb, ldb, pD);
        fh::set_sky(c, sky, nb, ldb);
        fh::run_incident(c, ldb, pB, pD, pH);
        fh::launch_unpad(c, pH, nb, ldb, dB);
        download(c, irradiance, dB, n);
    });
}

int fh_sweep_device(fh_context* c, uint32_t nb, const float* dc, const float* dr, const float* dt, const float* dBin,
                    float* dBout, uint32_t precision) {
    if (!c) return FH_INVALID_ARGUMENT;
    return guarded(c, [&] {
        require_graph(c, "sweep()");
        if (nb == 0u || nb > 4096u) fail(FH_INVALID_ARGUMENT, "band count must be in [1, 4096]");
        if (!dc || !dr || !dt || !dBin || !dBout) fail(FH_INVALID_ARGUMENT, "arrays must not be null");
        if (precision > FH_PRECISION_BF16_DC) fail(FH_INVALID_ARGUMENT, "a single sweep reads FP32 or BF16 weights");
        fh::launch_single_sweep(c, fh::band_stride(nb), dc, dr, dt, dBin, dBout, precision == FH_PRECISION_BF16_DC);
    });
}

// fh_sweep_device with the block partials in the fixed scratch array: no
// device allocation in the call
int fh_sweep_device_noalloc(fh_context* c, uint32_t nb, const float* dc, const float* dr, const float* dt,
                            const float* dBin, float* dBout, uint32_t precision) {
    if (!c) return FH_INVALID_ARGUMENT;
    return guarded(c, [&] {
        require_graph(c, "sweep()");
        if (nb == 0u || nb > 4096u) fail(FH_INVALID_ARGUMENT, "band count must be in [1, 4096]");
        if (!dc || !dr || !dt || !dBin || !dBout) fail(FH_INVALID_ARGUMENT, "arrays must not be null");
        if (precision > FH_PRECISION_BF16_DC) fail(FH_INVALID_ARGUMENT, "a single sweep reads FP32 or BF16 weights");
        if (!c->fixedPartials) fail(FH_RUNTIME_ERROR, "fixed scratch is not set up");
        fh::launch_single_sweep(c, fh::band_stride(nb), dc, dr, dt, dBin, dBout, precision == FH_PRECISION_BF16_DC,
                                nullptr, nullptr, nullptr, nullptr, false, c->fixedPartials, kFixedPartials);
    });
}

int fh_benchmark_sweep(fh_context* c, uint32_t nb, uint32_t precision, uint32_t materialTables, uint32_t reps,
                       float* us) {
    if (!c) return FH_INVALID_ARGUMENT;
    return guarded(c, [&] {
        require_graph(c, "sweep()");
        if (nb == 0u || nb > 4096u || reps == 0u || !us) fail(FH_INVALID_ARGUMENT, "invalid benchmark arguments");
        const uint32_t ldb = fh::band_stride(nb);
        const size_t n = size_t(c->S) * ldb;
        // the energy balance's form of the band-parallel sweep: r, t from two
        // material tables (first half of the facets, second half), c and D
        // from per-side direct factors, Chebyshev weights != 1 (the previous
        // iterate is read)
        fh::DBuf<uint32_t> tabMat;
        fh::DBuf<float> tabF;
        fh::SweepTables tables;
        const bool useTables = materialTables && fh::uses_bf16_gather(ldb);
        if (useTables) {
            std::vector<uint32_t> mat(c->F);
            for (uint32_t f = 0; f < c->F; ++f) mat[f] = f < c->F / 2u ? 0u : 1u;
            std::vector<float> host(4u * ldb + c->S, 0.0f);
            for (uint32_t b = 0; b < nb; ++b) {
                host[b] = 0.45f;               // rtab row 0
                host[ldb + b] = 0.15f;         // rtab row 1
                host[2u * ldb + b] = 0.40f;    // ttab row 0 (row 1: 0)
                host[3u * ldb + b] = 900.0f / float(nb);  // dni
            }
            for (uint32_t s = 0; s < c->S; ++s) host[4u * ldb + s] = (s & 1u) ? 0.0f : 0.6f;  // direct factors
            tabMat.ensure(c->F);
            tabF.ensure(host.size() + 2u * ldb + ldb);
            upload(c, tabMat.p, mat.data(), mat.size());
            upload(c, tabF.p, host.data(), host.size());
            float* omega = tabF.p + host.size();
            fh::launch_fill_f32(c, omega, ldb, 1.2f);
            std::vector<float> sky(nb, 60.0f / float(nb));
            fh::set_sky(c, sky.data(), nb, ldb);
            tables.fmat = tabMat.p;
            tables.rtab = tabF.p;
            tables.ttab = tabF.p + 2u * ldb;
            tables.dniv = tabF.p + 3u * ldb;
            tables.dirf = tabF.p + 4u * ldb;
            tables.omega = omega;
        }
        const fh::SweepTables* tab = useTables ? &tables : nullptr;
        fh::DBuf<float> buf;
        buf.ensure(5 * n);
        float* cc = buf.p;
        float* rr = cc + n;
        float* tt = rr + n;
        float* B0 = tt + n;
        float* B1 = B0 + n;
        fh::launch_fill_f32(c, cc, n, 25.0f);
        fh::launch_fill_f32(c, rr, n, 0.2f);
        fh::launch_fill_f32(c, tt, n, 0.1f);
        fh::launch_fill_f32(c, B0, n, 50.0f);
        fh::launch_fill_f32(c, B1, n, 0.0f);
        const bool q8 = precision == FH_PRECISION_FP8_INC && fh::uses_bf16_gather(ldb);
        // FH_PRECISION_BF16_F: the BF16 phase's sweep (BF16 F, BF16 copy of B)
        const bool wf = precision == FH_PRECISION_BF16_F;
        const bool w16 = precision == FH_PRECISION_BF16_DC || wf || q8;
        // BF16 solves of band-parallel strides run increment sweeps: BF16
        // increments gathered, accumulated gathers and carries in FP32; FP8
        // solves the same with FP8 increments
        fh::DBuf<uint16_t> h16;
        fh::DBuf<uint8_t> h8;
        fh::DBuf<float> acc;
        uint16_t* H0 = nullptr;
        uint16_t* H1 = nullptr;
        uint8_t* Q0 = nullptr;
        uint8_t* Q1 = nullptr;
        float* G = nullptr;
        float* K = nullptr;
        if (q8) {
            const size_t nq = size_t(c->S) * fh::q8_stride(ldb);
            h8.ensure(2 * nq);
            Q0 = h8.p;
            Q1 = h8.p + nq;
            fh::launch_fill_u32(c, reinterpret_cast<uint32_t*>(h8.p), nq / 2u, 0u);  // FP8 zeros, exponent 0
            acc.ensure(2 * n);
            G = acc.p;
            K = acc.p + n;
            fh::launch_fill_f32(c, G, n, 100.0f);
            fh::launch_fill_f32(c, K, n, 0.0f);
        } else if (w16 && fh::uses_bf16_gather(ldb)) {
            const size_t nh = size_t(c->S) * fh::bf16_stride(ldb);
            h16.ensure(2 * nh);
            H0 = h16.p;
            H1 = h16.p + nh;
            fh::launch_fill_u32(c, reinterpret_cast<uint32_t*>(h16.p), nh, 0x3c233c23u);  // 0.01 in BF16
            if (!wf) {
                acc.ensure(2 * n);
                G = acc.p;
                K = acc.p + n;
                fh::launch_fill_f32(c, G, n, 100.0f);
                fh::launch_fill_f32(c, K, n, 0.0f);
            }
        }
        for (int i = 0; i < 3; ++i)
            fh::launch_single_sweep(c, ldb, cc, rr, tt, B0, B1, w16, H0, H1, G, K, true, nullptr, 0u, tab, Q0, Q1);
        hipEvent_t e0, e1;
        FH_HIP(hipEventCreate(&e0));
        FH_HIP(hipEventCreate(&e1));
        FH_HIP(hipEventRecord(e0, c->stream));
        for (uint32_t i = 0; i < reps; ++i)
            fh::launch_single_sweep(c, ldb, cc, rr, tt, (i & 1u) ? B1 : B0, (i & 1u) ? B0 : B1, w16,
                                    (i & 1u) ? H1 : H0, (i & 1u) ? H0 : H1, G, K, true, nullptr, 0u, tab,
                                    (i & 1u) ? Q1 : Q0, (i & 1u) ? Q0 : Q1);
        FH_HIP(hipEventRecord(e1, c->stream));
        FH_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        FH_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *us = ms * 1000.0f / float(reps);
    });
}

__global__ void k_copy_calibration(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        b[i] = a[i];
}

int fh_benchmark_copy(fh_context* c, uint64_t bytes, uint32_t reps, float* us) {
    if (!c) return FH_INVALID_ARGUMENT;
    return guarded(c, [&] {
        if (bytes < 16u || reps == 0u || !us) fail(FH_INVALID_ARGUMENT, "invalid benchmark arguments");
        const size_t n = size_t(bytes / 16u);
        fh::DBuf<float4> a, b;
        a.ensure(n);
        b.ensure(n);
        fh::launch_fill_f32(c, reinterpret_cast<float*>(a.p), n * 4u, 1.0f);
        const dim3 grid(2048), block(256);
        hipLaunchKernelGGL(k_copy_calibration, grid, block, 0, c->stream, a.p, b.p, n);
        hipEvent_t e0, e1;
        FH_HIP(hipEventCreate(&e0));
        FH_HIP(hipEventCreate(&e1));
        FH_HIP(hipEventRecord(e0, c->stream));
        for (uint32_t i = 0; i < reps; ++i)
            hipLaunchKernelGGL(k_copy_calibration, grid, block, 0, c->stream, a.p, b.p, n);
        FH_HIP(hipEventRecord(e1, c->stream));
        FH_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        FH_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *us = ms * 1000.0f / float(reps);
    });
}

int fh_synchronize(fh_context* c) {
    if (!c) return FH_INVALID_ARGUMENT;
    return guarded(c, [&] { FH_HIP(hipStreamSynchronize(c->stream)); });
}

}  // extern "C"
