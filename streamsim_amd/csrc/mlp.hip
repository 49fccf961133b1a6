// mlp.hip -- the learned warm start: a 9 -> H -> H -> 1 GELU MLP that
// predicts each energy-balance unit's converged temperature from
//   [absorbed SW, sky-view fraction, ground-view fraction, LW emissivity,
//    material class, air temperature, wind speed, previous temperature,
//    sky LW irradiance],
// trained on the device by the kernels below (forward, backward, Adam) on
// temperatures generated by the plain coupled solver.  The guess only seeds
// the fixed point; the coupled iteration then runs to the same tolerance.
//
// Units: unit u < leafCount is leaf facet u (both sides share T); otherwise
// ground side 2 leafCount + (u - leafCount).
// GELU is the tanh form 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))).
#include <hip/hip_runtime.h>

#include <cmath>
#include <random>

#include "solver.hpp"

namespace fh {

constexpr int kIn = 9;
constexpr int kMaxHidden = 256;

__device__ __forceinline__ float gelu_f(float x) {
    const float k = 0.7978845608028654f;
    return 0.5f * x * (1.0f + tanhf(k * (x + 0.044715f * x * x * x)));
}
// GELU for inference on the matrix-core path: the same tanh form with
// tanh(u) = 1 - 2 / (1 + e^(2u)) on the hardware exp2 and reciprocal
// (about 1e-6 relative; the model is a warm-start predictor)
__device__ __forceinline__ float gelu_fast(float x) {
    const float k = 0.7978845608028654f;
    const float u = k * (x + 0.044715f * x * x * x);
    const float e = __builtin_amdgcn_exp2f(fminf(2.8853900817779268f * u, 126.0f));  // e^(2u)
    const float th = 1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + e);
    return 0.5f * x * (1.0f + th);
}
__device__ __forceinline__ float gelu_d(float x) {
    const float k = 0.7978845608028654f;
    const float th = tanhf(k * (x + 0.044715f * x * x * x));
    const float du = k * (1.0f + 3.0f * 0.044715f * x * x);
    return 0.5f * (1.0f + th) + 0.5f * x * (1.0f - th * th) * du;
}

struct MlpView {
    const float *W1, *b1, *W2, *b2, *W3, *b3;
};
static MlpView view(const float* p, uint32_t H) {
    MlpView v;
    v.W1 = p;
    v.b1 = v.W1 + size_t(H) * kIn;
    v.W2 = v.b1 + H;
    v.b2 = v.W2 + size_t(H) * H;
    v.W3 = v.b2 + H;
    v.b3 = v.W3 + H;
    return v;
}

uint32_t mlp_param_count(uint32_t H) { return H * kIn + H + H * H + H + H + 1u; }

// ---- training kernels (activations stored [sample][H]) ---------------------
__global__ void k_l1(const float* __restrict__ X, uint32_t n, uint32_t H, MlpView w, float* __restrict__ A1,
                     float* __restrict__ H1) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(n) * H) return;
    const uint32_t b = uint32_t(i / H), j = uint32_t(i - size_t(b) * H);
    float a = w.b1[j];
#pragma unroll
    for (int q = 0; q < kIn; ++q) a += w.W1[size_t(j) * kIn + q] * X[size_t(b) * kIn + q];
    A1[i] = a;
    H1[i] = gelu_f(a);
}

__global__ void k_l2(const float* __restrict__ H1, uint32_t n, uint32_t H, MlpView w, float* __restrict__ A2,
                     float* __restrict__ H2) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(n) * H) return;
    const uint32_t b = uint32_t(i / H), j = uint32_t(i - size_t(b) * H);
    float a = w.b2[j];
    const float* hr = H1 + size_t(b) * H;
    const float* wr = w.W2 + size_t(j) * H;
    for (uint32_t q = 0; q < H; ++q) a += wr[q] * hr[q];
    A2[i] = a;
    H2[i] = gelu_f(a);
}

// output, loss term and dA2; one wavefront per sample
__global__ __launch_bounds__(256) void k_out_back(const float* __restrict__ H2, const float* __restrict__ A2,
                                                  const float* __restrict__ target, uint32_t n, uint32_t H, MlpView w,
                                                  float scale, float* __restrict__ y, float* __restrict__ dA2,
                                                  float* __restrict__ lossPart) {
    const uint32_t b = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    if (b >= n) return;
    float acc = 0.0f;
    for (uint32_t j = lane; j < H; j += 64u) acc += w.W3[j] * H2[size_t(b) * H + j];
    for (int m = 32; m > 0; m >>= 1) acc += __shfl_xor(acc, m, 64);
    const float out = acc + w.b3[0];
    if (lane == 0) y[b] = out;
    if (!target) return;
    const float diff = out - target[b];
    const float go = scale * 2.0f * diff;
    if (lane == 0) lossPart[b] = diff * diff;
    for (uint32_t j = lane; j < H; j += 64u) dA2[size_t(b) * H + j] = go * w.W3[j] * gelu_d(A2[size_t(b) * H + j]);
}

__global__ void k_back1(const float* __restrict__ dA2, const float* __restrict__ A1, uint32_t n, uint32_t H, MlpView w,
                        float* __restrict__ dA1) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(n) * H) return;
    const uint32_t b = uint32_t(i / H), q = uint32_t(i - size_t(b) * H);
    float a = 0.0f;
    const float* d = dA2 + size_t(b) * H;
    for (uint32_t j = 0; j < H; ++j) a += d[j] * w.W2[size_t(j) * H + q];
    dA1[i] = a * gelu_d(A1[i]);
}

// parameter gradients: one lane per parameter, summed over the batch
__global__ void k_grads(const float* __restrict__ X, const float* __restrict__ H1, const float* __restrict__ H2,
                        const float* __restrict__ dA1, const float* __restrict__ dA2, const float* __restrict__ y,
                        const float* __restrict__ target, uint32_t n, uint32_t H, float scale, float* __restrict__ g) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nW1 = H * kIn, nb1 = H, nW2 = H * H, nb2 = H, nW3 = H;
    if (p >= nW1 + nb1 + nW2 + nb2 + nW3 + 1u) return;
    float acc = 0.0f;
    if (p < nW1) {
        const uint32_t j = p / kIn, q = p - j * kIn;
        for (uint32_t b = 0; b < n; ++b) acc += dA1[size_t(b) * H + j] * X[size_t(b) * kIn + q];
    } else if (p < nW1 + nb1) {
        const uint32_t j = p - nW1;
        for (uint32_t b = 0; b < n; ++b) acc += dA1[size_t(b) * H + j];
    } else if (p < nW1 + nb1 + nW2) {
        const uint32_t k = p - nW1 - nb1, j = k / H, q = k - j * H;
        for (uint32_t b = 0; b < n; ++b) acc += dA2[size_t(b) * H + j] * H1[size_t(b) * H + q];
    } else if (p < nW1 + nb1 + nW2 + nb2) {
        const uint32_t j = p - nW1 - nb1 - nW2;
        for (uint32_t b = 0; b < n; ++b) acc += dA2[size_t(b) * H + j];
    } else if (p < nW1 + nb1 + nW2 + nb2 + nW3) {
        const uint32_t j = p - nW1 - nb1 - nW2 - nb2;
        for (uint32_t b = 0; b < n; ++b) acc += scale * 2.0f * (y[b] - target[b]) * H2[size_t(b) * H + j];
    } else {
        for (uint32_t b = 0; b < n; ++b) acc += scale * 2.0f * (y[b] - target[b]);
    }
    g[p] = acc;
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       uint32_t np, float lr, float b1, float b2, float c1, float c2, float eps) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---- inference: one lane per unit, layer-1 activations in registers --------
template <int H>
__global__ __launch_bounds__(128) void k_infer(const float* __restrict__ X, uint32_t n, const float* __restrict__ params,
                                               float* __restrict__ y) {
    __shared__ float W[H * kIn + H + H * H + H + H + 1];
    const uint32_t np = H * kIn + H + H * H + H + H + 1;
    for (uint32_t i = threadIdx.x; i < np; i += blockDim.x) W[i] = params[i];
    __syncthreads();
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const float* W1 = W;
    const float* b1 = W1 + H * kIn;
    const float* W2 = b1 + H;
    const float* b2 = W2 + H * H;
    const float* W3 = b2 + H;
    const float b3 = W3[H];
    float x[kIn];
#pragma unroll
    for (int q = 0; q < kIn; ++q) x[q] = X[size_t(b) * kIn + q];
    float h1[H];
#pragma unroll
    for (int j = 0; j < H; ++j) {
        float a = b1[j];
#pragma unroll
        for (int q = 0; q < kIn; ++q) a += W1[j * kIn + q] * x[q];
        h1[j] = gelu_f(a);
    }
    float out = b3;
    for (int j = 0; j < H; ++j) {
        float a = b2[j];
#pragma unroll
        for (int q = 0; q < H; ++q) a += W2[j * H + q] * h1[q];
        out += W3[j] * gelu_f(a);
    }
    y[b] = out;
}

// ---- inference on the matrix cores (hidden width 128) -----------------------
// W2T[q][j] = W2[j][q]: the layer-2 B operand read row-contiguously
__global__ void k_transpose_w2(const float* __restrict__ params, uint32_t H, float* __restrict__ W2T) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= H * H) return;
    const float* W2 = params + H * kIn + H;
    const uint32_t q = i / H, j = i - q * H;
    W2T[i] = W2[j * H + q];
}

typedef float f32x16 __attribute__((ext_vector_type(16)));

// A block of 256 threads (4 waves) evaluates 64 units (their inputs staged
// in the LDS by coalesced loads first; 64 units and the staged inputs
// measured 0.86 ms against 1.05 ms for 96 units read per unit, 1.5 M
// units).  Layer 1 (9 -> 128) on the vector lanes into LDS; layer 2 (64 x
// 128 times 128 x 128) on the matrix cores with v_mfma_f32_32x32x2f32 (FP32
// in, FP32 accumulate): wave w owns output columns [32 w, 32 w + 32) for the
// block's 2 row tiles of 32 units, loading each B operand once for the 2
// tiles; the epilogue applies
// b2, GELU and the output layer per column, sums the 32 columns across the
// lanes and the 4 column tiles through LDS.
// Operand layout of the 32x32x2 instruction (wave64): lane l supplies
// A[l % 32][l / 32] and B[l / 32][l % 32]; accumulator register r of lane l
// holds C[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
constexpr int kMfmaUnits = 64;
__global__ __launch_bounds__(256) void k_infer_mfma(const float* __restrict__ X, uint32_t n,
                                                    const float* __restrict__ params, const float* __restrict__ W2T,
                                                    float* __restrict__ y) {
    constexpr int H = 128, TU = kMfmaUnits, RT = TU / 32;
    __shared__ float sH1[TU][H + 1];
    __shared__ float sOut[TU][4];
    const float* W1 = params;
    const float* b1 = W1 + H * kIn;
    const float* b2 = b1 + H + H * H;
    const float* W3 = b2 + H;
    const float b3 = W3[H];
    const uint32_t u0 = blockIdx.x * TU;
    __shared__ float sX[TU * kIn];  // the block's inputs
    for (uint32_t i = threadIdx.x; i < uint32_t(TU * kIn); i += 256u)
        sX[i] = size_t(u0) * kIn + i < size_t(n) * kIn ? X[size_t(u0) * kIn + i] : 0.0f;
    __syncthreads();
    // layer 1: thread t computes hidden unit j = t % 128 of units t / 128,
    // t / 128 + 2, ...: its row of W1 stays in registers
    {
        const uint32_t j = threadIdx.x & uint32_t(H - 1);
        float w1[kIn];
#pragma unroll
        for (int q = 0; q < kIn; ++q) w1[q] = W1[j * kIn + q];
        const float b1j = b1[j];
        for (uint32_t u = threadIdx.x / uint32_t(H); u < uint32_t(TU); u += 256u / uint32_t(H)) {
            const uint32_t unit = u0 + u;
            float a = b1j;
            if (unit < n) {
#pragma unroll
                for (int q = 0; q < kIn; ++q) a += w1[q] * sX[u * kIn + q];
            }
            sH1[u][j] = unit < n ? gelu_fast(a) : 0.0f;
        }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 31, lk = lane >> 5;
    // this lane's B operands for all 64 k-steps, loaded before the matrix
    // loop so that no global load latency sits between the MFMAs
    float bv[H / 2];
#pragma unroll
    for (int k = 0; k < H / 2; ++k) bv[k] = W2T[(2 * k + lk) * H + w * 32 + li];
    __syncthreads();
    f32x16 acc[RT];
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[t][r] = 0.0f;
#pragma unroll 8
    for (int k = 0; k < H / 2; ++k) {
#pragma unroll
        for (int t = 0; t < RT; ++t)
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(sH1[t * 32 + li][2 * k + lk], bv[k], acc[t], 0, 0, 0);
    }
    const int j = w * 32 + li;
    const float bj = b2[j], wj = W3[j];
#pragma unroll
    for (int t = 0; t < RT; ++t) {
        float part[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) part[r] = wj * gelu_fast(acc[t][r] + bj);
#pragma unroll
        for (int r = 0; r < 16; ++r)
#pragma unroll
            for (int m = 16; m > 0; m >>= 1) part[r] += __shfl_xor(part[r], m, 64);
        if (li == 0)
#pragma unroll
            for (int r = 0; r < 16; ++r) sOut[t * 32 + 8 * (r >> 2) + 4 * lk + (r & 3)][w] = part[r];
    }
    __syncthreads();
    if (threadIdx.x < uint32_t(TU)) {
        const uint32_t unit = u0 + threadIdx.x;
        if (unit < n) y[unit] = b3 + sOut[threadIdx.x][0] + sOut[threadIdx.x][1] + sOut[threadIdx.x][2] +
                                sOut[threadIdx.x][3];
    }
}

static void infer(fh_context* c, const float* X, uint32_t n, float* y) {
    const uint32_t H = c->mlp.hidden;
    if (H == 128) {
        c->mlp.W2T.ensure(size_t(H) * H);
        hipLaunchKernelGGL(k_transpose_w2, dim3((H * H + 255u) / 256u), dim3(256), 0, c->stream, c->mlp.params.p, H,
                           c->mlp.W2T.p);
        const uint32_t blocks = (n + kMfmaUnits - 1u) / kMfmaUnits;
        hipLaunchKernelGGL(k_infer_mfma, dim3(blocks ? blocks : 1u), dim3(256), 0, c->stream, X, n, c->mlp.params.p,
                           c->mlp.W2T.p, y);
        FH_HIP(hipGetLastError());
        return;
    }
    const dim3 grid((n + 127u) / 128u ? (n + 127u) / 128u : 1u);
    if (H == 32) hipLaunchKernelGGL(k_infer<32>, grid, dim3(128), 0, c->stream, X, n, c->mlp.params.p, y);
    else if (H == 64) hipLaunchKernelGGL(k_infer<64>, grid, dim3(128), 0, c->stream, X, n, c->mlp.params.p, y);
    else fail(FH_INVALID_ARGUMENT, "inference supports hidden widths 32, 64 and 128");
    FH_HIP(hipGetLastError());
}

// ---- energy-balance features ------------------------------------------------
// ground view fraction per side: sum of weights of entries that hit ground
__global__ __launch_bounds__(256) void k_ground_view(const uint32_t* __restrict__ rowptr,
                                                     const uint32_t* __restrict__ col, const float* __restrict__ w,
                                                     uint32_t S, uint32_t groundFirstSide, float* __restrict__ gv) {
    const uint32_t s = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (s >= S) return;
    const uint32_t lane = threadIdx.x & 63u;
    float acc = 0.0f;
    for (uint32_t e = rowptr[s] + lane; e < rowptr[s + 1]; e += 64u)
        if (col[e] >= groundFirstSide) acc += w[e];
    for (int m = 32; m > 0; m >>= 1) acc += __shfl_xor(acc, m, 64);
    if (lane == 0) gv[s] = acc;
}

struct FeatArgs {
    float Ta, wind, sky;
    float inMean[kIn], inStd[kIn];
};

__global__ void k_features(uint32_t units, uint32_t leafCount, const uint32_t* __restrict__ mat,
                           const float* __restrict__ absorbed, const float* __restrict__ fsky,
                           const float* __restrict__ gv, const float* __restrict__ eps, const float* __restrict__ Tprev,
                           FeatArgs a, float* __restrict__ raw, float* __restrict__ norm) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= units) return;
    float f[kIn];
    if (u < leafCount) {
        const uint32_t s = 2u * u;
        f[0] = absorbed[s] + absorbed[s + 1];
        f[1] = 0.5f * (fsky[s] + fsky[s + 1]);
        f[2] = 0.5f * (gv[s] + gv[s + 1]);
        f[3] = eps[s];
        f[4] = float(mat[u]);
        f[7] = Tprev[s];
    } else {
        const uint32_t s = 2u * leafCount + (u - leafCount);
        f[0] = absorbed[s];
        f[1] = fsky[s];
        f[2] = gv[s];
        f[3] = eps[s];
        f[4] = float(mat[s >> 1]);
        f[7] = Tprev[s];
    }
    f[5] = a.Ta;
    f[6] = a.wind;
    f[8] = a.sky;
#pragma unroll
    for (int q = 0; q < kIn; ++q) {
        raw[size_t(u) * kIn + q] = f[q];
        norm[size_t(u) * kIn + q] = (f[q] - a.inMean[q]) / a.inStd[q];
    }
}

// T = Ta + outMean + outStd * y, clamped like the Newton update
__global__ void k_apply_guess(uint32_t units, uint32_t leafCount, const float* __restrict__ y, float Ta, float mean,
                              float sd, float* __restrict__ T) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= units) return;
    float t = Ta + mean + sd * y[u];
    t = fminf(fmaxf(t, 220.0f), 360.0f);
    if (u < leafCount) {
        T[2u * u] = t;
        T[2u * u + 1u] = t;
    } else {
        T[2u * leafCount + (u - leafCount)] = t;
    }
}

// Residual form: T += (Ta - TaPrev) + sd (y - yPrev), the change the network
// predicts between the previous node and this one, applied to the previous
// node's converged state (the network's own bias per unit cancels)
__global__ void k_apply_delta(uint32_t units, uint32_t leafCount, const float* __restrict__ y,
                              const float* __restrict__ yPrev, float dTa, float sd, float* __restrict__ T) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= units) return;
    const float d = dTa + sd * (y[u] - yPrev[u]);
    if (u < leafCount) {
        const float t = fminf(fmaxf(T[2u * u] + d, 220.0f), 360.0f);
        T[2u * u] = t;
        T[2u * u + 1u] = t;
    } else {
        const uint32_t s = 2u * leafCount + (u - leafCount);
        T[s] = fminf(fmaxf(T[s] + d, 220.0f), 360.0f);
    }
}

static uint32_t unit_count(const fh_context* c) { return c->eb.leafCount + 2u * (c->F - c->eb.leafCount); }

// Computes the raw + normalised features of every unit for the current node
// (after the SW step, with Tprev = the node's starting state).
void eb_features(fh_context* c, float Ta, float wind, float sky) {
    EbState& e = c->eb;
    const uint32_t units = unit_count(c);
    if (!e.gview.p || e.gview.n != c->S) {
        e.gview.ensure(c->S);
        hipLaunchKernelGGL(k_ground_view, dim3((c->S + 3u) / 4u), dim3(256), 0, c->stream, c->rowptr.p, c->col.p,
                           c->val32.p, c->S, 2u * e.leafCount, e.gview.p);
    }
    e.feat.ensure(size_t(units) * kIn);
    e.featN.ensure(size_t(units) * kIn);
    e.guess.ensure(units);
    FeatArgs a{};
    a.Ta = Ta;
    a.wind = wind;
    a.sky = sky;
    for (int q = 0; q < kIn; ++q) {
        a.inMean[q] = c->mlp.ready ? c->mlp.inMean[q] : 0.0f;
        a.inStd[q] = c->mlp.ready && c->mlp.inStd[q] > 0.0f ? c->mlp.inStd[q] : 1.0f;
    }
    hipLaunchKernelGGL(k_features, dim3((units + 255u) / 256u), dim3(256), 0, c->stream, units, e.leafCount,
                       e.material.p, e.absorbed.p, c->fsky.p, e.gview.p, e.eps.p, e.Tprev.p, a, e.feat.p, e.featN.p);
    FH_HIP(hipGetLastError());
    e.featValid = true;
}

struct EbConst;
void surrogate_guess(fh_context* c, const EbConst&, float windSpeed, float skyLongwave, float Ta) {
    if (!c->mlp.ready) fail(FH_LOGIC_ERROR, "surrogate warm start requested but no surrogate parameters are set");
    EbState& e = c->eb;
    const uint32_t units = unit_count(c);
    eb_features(c, Ta, windSpeed, skyLongwave);
    infer(c, e.featN.p, units, e.guess.p);
    const dim3 grid((units + 255u) / 256u);
    if (e.guessPrevValid) {
        // the previous node was also warm-started: predict the change
        hipLaunchKernelGGL(k_apply_delta, grid, dim3(256), 0, c->stream, units, e.leafCount, e.guess.p,
                           e.guessPrev.p, Ta - e.guessPrevTa, c->mlp.outStd, e.T.p);
    } else {
        hipLaunchKernelGGL(k_apply_guess, grid, dim3(256), 0, c->stream, units, e.leafCount, e.guess.p, Ta,
                           c->mlp.outMean, c->mlp.outStd, e.T.p);
    }
    FH_HIP(hipGetLastError());
    e.guessPrev.ensure(units);
    FH_HIP(hipMemcpyAsync(e.guessPrev.p, e.guess.p, sizeof(float) * units, hipMemcpyDeviceToDevice, c->stream));
    e.guessPrevTa = Ta;
    e.guessPrevValid = true;
}

// ---- training driver -----------------------------------------------------------
static void ensure_train(fh_context* c, uint32_t batch) {
    MlpState& m = c->mlp;
    const size_t nh = size_t(batch) * m.hidden;
    for (DBuf<float>* b : {&m.A1, &m.H1, &m.A2, &m.H2, &m.dA2, &m.dA1}) b->ensure(nh);
    m.y.ensure(batch);
    m.red.ensure(batch);
    m.grad.ensure(mlp_param_count(m.hidden));
}

// forward + backward over one batch of device samples; grad in m.grad
static void batch_grad(fh_context* c, const float* X, const float* T, uint32_t n) {
    MlpState& m = c->mlp;
    const uint32_t H = m.hidden;
    const MlpView w = view(m.params.p, H);
    const size_t nh = size_t(n) * H;
    const float scale = 1.0f / float(n);
    hipLaunchKernelGGL(k_l1, dim3(uint32_t((nh + 255) / 256)), dim3(256), 0, c->stream, X, n, H, w, m.A1.p, m.H1.p);
    hipLaunchKernelGGL(k_l2, dim3(uint32_t((nh + 255) / 256)), dim3(256), 0, c->stream, m.H1.p, n, H, w, m.A2.p,
                       m.H2.p);
    hipLaunchKernelGGL(k_out_back, dim3((n + 3u) / 4u), dim3(256), 0, c->stream, m.H2.p, m.A2.p, T, n, H, w, scale,
                       m.y.p, m.dA2.p, m.red.p);
    hipLaunchKernelGGL(k_back1, dim3(uint32_t((nh + 255) / 256)), dim3(256), 0, c->stream, m.dA2.p, m.A1.p, n, H, w,
                       m.dA1.p);
    const uint32_t np = mlp_param_count(H);
    hipLaunchKernelGGL(k_grads, dim3((np + 255u) / 256u), dim3(256), 0, c->stream, X, m.H1.p, m.H2.p, m.dA1.p, m.dA2.p,
                       m.y.p, T, n, H, scale, m.grad.p);
    FH_HIP(hipGetLastError());
}

void mlp_set(fh_context* c, uint32_t H, const float* params, const float* inMean, const float* inStd, float outMean,
             float outStd) {
    MlpState& m = c->mlp;
    m.ready = false;
    m.hidden = H;
    const uint32_t np = mlp_param_count(H);
    m.params.ensure(np);
    FH_HIP(hipMemcpyAsync(m.params.p, params, sizeof(float) * np, hipMemcpyHostToDevice, c->stream));
    for (int q = 0; q < kIn; ++q) {
        m.inMean[q] = inMean ? inMean[q] : 0.0f;
        m.inStd[q] = inStd ? inStd[q] : 1.0f;
    }
    m.outMean = outMean;
    m.outStd = outStd;
    m.adamStep = 0;
    m.m.ensure(np);
    m.v.ensure(np);
    launch_fill_f32(c, m.m.p, np, 0.0f);
    launch_fill_f32(c, m.v.p, np, 0.0f);
    FH_HIP(hipStreamSynchronize(c->stream));
    m.ready = true;
}

}  // namespace fh

using fh::fail;

namespace {
template <class Fn>
int guarded_mlp(fh_context* c, Fn&& body) {
    if (!c) return FH_INVALID_ARGUMENT;
    try {
        FH_HIP(hipSetDevice(c->device));
        body();
        c->err.clear();
        return FH_OK;
    } catch (const fh::Error& e) {
        c->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        c->err = e.what();
        return FH_RUNTIME_ERROR;
    }
}
}  // namespace

extern "C" {

uint32_t fh_mlp_param_count(uint32_t hidden) { return fh::mlp_param_count(hidden); }

int fh_mlp_set_params(fh_context* c, uint32_t hidden, const float* params, const float* featureMean,
                      const float* featureStd, float targetMean, float targetStd) {
    return guarded_mlp(c, [&] {
        if (hidden == 0u || hidden > fh::kMaxHidden) fail(FH_INVALID_ARGUMENT, "hidden width must be in [1, 256]");
        if (!params) fail(FH_INVALID_ARGUMENT, "params must not be null");
        if (!(targetStd > 0.0f) || !std::isfinite(targetMean)) fail(FH_INVALID_ARGUMENT, "invalid target scaling");
        for (uint32_t i = 0; i < fh::mlp_param_count(hidden); ++i)
            if (!std::isfinite(params[i])) fail(FH_INVALID_ARGUMENT, "params must be finite");
        if (featureStd)
            for (int q = 0; q < fh::kIn; ++q)
                if (!(featureStd[q] > 0.0f)) fail(FH_INVALID_ARGUMENT, "feature scales must be positive");
        fh::mlp_set(c, hidden, params, featureMean, featureStd, targetMean, targetStd);
    });
}

int fh_mlp_get_params(fh_context* c, float* params) {
    return guarded_mlp(c, [&] {
        if (!c->mlp.ready) fail(FH_LOGIC_ERROR, "no surrogate parameters are set");
        if (!params) fail(FH_INVALID_ARGUMENT, "params must not be null");
        FH_HIP(hipMemcpyAsync(params, c->mlp.params.p, sizeof(float) * fh::mlp_param_count(c->mlp.hidden),
                              hipMemcpyDeviceToHost, c->stream));
        FH_HIP(hipStreamSynchronize(c->stream));
    });
}

int fh_mlp_forward(fh_context* c, const float* x, uint32_t n, float* y) {
    return guarded_mlp(c, [&] {
        if (!c->mlp.ready) fail(FH_LOGIC_ERROR, "no surrogate parameters are set");
        if (!x || !y || n == 0u) fail(FH_INVALID_ARGUMENT, "inputs must be given");
        fh::DBuf<float> dx, dy;
        dx.ensure(size_t(n) * fh::kIn);
        dy.ensure(n);
        FH_HIP(hipMemcpyAsync(dx.p, x, sizeof(float) * n * fh::kIn, hipMemcpyHostToDevice, c->stream));
        fh::infer(c, dx.p, n, dy.p);
        FH_HIP(hipMemcpyAsync(y, dy.p, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
        FH_HIP(hipStreamSynchronize(c->stream));
    });
}

int fh_mlp_loss_grad(fh_context* c, const float* x, const float* target, uint32_t n, float* grad, float* loss) {
    return guarded_mlp(c, [&] {
        if (!c->mlp.ready) fail(FH_LOGIC_ERROR, "no surrogate parameters are set");
        if (!x || !target || n == 0u) fail(FH_INVALID_ARGUMENT, "inputs must be given");
        fh::DBuf<float> dx, dt;
        dx.ensure(size_t(n) * fh::kIn);
        dt.ensure(n);
        FH_HIP(hipMemcpyAsync(dx.p, x, sizeof(float) * n * fh::kIn, hipMemcpyHostToDevice, c->stream));
        FH_HIP(hipMemcpyAsync(dt.p, target, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
        fh::ensure_train(c, n);
        fh::batch_grad(c, dx.p, dt.p, n);
        const uint32_t np = fh::mlp_param_count(c->mlp.hidden);
        if (grad) FH_HIP(hipMemcpyAsync(grad, c->mlp.grad.p, sizeof(float) * np, hipMemcpyDeviceToHost, c->stream));
        std::vector<float> parts(n);
        FH_HIP(hipMemcpyAsync(parts.data(), c->mlp.red.p, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
        FH_HIP(hipStreamSynchronize(c->stream));
        double acc = 0.0;
        for (float p : parts) acc += p;
        if (loss) *loss = float(acc / n);
    });
}

int fh_mlp_train(fh_context* c, const float* x, const float* target, uint32_t n, uint32_t batch, uint32_t epochs,
                 float lr, float lrDecay, uint32_t seed, float* lastLoss) {
    return guarded_mlp(c, [&] {
        if (!c->mlp.ready) fail(FH_LOGIC_ERROR, "no surrogate parameters are set");
        if (!x || !target || n == 0u || batch == 0u) fail(FH_INVALID_ARGUMENT, "training data must be given");
        if (!(lr > 0.0f) || !(lrDecay > 0.0f)) fail(FH_INVALID_ARGUMENT, "learning rate must be positive");
        batch = std::min(batch, n);
        fh::DBuf<float> dx, dt;
        dx.ensure(size_t(n) * fh::kIn);
        dt.ensure(n);
        FH_HIP(hipMemcpyAsync(dx.p, x, sizeof(float) * n * fh::kIn, hipMemcpyHostToDevice, c->stream));
        FH_HIP(hipMemcpyAsync(dt.p, target, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
        fh::ensure_train(c, batch);
        const uint32_t np = fh::mlp_param_count(c->mlp.hidden);
        const uint32_t steps = n / batch;
        std::mt19937 rng(seed);
        std::vector<uint32_t> order(steps);
        for (uint32_t i = 0; i < steps; ++i) order[i] = i;
        const float b1 = 0.9f, b2 = 0.999f;
        float elr = lr;
        for (uint32_t ep = 0; ep < epochs; ++ep, elr *= lrDecay) {
            std::shuffle(order.begin(), order.end(), rng);
            for (uint32_t k = 0; k < steps; ++k) {
                const size_t off = size_t(order[k]) * batch;
                fh::batch_grad(c, dx.p + off * fh::kIn, dt.p + off, batch);
                const double t = double(++c->mlp.adamStep);
                const float c1 = float(1.0 - std::pow(double(b1), t));
                const float c2 = float(1.0 - std::pow(double(b2), t));
                hipLaunchKernelGGL(fh::k_adam, dim3((np + 255u) / 256u), dim3(256), 0, c->stream, c->mlp.params.p,
                                   c->mlp.grad.p, c->mlp.m.p, c->mlp.v.p, np, elr, b1, b2, c1, c2, 1.0e-8f);
                FH_HIP(hipGetLastError());
            }
        }
        std::vector<float> parts(batch);
        FH_HIP(hipMemcpyAsync(parts.data(), c->mlp.red.p, sizeof(float) * batch, hipMemcpyDeviceToHost, c->stream));
        FH_HIP(hipStreamSynchronize(c->stream));
        double acc = 0.0;
        for (float p : parts) acc += p;
        if (lastLoss) *lastLoss = float(acc / batch);
    });
}

uint32_t fh_eb_unit_count(const fh_context* c) { return c && c->eb.ready ? fh::unit_count(c) : 0u; }

int fh_eb_get_features(fh_context* c, float* features, float* temperature) {
    return guarded_mlp(c, [&] {
        fh::EbState& e = c->eb;
        if (!e.ready || !e.featValid) fail(FH_LOGIC_ERROR, "no node has recorded features");
        const uint32_t units = fh::unit_count(c);
        if (features)
            FH_HIP(hipMemcpyAsync(features, e.feat.p, sizeof(float) * units * fh::kIn, hipMemcpyDeviceToHost,
                                  c->stream));
        if (temperature) {
            std::vector<float> T(c->S);
            FH_HIP(hipMemcpyAsync(T.data(), e.T.p, sizeof(float) * c->S, hipMemcpyDeviceToHost, c->stream));
            FH_HIP(hipStreamSynchronize(c->stream));
            for (uint32_t u = 0; u < units; ++u)
                temperature[u] = u < e.leafCount ? T[2u * u] : T[2u * e.leafCount + (u - e.leafCount)];
        }
        FH_HIP(hipStreamSynchronize(c->stream));
    });
}

}  // extern "C"
