// Learned initial guess of the implicit step: a per-fiber multilayer
// perceptron (input 384 = 31 points x 3 coordinates x {position relative to
// the centroid, rhs, previous solution, block-Jacobi correction of M^-1 b},
// zero-padded; three hidden layers of
// 256 with GELU; output 96) that predicts the correction of the block-Jacobi
// solution: x0 = M^-1 b + s * MLP(features), s = |rhs|_inf of the fiber.
// It has no counterpart in the reference (meca.cc:704 starts from zero); the
// solver still iterates to the same tolerance from this guess.
//
// Weights are FP32 masters with BF16 copies; the GEMMs run on MFMA
// (v_mfma_f32_16x16x16bf16_1k: BF16 operands, FP32 accumulation), activations
// are stored in BF16 between layers, pre-activations in FP32.  Training
// (forward, backward, Adam) runs on the device on (features, target) pairs
// recorded from converged solves.
#include <cmath>
#include <cstring>
#include <random>

#include "args.h"
#include "mlp.h"

namespace cymg {

static_assert(kMlpWeights == CYMG_GUESS_WEIGHTS, "weight count of the header");
static_assert(kMlpIn == CYMG_GUESS_INPUTS && kMlpOut == CYMG_GUESS_OUTPUTS && kMlpPoints == CYMG_GUESS_POINTS,
              "sizes of the header");

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t to_bf16(float f)
{
    // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);   // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float z)
{
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// C[i][j] (+)= sum_r A(i, r) B(r, j) (+ bias[j]); A(i, r) = A[i*ai + r*ar],
// B(r, j) = B[r*br + j*bj], both BF16; C FP32 with row stride ldc.  M, N, R
// are multiples of 16.  One wavefront per 16x16 tile of C, four per
// workgroup along M.  MFMA lane layout (16x16x16, checked on the device):
// lane l holds A(l%16, 4(l/16)+q), B(4(l/16)+q, l%16), C(4(l/16)+q, l%16).
__global__ __launch_bounds__(256) void k_gemm_bf16(const uint16_t* __restrict__ A, long ai, long ar,
                                                   const uint16_t* __restrict__ B, long br, long bj,
                                                   float* __restrict__ C, long ldc, const float* __restrict__ bias,
                                                   int M, int N, int R, int accumulate)
{
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int tm = blockIdx.x * 4 + wave, tn = blockIdx.y;
    if (tm * 16 >= M || tn * 16 >= N) return;
    const long i = tm * 16 + (l & 15);
    const long j = tn * 16 + (l & 15);
    const int kq = 4 * (l >> 4);
    f32x4 c = { 0.f, 0.f, 0.f, 0.f };
    for (int r0 = 0; r0 < R; r0 += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const long r = r0 + kq + q;
            a[q] = (short)A[i * ai + r * ar];
            b[q] = (short)B[r * br + j * bj];
        }
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const long row = tm * 16 + kq + q;
        float v = c[q];
        if (bias) v += bias[j];
        float* dst = C + row * ldc + j;
        *dst = accumulate ? *dst + v : v;
    }
}

// H = bf16(gelu(Z)), elementwise over count values
__global__ void k_gelu_fwd(const float* __restrict__ Z, uint16_t* __restrict__ H, long count)
{
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t < count) H[t] = to_bf16(gelu(Z[t]));
}

// dZ = dH * gelu'(Z); the FP32 value for the bias gradient, BF16 for the GEMMs
__global__ void k_gelu_bwd(const float* __restrict__ dH, const float* __restrict__ Z, float* __restrict__ dZ,
                           uint16_t* __restrict__ dZb, long count)
{
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t < count) {
        const float v = dH[t] * gelu_grad(Z[t]);
        dZ[t] = v;
        dZb[t] = to_bf16(v);
    }
}

// column sums of a [rows x cols] FP32 matrix (bias gradients)
__global__ void k_colsum(const float* __restrict__ X, int rows, int cols, float* __restrict__ out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    float s = 0.f;
    for (int i = 0; i < rows; ++i) s += X[(long)i * cols + j];
    out[j] = s;
}

// loss = sum over samples and valid outputs of (Y - T)^2 / B; dY = 2 (Y - T) / B.
// T[.][kOut-1] holds the number of valid outputs of the sample.
__global__ void k_mse(const float* __restrict__ Y, const float* __restrict__ T, int B, float* __restrict__ dY,
                      uint16_t* __restrict__ dYb, float* __restrict__ loss)
{
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    float e2 = 0.f;
    if (t < (long)B * kMlpOut) {
        const long row = t / kMlpOut;
        const int o = (int)(t - row * kMlpOut);
        const int nvalid = (int)T[row * kMlpOut + kMlpOut - 1];
        float g = 0.f;
        if (o < nvalid) {
            const float e = Y[t] - T[t];
            e2 = e * e / (float)B;
            g = 2.f * e / (float)B;
        }
        dY[t] = g;
        dYb[t] = to_bf16(g);
    }
    for (int o = 32; o > 0; o >>= 1) e2 += __shfl_xor(e2, o, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0 && e2 != 0.f) atomicAdd(loss, e2);
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, uint16_t* __restrict__ wb, long count, float lr, float c1, float c2)
{
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= count) return;
    const float gi = g[t];
    const float mi = 0.9f * m[t] + 0.1f * gi;
    const float vi = 0.999f * v[t] + 0.001f * gi * gi;
    m[t] = mi;
    v[t] = vi;
    const float wi = w[t] - lr * (mi * c1) / (sqrtf(vi * c2) + 1e-8f);
    w[t] = wi;
    wb[t] = to_bf16(wi);
}

__global__ void k_to_bf16(const float* __restrict__ x, uint16_t* __restrict__ y, long count)
{
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t < count) y[t] = to_bf16(x[t]);
}

// Features of each fiber (one wavefront per fiber; fibers of more than
// kMlpPoints points get zero features and scale 0): per point, the position
// about the centroid / 5, then rhs, previous solution and the block-Jacobi
// correction z1 = M^-1 (b - A M^-1 b), each / s with s = |rhs|_inf of the
// fiber.  z1 is the coupling term: the fiber's own block is solved exactly by
// M^-1 b, so its residual holds what the linked fibers add through mB/mC.
template <int D>
__global__ __launch_bounds__(kWave) void k_features(FiberArgs a, const double* __restrict__ pts,
                                                    const double* __restrict__ rhs, const double* __restrict__ prev,
                                                    const double* __restrict__ corr, uint16_t* __restrict__ X,
                                                    float* __restrict__ scale)
{
    const uint32_t f = blockIdx.x;
    const int lane = threadIdx.x;
    const int np = (int)a.np[f];
    const uint32_t o = D * a.idx[f];
    uint16_t* row = X + (long)f * kMlpIn;
    const bool ok = np <= kMlpPoints;
    double c[D], r[D], q[D], z[D], cs[D];
    double rm = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        c[d] = r[d] = q[d] = z[d] = 0.0;
        if (ok && lane < np) {
            c[d] = pts[o + D * lane + d];
            r[d] = rhs[o + D * lane + d];
            q[d] = prev[o + D * lane + d];
            z[d] = corr[o + D * lane + d];
            rm = fmax(rm, fabs(r[d]));
        }
        cs[d] = wave_sum(c[d]) / (double)np;
    }
    rm = wave_max(rm);
    const double s = (ok && rm > 0.0) ? rm : 0.0;
    const double inv = s > 0.0 ? 1.0 / s : 0.0;
    for (int t = lane; t < kMlpIn; t += kWave) row[t] = 0;
    if (ok && lane < np) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            row[3 * lane + d] = to_bf16((float)((c[d] - cs[d]) * 0.2));
            row[3 * kMlpPoints + 3 * lane + d] = to_bf16((float)(r[d] * inv));
            row[6 * kMlpPoints + 3 * lane + d] = to_bf16((float)(q[d] * inv));
            row[9 * kMlpPoints + 3 * lane + d] = to_bf16((float)(z[d] * inv));
        }
    }
    if (lane == 0) scale[f] = (float)s;
}

// x0 = y + s * out for the points of eligible fibers, y elsewhere
template <int D>
__global__ __launch_bounds__(kWave) void k_guess(FiberArgs a, const double* __restrict__ y,
                                                 const float* __restrict__ out, const float* __restrict__ scale,
                                                 double* __restrict__ x0)
{
    const uint32_t f = blockIdx.x;
    const int np = (int)a.np[f];
    const uint32_t o = D * a.idx[f];
    const float s = scale[f];
    for (int t = threadIdx.x; t < D * np; t += kWave) {
        const int p = t / D, d = t - D * (t / D);
        double v = y[o + t];
        if (s > 0.f && np <= kMlpPoints) v += (double)s * (double)out[(long)f * kMlpOut + 3 * p + d];
        x0[o + t] = v;
    }
}

// targets (x - y) / s of each fiber; the last column holds the valid count
template <int D>
__global__ __launch_bounds__(kWave) void k_targets(FiberArgs a, const double* __restrict__ x,
                                                   const double* __restrict__ y, const float* __restrict__ scale,
                                                   float* __restrict__ T)
{
    const uint32_t f = blockIdx.x;
    const int np = (int)a.np[f];
    const uint32_t o = D * a.idx[f];
    const float s = scale[f];
    float* row = T + (long)f * kMlpOut;
    for (int t = threadIdx.x; t < kMlpOut; t += kWave) row[t] = 0.f;
    bool ok = s > 0.f && np <= kMlpPoints;
    float big = 0.f;
    if (ok)
        for (int t = threadIdx.x; t < D * np; t += kWave) {
            const int p = t / D, d = t - D * (t / D);
            const float v = (float)((x[o + t] - y[o + t]) / (double)s);
            row[3 * p + d] = v;
            big = fmaxf(big, fabsf(v));
        }
    // a fiber whose correction is far beyond its own load scale (a lightly
    // loaded fiber moved by its links) is not a sample: it would dominate the
    // squared loss
    for (int w = 32; w > 0; w >>= 1) big = fmaxf(big, __shfl_xor(big, w, kWave));
    if (!(big <= kMlpTargetClip)) ok = false;
    // valid outputs: 3 per point (the third coordinate stays zero in 2D)
    if (threadIdx.x == 0) row[kMlpOut - 1] = ok ? (float)(3 * np) : 0.f;
}

// random minibatch rows from the dataset (hash of seed, step and row)
__global__ void k_gather_batch(const uint16_t* __restrict__ dsx, const float* __restrict__ dst, uint32_t count,
                               uint64_t seed, uint32_t step, int B, uint16_t* __restrict__ X, float* __restrict__ T)
{
    const int b = blockIdx.x;
    uint64_t h = seed ^ (0x9e3779b97f4a7c15ull * (step + 1)) ^ (0xbf58476d1ce4e5b9ull * (uint64_t)(b + 1));
    h ^= h >> 31; h *= 0x94d049bb133111ebull; h ^= h >> 29;
    const uint32_t src = (uint32_t)(h % count);
    for (int t = threadIdx.x; t < kMlpIn; t += blockDim.x) X[(long)b * kMlpIn + t] = dsx[(long)src * kMlpIn + t];
    for (int t = threadIdx.x; t < kMlpOut; t += blockDim.x) T[(long)b * kMlpOut + t] = dst[(long)src * kMlpOut + t];
}

// ---------------------------------------------------------------------------
static void gemm(Context& c, const uint16_t* A, long ai, long ar, const uint16_t* B, long br, long bj, float* C,
                 long ldc, const float* bias, int M, int N, int R, int acc)
{
    const dim3 g((M / 16 + 3) / 4, N / 16), b(256);
    hipLaunchKernelGGL(k_gemm_bf16, g, b, 0, c.stream, A, ai, ar, B, br, bj, C, ldc, bias, M, N, R, acc);
}

static dim3 grid1(long n, int b = 256) { return dim3((unsigned)((n + b - 1) / b)); }

static const int kLayerIn[kMlpLayers] = { kMlpIn, kMlpHidden, kMlpHidden, kMlpHidden };
static const int kLayerOut[kMlpLayers] = { kMlpHidden, kMlpHidden, kMlpHidden, kMlpOut };

static long weight_offset(int layer)
{
    long off = 0;
    for (int q = 0; q < layer; ++q) off += (long)kLayerIn[q] * kLayerOut[q] + kLayerOut[q];
    return off;
}

static void ensure_work(Context& c, int M)
{
    MlpState& s = c.mlp;
    if (s.work_rows >= M) return;
    for (int q = 0; q < kMlpLayers; ++q) {
        s.z[q].resize((size_t)M * kLayerOut[q]);
        s.h[q].resize((size_t)M * kLayerIn[q]);
        s.dz[q].resize((size_t)M * kLayerOut[q]);
        s.dzb[q].resize((size_t)M * kLayerOut[q]);
    }
    s.dh.resize((size_t)M * kMlpHidden);
    s.tgt.resize((size_t)M * kMlpOut);
    s.loss.resize(1);
    s.work_rows = M;
}

// forward over M rows of s.h[0]; s.z[last] holds the output
static void forward(Context& c, int M)
{
    MlpState& s = c.mlp;
    for (int q = 0; q < kMlpLayers; ++q) {
        const long off = weight_offset(q);
        const uint16_t* W = s.wb.get() + off;
        const float* bias = s.w.get() + off + (long)kLayerIn[q] * kLayerOut[q];
        gemm(c, s.h[q].get(), kLayerIn[q], 1, W, kLayerOut[q], 1, s.z[q].get(), kLayerOut[q], bias, M,
             kLayerOut[q], kLayerIn[q], 0);
        if (q + 1 < kMlpLayers) {
            const long cnt = (long)M * kLayerOut[q];
            hipLaunchKernelGGL(k_gelu_fwd, grid1(cnt), dim3(256), 0, c.stream, s.z[q].get(), s.h[q + 1].get(), cnt);
        }
    }
}

// backward of the MSE loss against s.tgt over B rows; gradients into s.g
static void backward(Context& c, int B)
{
    MlpState& s = c.mlp;
    const int L = kMlpLayers - 1;
    CYMG_HIP(hipMemsetAsync(s.loss.get(), 0, sizeof(float), c.stream));
    const long cnt = (long)B * kMlpOut;
    hipLaunchKernelGGL(k_mse, grid1(cnt), dim3(256), 0, c.stream, s.z[L].get(), s.tgt.get(), B, s.dz[L].get(),
                       s.dzb[L].get(), s.loss.get());
    for (int q = L; q >= 0; --q) {
        const long off = weight_offset(q);
        const int K = kLayerIn[q], N = kLayerOut[q];
        float* gW = s.g.get() + off;
        float* gb = gW + (long)K * N;
        // dW = H^T dZ  (K x N), reduction over the batch
        gemm(c, s.h[q].get(), 1, K, s.dzb[q].get(), N, 1, gW, N, nullptr, K, N, B, 0);
        hipLaunchKernelGGL(k_colsum, grid1(N, 128), dim3(128), 0, c.stream, s.dz[q].get(), B, N, gb);
        if (q == 0) break;
        // dH = dZ W^T (B x K), then through the previous layer's GELU
        const uint16_t* W = s.wb.get() + off;
        gemm(c, s.dzb[q].get(), N, 1, W, 1, N, s.dh.get(), K, nullptr, B, K, N, 0);
        const long cq = (long)B * K;
        hipLaunchKernelGGL(k_gelu_bwd, grid1(cq), dim3(256), 0, c.stream, s.dh.get(), s.z[q - 1].get(),
                           s.dz[q - 1].get(), s.dzb[q - 1].get(), cq);
    }
}

void mlp_init(Context& c, uint64_t seed)
{
    MlpState& s = c.mlp;
    const long n = kMlpWeights;
    s.w.resize(n); s.wb.resize(n); s.m.resize(n); s.v.resize(n); s.g.resize(n);
    std::vector<float> w(n, 0.f);
    std::mt19937_64 rng(seed);
    for (int q = 0; q < kMlpLayers; ++q) {
        const long off = weight_offset(q);
        const float a = 1.0f / std::sqrt((float)kLayerIn[q]);
        std::uniform_real_distribution<float> U(-a, a);
        for (long t = 0; t < (long)kLayerIn[q] * kLayerOut[q]; ++t) w[off + t] = U(rng);
    }
    mlp_set_weights(c, w.data());
}

void mlp_set_weights(Context& c, const float* w)
{
    MlpState& s = c.mlp;
    const long n = kMlpWeights;
    s.w.resize(n); s.wb.resize(n); s.m.resize(n); s.v.resize(n); s.g.resize(n);
    CYMG_HIP(hipMemcpyAsync(s.w.get(), w, n * sizeof(float), hipMemcpyHostToDevice, c.stream));
    CYMG_HIP(hipMemsetAsync(s.m.get(), 0, n * sizeof(float), c.stream));
    CYMG_HIP(hipMemsetAsync(s.v.get(), 0, n * sizeof(float), c.stream));
    hipLaunchKernelGGL(k_to_bf16, grid1(n), dim3(256), 0, c.stream, s.w.get(), s.wb.get(), n);
    CYMG_HIP(hipStreamSynchronize(c.stream));   // pageable source
    s.adam_t = 0;
    s.ready = true;
}

void mlp_get_weights(Context& c, float* w)
{
    CYMG_HIP(hipMemcpyAsync(w, c.mlp.w.get(), kMlpWeights * sizeof(float), hipMemcpyDeviceToHost, c.stream));
    CYMG_HIP(hipStreamSynchronize(c.stream));
}

double mlp_check(Context& c, const float* X, const float* T, int B, float* grads)
{
    if (B <= 0 || B % 16) throw Error(CYMG_ERR_INVALID, "batch must be a positive multiple of 16");
    MlpState& s = c.mlp;
    ensure_work(c, B);
    DevBuf<float> xf;
    xf.resize((size_t)B * kMlpIn);
    CYMG_HIP(hipMemcpyAsync(xf.get(), X, (size_t)B * kMlpIn * 4, hipMemcpyHostToDevice, c.stream));
    CYMG_HIP(hipMemcpyAsync(s.tgt.get(), T, (size_t)B * kMlpOut * 4, hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_to_bf16, grid1((long)B * kMlpIn), dim3(256), 0, c.stream, xf.get(), s.h[0].get(),
                       (long)B * kMlpIn);
    forward(c, B);
    backward(c, B);
    float loss = 0.f;
    CYMG_HIP(hipMemcpyAsync(&loss, s.loss.get(), 4, hipMemcpyDeviceToHost, c.stream));
    CYMG_HIP(hipMemcpyAsync(grads, s.g.get(), kMlpWeights * sizeof(float), hipMemcpyDeviceToHost, c.stream));
    CYMG_HIP(hipStreamSynchronize(c.stream));
    return loss;
}

double mlp_train(Context& c, int steps, int B, double lr, uint64_t seed)
{
    MlpState& s = c.mlp;
    if (B <= 0 || B % 16) throw Error(CYMG_ERR_INVALID, "batch must be a positive multiple of 16");
    if (s.ds_count == 0) throw Error(CYMG_ERR_INVALID, "no recorded samples to train on");
    ensure_work(c, B);
    double last = 0.0;
    for (int t = 0; t < steps; ++t) {
        hipLaunchKernelGGL(k_gather_batch, dim3(B), dim3(128), 0, c.stream, s.ds_x.get(), s.ds_t.get(), s.ds_count,
                           seed, (uint32_t)t, B, s.h[0].get(), s.tgt.get());
        forward(c, B);
        backward(c, B);
        s.adam_t += 1;
        const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, s.adam_t)));
        const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, s.adam_t)));
        hipLaunchKernelGGL(k_adam, grid1(kMlpWeights), dim3(256), 0, c.stream, s.w.get(), s.g.get(), s.m.get(),
                           s.v.get(), s.wb.get(), (long)kMlpWeights, (float)lr, c1, c2);
        if (t + 1 == steps) {
            float loss = 0.f;
            CYMG_HIP(hipMemcpyAsync(&loss, s.loss.get(), 4, hipMemcpyDeviceToHost, c.stream));
            CYMG_HIP(hipStreamSynchronize(c.stream));
            last = loss;
        }
    }
    CYMG_HIP(hipGetLastError());
    return last;
}

double mlp_eval(Context& c, int B, uint64_t seed)
{
    MlpState& s = c.mlp;
    if (B <= 0 || B % 16) throw Error(CYMG_ERR_INVALID, "batch must be a positive multiple of 16");
    if (s.ds_count == 0) throw Error(CYMG_ERR_INVALID, "no recorded samples to evaluate on");
    ensure_work(c, B);
    hipLaunchKernelGGL(k_gather_batch, dim3(B), dim3(128), 0, c.stream, s.ds_x.get(), s.ds_t.get(), s.ds_count, seed,
                       0u, B, s.h[0].get(), s.tgt.get());
    forward(c, B);
    const int L = kMlpLayers - 1;
    CYMG_HIP(hipMemsetAsync(s.loss.get(), 0, sizeof(float), c.stream));
    const long cnt = (long)B * kMlpOut;
    hipLaunchKernelGGL(k_mse, grid1(cnt), dim3(256), 0, c.stream, s.z[L].get(), s.tgt.get(), B, s.dz[L].get(),
                       s.dzb[L].get(), s.loss.get());
    CYMG_HIP(hipGetLastError());
    float loss = 0.f;
    CYMG_HIP(hipMemcpyAsync(&loss, s.loss.get(), 4, hipMemcpyDeviceToHost, c.stream));
    CYMG_HIP(hipStreamSynchronize(c.stream));
    return loss;
}

// guard of the learned guess: per block, max |b - Ax| and max |b| (NaN kept)
__global__ __launch_bounds__(256) void k_guard_part(uint32_t n, const double* __restrict__ b,
                                                    const double* __restrict__ ax, double* __restrict__ part)
{
    double rm = 0.0, bm = 0.0;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const double r = fabs(b[i] - ax[i]);
        rm = (r > rm || r != r) ? r : rm;
        bm = fmax(bm, fabs(b[i]));
    }
    for (int w = 32; w > 0; w >>= 1) {
        const double o = __shfl_xor(rm, w, kWave);
        rm = (o > rm || o != o) ? o : rm;
        bm = fmax(bm, __shfl_xor(bm, w, kWave));
    }
    __shared__ double sr[4], sb[4];
    if ((threadIdx.x & 63) == 0) {
        sr[threadIdx.x / 64] = rm;
        sb[threadIdx.x / 64] = bm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < 4; ++q) {
            rm = (sr[q] > rm || sr[q] != sr[q]) ? sr[q] : rm;
            bm = fmax(bm, sb[q]);
        }
        part[2 * blockIdx.x] = rm;
        part[2 * blockIdx.x + 1] = bm;
    }
}

// one wavefront: the guess is dropped (flag 1) when max |b - A x0| is not below max |b|
__global__ void k_guard_decide(uint32_t nb, const double* __restrict__ part, int* __restrict__ flag)
{
    double rm = 0.0, bm = 0.0;
    for (uint32_t i = threadIdx.x; i < nb; i += kWave) {
        const double r = part[2 * i];
        rm = (r > rm || r != r) ? r : rm;
        bm = fmax(bm, part[2 * i + 1]);
    }
    for (int w = 32; w > 0; w >>= 1) {
        const double o = __shfl_xor(rm, w, kWave);
        rm = (o > rm || o != o) ? o : rm;
        bm = fmax(bm, __shfl_xor(bm, w, kWave));
    }
    if (threadIdx.x == 0) *flag = (rm < bm) ? 0 : 1;
}

__global__ void k_guard_apply(uint32_t n, const int* __restrict__ flag, double* __restrict__ x0)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && *flag) x0[i] = 0.0;
}

// r = b - r
__global__ void k_residual(uint32_t n, const double* __restrict__ b, double* __restrict__ r)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = b[i] - r[i];
}

// y = M^-1 b (0 without preconditioner), z1 = M^-1 (b - A y) and the
// features of the step
void mlp_prepare_step(Context& c, int precond)
{
    MlpState& s = c.mlp;
    const uint32_t F = c.nfib;
    const uint32_t n = c.n();
    const int M = (int)((F + 15) / 16 * 16);
    ensure_work(c, M);
    s.fx.resize((size_t)M * kMlpIn);
    s.fs.resize(M);
    s.y.resize(n);
    s.x0.resize(n);
    s.z1.resize(n);
    CYMG_HIP(hipMemsetAsync(s.fx.get(), 0, (size_t)M * kMlpIn * 2, c.stream));
    if (precond) {
        launch_precond_apply(c, c.d_rhs.get(), s.y.get());
        launch_matvec(c, s.y.get(), s.x0.get());          // x0 holds A y until the guess is formed
        hipLaunchKernelGGL(k_residual, grid1(n), dim3(256), 0, c.stream, n, c.d_rhs.get(), s.x0.get());
        launch_precond_apply(c, s.x0.get(), s.z1.get());
    } else {
        CYMG_HIP(hipMemsetAsync(s.y.get(), 0, (size_t)n * 8, c.stream));
        CYMG_HIP(hipMemsetAsync(s.z1.get(), 0, (size_t)n * 8, c.stream));
    }
    const FiberArgs a = make_fiber_args(c);
    if (c.dim == 3)
        hipLaunchKernelGGL(k_features<3>, dim3(F), dim3(kWave), 0, c.stream, a, c.d_pts.get(), c.d_rhs.get(),
                           c.d_prev_sol.get(), s.z1.get(), s.fx.get(), s.fs.get());
    else
        hipLaunchKernelGGL(k_features<2>, dim3(F), dim3(kWave), 0, c.stream, a, c.d_pts.get(), c.d_rhs.get(),
                           c.d_prev_sol.get(), s.z1.get(), s.fx.get(), s.fs.get());
    CYMG_HIP(hipGetLastError());
    s.step_ready = true;
}

// the learned guess of the prepared step, in s.x0
const double* mlp_guess(Context& c)
{
    MlpState& s = c.mlp;
    if (!s.ready) throw Error(CYMG_ERR_UNSUPPORTED, "learned guess not trained or loaded");
    const uint32_t F = c.nfib;
    const int M = (int)((F + 15) / 16 * 16);
    CYMG_HIP(hipMemcpyAsync(s.h[0].get(), s.fx.get(), (size_t)M * kMlpIn * 2, hipMemcpyDeviceToDevice, c.stream));
    forward(c, M);
    if (c.nsol) {
        // Solids are seeded with M^-1 b
        const size_t o = (size_t)c.dim * c.fib_pts;
        CYMG_HIP(hipMemcpyAsync(s.x0.get() + o, s.y.get() + o, (c.n() - o) * 8, hipMemcpyDeviceToDevice, c.stream));
    }
    const FiberArgs a = make_fiber_args(c);
    if (c.dim == 3)
        hipLaunchKernelGGL(k_guess<3>, dim3(F), dim3(kWave), 0, c.stream, a, s.y.get(),
                           s.z[kMlpLayers - 1].get(), s.fs.get(), s.x0.get());
    else
        hipLaunchKernelGGL(k_guess<2>, dim3(F), dim3(kWave), 0, c.stream, a, s.y.get(),
                           s.z[kMlpLayers - 1].get(), s.fs.get(), s.x0.get());
    // the plain method stays the judge: a guess whose residual is not below
    // that of the zero guess (|b - A x0|_inf >= |b|_inf) is replaced by zero,
    // on the device, without a host round trip
    const uint32_t n = c.n();
    const uint32_t nb = std::min<uint32_t>(1024u, (n + 255) / 256);
    s.gpart.resize(2 * (size_t)nb);
    s.gflag.resize(1);
    if (!s.h_gflag) CYMG_HIP(hipHostMalloc((void**)&s.h_gflag, sizeof(int), hipHostMallocDefault));
    if (!s.guard) {
        *s.h_gflag = 0;
        return s.x0.get();
    }
    launch_matvec(c, s.x0.get(), s.z1.get());       // z1 is free once the features are built
    hipLaunchKernelGGL(k_guard_part, dim3(nb), dim3(256), 0, c.stream, n, c.d_rhs.get(), s.z1.get(),
                       s.gpart.get());
    hipLaunchKernelGGL(k_guard_decide, dim3(1), dim3(kWave), 0, c.stream, nb, s.gpart.get(), s.gflag.get());
    hipLaunchKernelGGL(k_guard_apply, grid1(n), dim3(256), 0, c.stream, n, s.gflag.get(), s.x0.get());
    CYMG_HIP(hipMemcpyAsync(s.h_gflag, s.gflag.get(), sizeof(int), hipMemcpyDeviceToHost, c.stream));
    CYMG_HIP(hipGetLastError());
    return s.x0.get();
}

// whether the last guess was dropped (valid once the stream has passed mlp_guess)
bool mlp_guess_dropped(Context& c)
{
    return c.mlp.h_gflag && *c.mlp.h_gflag != 0;
}

// append the prepared step's features and the targets of the converged solution
void mlp_record(Context& c)
{
    MlpState& s = c.mlp;
    if (!s.recording || !s.step_ready) return;
    const uint32_t F = c.nfib;
    if (F > s.ds_cap) return;
    DevBuf<float>& T = s.tgt;
    const int M = (int)((F + 15) / 16 * 16);
    ensure_work(c, M);
    const FiberArgs a = make_fiber_args(c);
    if (c.dim == 3)
        hipLaunchKernelGGL(k_targets<3>, dim3(F), dim3(kWave), 0, c.stream, a, c.d_sol.get(), s.y.get(), s.fs.get(),
                           T.get());
    else
        hipLaunchKernelGGL(k_targets<2>, dim3(F), dim3(kWave), 0, c.stream, a, c.d_sol.get(), s.y.get(), s.fs.get(),
                           T.get());
    // ring buffer: copy in at most two pieces
    uint32_t done = 0;
    while (done < F) {
        const uint32_t at = s.ds_next;
        const uint32_t n = std::min(F - done, s.ds_cap - at);
        CYMG_HIP(hipMemcpyAsync(s.ds_x.get() + (size_t)at * kMlpIn, s.fx.get() + (size_t)done * kMlpIn,
                                (size_t)n * kMlpIn * 2, hipMemcpyDeviceToDevice, c.stream));
        CYMG_HIP(hipMemcpyAsync(s.ds_t.get() + (size_t)at * kMlpOut, T.get() + (size_t)done * kMlpOut,
                                (size_t)n * kMlpOut * 4, hipMemcpyDeviceToDevice, c.stream));
        done += n;
        s.ds_next = (at + n) % s.ds_cap;
    }
    s.ds_count = std::min(s.ds_cap, s.ds_count + F);
    CYMG_HIP(hipGetLastError());
}

void mlp_features(Context& c, float* X, float* scale)
{
    MlpState& s = c.mlp;
    const uint32_t F = c.nfib;
    std::vector<uint16_t> h((size_t)F * kMlpIn);
    CYMG_HIP(hipMemcpyAsync(h.data(), s.fx.get(), h.size() * 2, hipMemcpyDeviceToHost, c.stream));
    CYMG_HIP(hipMemcpyAsync(scale, s.fs.get(), (size_t)F * 4, hipMemcpyDeviceToHost, c.stream));
    CYMG_HIP(hipStreamSynchronize(c.stream));
    for (size_t t = 0; t < h.size(); ++t) {
        const uint32_t u = (uint32_t)h[t] << 16;
        std::memcpy(X + t, &u, 4);
    }
}

void mlp_set_record(Context& c, uint32_t capacity)
{
    MlpState& s = c.mlp;
    s.recording = capacity > 0;
    if (!capacity) return;   // stop; the samples stay for training
    s.ds_cap = capacity;
    s.ds_count = s.ds_next = 0;
    s.ds_x.resize((size_t)capacity * kMlpIn);
    s.ds_t.resize((size_t)capacity * kMlpOut);
}

}  // namespace cymg
