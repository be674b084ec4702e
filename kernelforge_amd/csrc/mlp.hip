// mlp.hip -- the preconditioner network: a 2-hidden-layer ReLU MLP with BF16 weights
// and FP32 accumulation, forward / backward / Adam on the GPU.
//
//   Z1 = X W1 + b1, A1 = relu(Z1); Z2 = A1 W2 + b2, A2 = relu(Z2); Y = A2 W3 + b3
// X (B x D0) float, W_l (K x M) bf16 (row-major, K = fan-in), b_l float.  Training keeps
// FP32 master weights; kf_mlp_adam updates them and refreshes the BF16 copies
// (round to nearest even).  Shapes are small (B ~ 10^2-10^3, D0 ~ 722, H = 128), so
// each kernel is one thread per output element with a K-loop in FP32.
#include <cmath>
#include <cstdint>

#include "kf_common.hpp"

namespace kf {
namespace {

__device__ inline float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__host__ __device__ inline uint16_t f32_to_bf16(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u)  // inf / nan: truncate, keep nan quiet
        return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Z[b, m] = sum_k X[b, k] W[k, m] + bias[m]; A = relu(Z) when A != null
__global__ void linear_fwd(const float *__restrict__ X, int B, int K, const uint16_t *__restrict__ W,
                           int M, const float *__restrict__ bias, float *__restrict__ Z,
                           float *__restrict__ A) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= B * M) return;
    const int b = idx / M, m = idx - b * M;
    const float *x = X + (size_t)b * K;
    float s = bias ? bias[m] : 0.0f;
    for (int k = 0; k < K; ++k) s = fmaf(x[k], bf16_to_f32(W[(size_t)k * M + m]), s);
    Z[idx] = s;
    if (A) A[idx] = s > 0.0f ? s : 0.0f;
}

// dX[b, k] = sum_m dZ[b, m] W[k, m]  (then times relu'(Zprev) when Zprev != null)
__global__ void linear_bwd_in(const float *__restrict__ dZ, int B, int M,
                              const uint16_t *__restrict__ W, int K,
                              const float *__restrict__ Zprev, float *__restrict__ dX) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= B * K) return;
    const int b = idx / K, k = idx - b * K;
    const float *dz = dZ + (size_t)b * M;
    const uint16_t *w = W + (size_t)k * M;
    float s = 0.0f;
    for (int m = 0; m < M; ++m) s = fmaf(dz[m], bf16_to_f32(w[m]), s);
    if (Zprev && !(Zprev[idx] > 0.0f)) s = 0.0f;
    dX[idx] = s;
}

// dW[k, m] = sum_b A[b, k] dZ[b, m] (A = relu(Z) when relu_in, else the raw input);
// db[m] = sum_b dZ[b, m] (thread k == 0 of each column)
__global__ void linear_bwd_w(const float *__restrict__ A, int B, int K, int relu_in,
                             const float *__restrict__ dZ, int M, float *__restrict__ dW,
                             float *__restrict__ db) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= K * M) return;
    const int k = idx / M, m = idx - k * M;
    float s = 0.0f, sb = 0.0f;
    for (int b = 0; b < B; ++b) {
        float a = A[(size_t)b * K + k];
        if (relu_in && !(a > 0.0f)) a = 0.0f;
        const float g = dZ[(size_t)b * M + m];
        s = fmaf(a, g, s);
        sb += g;
    }
    dW[idx] = s;
    if (k == 0 && db) db[m] = sb;
}

__global__ void adam_kernel(float *__restrict__ p, const float *__restrict__ g,
                            float *__restrict__ m, float *__restrict__ v, int n, float lr,
                            float b1, float b2, float eps, float c1, float c2,
                            uint16_t *__restrict__ pb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    if (pb) pb[i] = f32_to_bf16(pi);
}

__global__ void cast_bf16_kernel(const float *__restrict__ p, int n, uint16_t *__restrict__ pb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pb[i] = f32_to_bf16(p[i]);
}

inline dim3 grid1(long n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace
}  // namespace kf

using namespace kf;

extern "C" int kf_mlp_forward(const float *X, int B, int D0, int H, int D3, const uint16_t *W1,
                              const float *b1, const uint16_t *W2, const float *b2,
                              const uint16_t *W3, const float *b3, float *Z1, float *Z2,
                              float *Y, void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    KF_REQUIRE(B >= 0 && D0 > 0 && H > 0 && D3 > 0, KF_ERR_ARG, "bad MLP shape");
    if (B == 0) return KF_OK;
    KF_REQUIRE(X && W1 && W2 && W3 && Z1 && Z2 && Y, KF_ERR_ARG, "null pointer");
    hipStream_t st = as_stream(stream);
    // activations A1, A2 live in scratch; the backward recomputes relu from Z1, Z2
    float *A1 = static_cast<float *>(workspace(WS_NET, sizeof(float) * (size_t)2 * B * H, st));
    KF_REQUIRE(A1, KF_ERR_NOMEM, "workspace allocation failed");
    float *A2 = A1 + (size_t)B * H;
    hipLaunchKernelGGL(linear_fwd, grid1((long)B * H), dim3(256), 0, st, X, B, D0, W1, H, b1, Z1,
                       A1);
    hipLaunchKernelGGL(linear_fwd, grid1((long)B * H), dim3(256), 0, st, A1, B, H, W2, H, b2, Z2,
                       A2);
    hipLaunchKernelGGL(linear_fwd, grid1((long)B * D3), dim3(256), 0, st, A2, B, H, W3, D3, b3, Y,
                       (float *)nullptr);
    KF_HIP_TRY(hipGetLastError());
    return KF_OK;
}

extern "C" int kf_mlp_backward(const float *X, int B, int D0, int H, int D3, const uint16_t *W1,
                               const uint16_t *W2, const uint16_t *W3, const float *Z1,
                               const float *Z2, const float *dY, float *dW1, float *db1,
                               float *dW2, float *db2, float *dW3, float *db3, void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    KF_REQUIRE(B >= 0 && D0 > 0 && H > 0 && D3 > 0, KF_ERR_ARG, "bad MLP shape");
    if (B == 0) return KF_OK;
    KF_REQUIRE(X && W1 && W2 && W3 && Z1 && Z2 && dY && dW1 && dW2 && dW3, KF_ERR_ARG,
               "null pointer");
    (void)W1;
    hipStream_t st = as_stream(stream);
    float *d2 = static_cast<float *>(workspace(WS_NET, sizeof(float) * (size_t)2 * B * H, st));
    KF_REQUIRE(d2, KF_ERR_NOMEM, "workspace allocation failed");
    float *d1 = d2 + (size_t)B * H;
    // layer 3: dW3 = relu(Z2)^T dY ; dZ2 = (dY W3^T) * relu'(Z2)
    hipLaunchKernelGGL(linear_bwd_w, grid1((long)H * D3), dim3(256), 0, st, Z2, B, H, 1, dY, D3,
                       dW3, db3);
    hipLaunchKernelGGL(linear_bwd_in, grid1((long)B * H), dim3(256), 0, st, dY, B, D3, W3, H, Z2,
                       d2);
    // layer 2: dW2 = relu(Z1)^T dZ2 ; dZ1 = (dZ2 W2^T) * relu'(Z1)
    hipLaunchKernelGGL(linear_bwd_w, grid1((long)H * H), dim3(256), 0, st, Z1, B, H, 1, d2, H, dW2,
                       db2);
    hipLaunchKernelGGL(linear_bwd_in, grid1((long)B * H), dim3(256), 0, st, d2, B, H, W2, H, Z1,
                       d1);
    // layer 1: dW1 = X^T dZ1
    hipLaunchKernelGGL(linear_bwd_w, grid1((long)D0 * H), dim3(256), 0, st, X, B, D0, 0, d1, H,
                       dW1, db1);
    KF_HIP_TRY(hipGetLastError());
    return KF_OK;
}

extern "C" int kf_mlp_adam(float *p, const float *g, float *m, float *v, int n, float lr,
                           float beta1, float beta2, float eps, int t, uint16_t *p_bf16,
                           void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    KF_REQUIRE(n >= 0 && t >= 1, KF_ERR_ARG, "bad Adam arguments");
    if (n == 0) return KF_OK;
    const float c1 = 1.0f - std::pow(beta1, (float)t), c2 = 1.0f - std::pow(beta2, (float)t);
    hipLaunchKernelGGL(adam_kernel, grid1(n), dim3(256), 0, as_stream(stream), p, g, m, v, n, lr,
                       beta1, beta2, eps, c1, c2, p_bf16);
    KF_HIP_TRY(hipGetLastError());
    return KF_OK;
}

extern "C" int kf_cast_bf16(const float *p, int n, uint16_t *out, void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    KF_REQUIRE(n >= 0, KF_ERR_ARG, "bad size");
    if (n == 0) return KF_OK;
    hipLaunchKernelGGL(cast_bf16_kernel, grid1(n), dim3(256), 0, as_stream(stream), p, n, out);
    KF_HIP_TRY(hipGetLastError());
    return KF_OK;
}
