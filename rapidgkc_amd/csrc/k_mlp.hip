// Load predictor network on gfx950: a 3-layer MLP (ReLU) with BF16 weights and FP32
// accumulation, trained with masked MSE and Adam. Small and off the counting path's critical
// path; the kernels are plain LDS-tiled GEMMs (64x64 tiles, 4x4 per thread).
//
// Parameters (flat FP32 master copy, the BF16 copy is what the GEMMs read):
//   W1[n_in x h1] b1[h1] W2[h1 x h2] b2[h2] W3[h2 x n_out] b3[n_out]   (row-major, x @ W)
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rgkc_internal.h"

namespace rgkc {
namespace {

constexpr int TM = 64, TN = 64, TK = 16;

__device__ __forceinline__ float ld(const float* p, size_t i) { return p[i]; }
__device__ __forceinline__ float ld(const __hip_bfloat16* p, size_t i) { return __bfloat162float(p[i]); }

// C[M x N] (+)= op(A)[M x K] * op(B)[K x N]; A is FP32 activations/gradients, B is FP32 or BF16.
// op(A)(m, kk) = TA ? A[kk*lda + m] : A[m*lda + kk]; op(B)(kk, n) = TB ? B[n*ldb + kk] : B[kk*ldb + n].
// Epilogue: + bias[n] if bias, ReLU if relu; accumulate adds into C.
template <bool TA, bool TB, typename BT>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const float* __restrict__ A, int lda,
                                              const BT* __restrict__ B, int ldb, float* __restrict__ C, int ldc,
                                              const float* __restrict__ bias, int relu, int accumulate) {
    __shared__ float As[TK][TM + 1];
    __shared__ float Bs[TK][TN + 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    float acc[4][4] = {};
    for (int k0 = 0; k0 < K; k0 += TK) {
        for (int i = threadIdx.x; i < TM * TK; i += 256) {
            const int mm = i / TK, kk = i % TK;
            const int gm = m0 + mm, gk = k0 + kk;
            float v = 0.f;
            if (gm < M && gk < K) v = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
            As[kk][mm] = v;
        }
        for (int i = threadIdx.x; i < TN * TK; i += 256) {
            const int nn = i % TN, kk = i / TN;
            const int gn = n0 + nn, gk = k0 + kk;
            float v = 0.f;
            if (gn < N && gk < K) v = TB ? ld(B, (size_t)gn * ldb + gk) : ld(B, (size_t)gk * ldb + gn);
            Bs[kk][nn] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TK; kk++) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; i++) { a[i] = As[kk][ty * 4 + i]; b[i] = Bs[kk][tx * 4 + i]; }
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int gm = m0 + ty * 4 + i;
        if (gm >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int gn = n0 + tx * 4 + j;
            if (gn >= N) continue;
            float v = acc[i][j];
            if (bias) v += bias[gn];
            if (accumulate) v += C[(size_t)gm * ldc + gn];
            if (relu) v = v > 0.f ? v : 0.f;
            C[(size_t)gm * ldc + gn] = v;
        }
    }
}

// out[n] = sum_m G[m, n]
__global__ void k_colsum(int M, int N, const float* __restrict__ G, float* __restrict__ out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.f;
    for (int m = 0; m < M; m++) s += G[(size_t)m * N + n];
    out[n] = s;
}

// G[i] *= (H[i] > 0)  (ReLU backward on the stored activation)
__global__ void k_relu_bwd(size_t n, const float* __restrict__ H, float* __restrict__ G) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(H[i] > 0.f)) G[i] = 0.f;
}

// Masked MSE: loss = sum(mask * (Y - T)^2) / max(1, sum(mask)); dY = 2 mask (Y - T) / max(1, sum mask).
__global__ void k_mse(size_t n, const float* __restrict__ Y, const float* __restrict__ T, const float* __restrict__ mask,
                      float inv_cnt, float* __restrict__ dY, float* __restrict__ loss) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (i < n) {
        const float m = mask ? mask[i] : 1.f;
        const float d = (Y[i] - T[i]) * m;
        dY[i] = 2.f * d * inv_cnt;
        l = d * d * inv_cnt;
    }
    for (int o = 32; o > 0; o >>= 1) l += __shfl_down(l, o, 64);
    if ((threadIdx.x & 63) == 0 && l != 0.f) atomicAdd(loss, l);
}

// Adam on the FP32 master copy, then refresh the BF16 copy.
__global__ void k_adam(size_t n, float* __restrict__ w, __hip_bfloat16* __restrict__ wb, const float* __restrict__ g,
                       float* __restrict__ m1, float* __restrict__ m2, float lr, float b1, float b2, float eps,
                       float c1, float c2, float wd) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] + wd * w[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    const float mh = a / c1, vh = b / c2;
    const float nw = w[i] - lr * mh / (sqrtf(vh) + eps);
    w[i] = nw;
    wb[i] = __float2bfloat16(nw);
}

__global__ void k_to_bf16(size_t n, const float* __restrict__ w, __hip_bfloat16* __restrict__ wb) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wb[i] = __float2bfloat16(w[i]);
}

template <bool TA, bool TB, typename BT>
hipError_t gemm(int M, int N, int K, const float* A, int lda, const BT* B, int ldb, float* C, int ldc, const float* bias,
                bool relu, bool acc, hipStream_t s) {
    dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM);
    hipLaunchKernelGGL((k_gemm<TA, TB, BT>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, bias,
                       relu ? 1 : 0, acc ? 1 : 0);
    return hipGetLastError();
}

inline unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// ---- host-side operations on device buffers (called by the C ABI layer) --------------------
hipError_t mlp_to_bf16(const float* w, void* wb, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk(n)), dim3(256), 0, s, n, w, (__hip_bfloat16*)wb);
    return hipGetLastError();
}

// Forward: X[B x I] -> H1, H2 (post-ReLU) and Y[B x O].
hipError_t mlp_forward(const MlpDims& d, const void* wb, const float* w32, const float* X, int B, float* H1, float* H2,
                       float* Y, hipStream_t s) {
    const __hip_bfloat16* W = (const __hip_bfloat16*)wb;
    const __hip_bfloat16 *W1 = W + d.off_W1(), *W2 = W + d.off_W2(), *W3 = W + d.off_W3();
    const float *b1 = w32 + d.off_b1(), *b2 = w32 + d.off_b2(), *b3 = w32 + d.off_b3();
    hipError_t e = gemm<false, false>(B, d.h1, d.n_in, X, d.n_in, W1, d.h1, H1, d.h1, b1, true, false, s);
    if (e) return e;
    e = gemm<false, false>(B, d.h2, d.h1, H1, d.h1, W2, d.h2, H2, d.h2, b2, true, false, s);
    if (e) return e;
    return gemm<false, false>(B, d.n_out, d.h2, H2, d.h2, W3, d.n_out, Y, d.n_out, b3, false, false, s);
}

// Backward of masked MSE through the network; G (flat, same layout as the parameters) receives
// the gradients w.r.t. the BF16 weights and FP32 biases. dY/dH1/dH2 are scratch.
hipError_t mlp_backward(const MlpDims& d, const void* wb, const float* X, int B, const float* H1, const float* H2,
                        const float* Y, const float* T, const float* mask, float mask_sum, float* dY, float* dH1,
                        float* dH2, float* G, float* loss, hipStream_t s) {
    const __hip_bfloat16* W = (const __hip_bfloat16*)wb;
    const __hip_bfloat16 *W2 = W + d.off_W2(), *W3 = W + d.off_W3();
    float *gW1 = G + d.off_W1(), *gb1 = G + d.off_b1(), *gW2 = G + d.off_W2(), *gb2 = G + d.off_b2();
    float *gW3 = G + d.off_W3(), *gb3 = G + d.off_b3();
    const size_t nY = (size_t)B * d.n_out;
    hipError_t e = hipMemsetAsync(loss, 0, sizeof(float), s);
    if (e) return e;
    hipLaunchKernelGGL(k_mse, dim3(nblk(nY)), dim3(256), 0, s, nY, Y, T, mask, 1.f / (mask_sum > 1.f ? mask_sum : 1.f),
                       dY, loss);
    // layer 3: gW3 = H2^T dY ; gb3 = colsum dY ; dH2 = dY W3^T (masked by ReLU)
    if ((e = gemm<true, false>(d.h2, d.n_out, B, H2, d.h2, dY, d.n_out, gW3, d.n_out, nullptr, false, false, s))) return e;
    hipLaunchKernelGGL(k_colsum, dim3(nblk(d.n_out)), dim3(256), 0, s, B, d.n_out, dY, gb3);
    if ((e = gemm<false, true>(B, d.h2, d.n_out, dY, d.n_out, W3, d.n_out, dH2, d.h2, nullptr, false, false, s))) return e;
    hipLaunchKernelGGL(k_relu_bwd, dim3(nblk((size_t)B * d.h2)), dim3(256), 0, s, (size_t)B * d.h2, H2, dH2);
    // layer 2
    if ((e = gemm<true, false>(d.h1, d.h2, B, H1, d.h1, dH2, d.h2, gW2, d.h2, nullptr, false, false, s))) return e;
    hipLaunchKernelGGL(k_colsum, dim3(nblk(d.h2)), dim3(256), 0, s, B, d.h2, dH2, gb2);
    if ((e = gemm<false, true>(B, d.h1, d.h2, dH2, d.h2, W2, d.h2, dH1, d.h1, nullptr, false, false, s))) return e;
    hipLaunchKernelGGL(k_relu_bwd, dim3(nblk((size_t)B * d.h1)), dim3(256), 0, s, (size_t)B * d.h1, H1, dH1);
    // layer 1
    if ((e = gemm<true, false>(d.n_in, d.h1, B, X, d.n_in, dH1, d.h1, gW1, d.h1, nullptr, false, false, s))) return e;
    hipLaunchKernelGGL(k_colsum, dim3(nblk(d.h1)), dim3(256), 0, s, B, d.h1, dH1, gb1);
    return hipGetLastError();
}

hipError_t mlp_adam(size_t n, float* w, void* wb, const float* g, float* m1, float* m2, float lr, int step, float wd,
                    hipStream_t s) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(nblk(n)), dim3(256), 0, s, n, w, (__hip_bfloat16*)wb, g, m1, m2, lr, b1, b2, eps,
                       c1, c2, wd);
    return hipGetLastError();
}

}  // namespace rgkc
