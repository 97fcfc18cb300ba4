// mlp_kernels.hip -- the multilayer perceptron of the learned control
// variate: GEMMs on BF16 MFMA with FP32 accumulation (weights and
// activations kept in FP32, rounded to BF16 as they are loaded), exact-erf
// GELU, mean-squared-error gradient, column sums and Adam on FP32 master
// weights.  All kernels are deterministic (no atomics).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mbs {

typedef float mlp_f32x16 __attribute__((ext_vector_type(16)));
typedef short mlp_bf16x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (finite inputs).
__device__ __forceinline__ short to_bf16(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

// C[M x N] = beta * C + op(A) op(B) (+ bias[N]); op(A) is M x K, op(B) K x N.
//   TA = 0: A row-major [M][lda]; TA = 1: A stored [K][lda] (op(A) = A^T)
//   TB = 0: B row-major [K][ldb]; TB = 1: B stored [N][ldb] (op(B) = B^T)
// Workgroup = 4 waves = a 64 x 64 tile of C, each wave 32 x 32 on
// v_mfma_f32_32x32x8bf16_1k: lane l supplies A[l%32][4(l/32) + q] and
// B[4(l/32) + q][l%32], q = 0..3, for each K step of 8.  Operands outside
// the matrices read as zero; only in-range C elements are written.
template <int TA, int TB>
__global__ void __launch_bounds__(256) mlp_gemm_kernel(int M, int N, int K,
                                                       const float *__restrict__ A, int lda,
                                                       const float *__restrict__ B, int ldb,
                                                       float *__restrict__ C, int ldc,
                                                       const float *__restrict__ bias, int beta)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int row0 = blockIdx.y * 64 + (w >> 1) * 32;
    const int col0 = blockIdx.x * 64 + (w & 1) * 32;
    const int ar = row0 + (lane & 31);  // this lane's A row
    const int bc = col0 + (lane & 31);  // this lane's B column
    const int kq = 4 * (lane >> 5);
    mlp_f32x16 acc = {};
    for (int k0 = 0; k0 < K; k0 += 8) {
        mlp_bf16x4 av, bv;
        for (int q = 0; q < 4; ++q) {
            const int kk = k0 + kq + q;
            float a = 0.0f, b = 0.0f;
            if (ar < M && kk < K)
                a = TA ? A[(size_t)kk * lda + ar] : A[(size_t)ar * lda + kk];
            if (bc < N && kk < K)
                b = TB ? B[(size_t)bc * ldb + kk] : B[(size_t)kk * ldb + bc];
            av[q] = to_bf16(a);
            bv[q] = to_bf16(b);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
    }
    const int c = col0 + (lane & 31);
    if (c >= N)
        return;
    const float bb = bias ? bias[c] : 0.0f;
    for (int v = 0; v < 16; ++v) {
        const int r = row0 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        if (r < M) {
            float *dst = C + (size_t)r * ldc + c;
            *dst = (beta ? *dst : 0.0f) + acc[v] + bb;
        }
    }
}

// GELU(z) = z Phi(z), Phi the standard normal distribution function.
__global__ void mlp_gelu_kernel(size_t n, const float *__restrict__ z, float *__restrict__ h)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float x = z[i];
        h[i] = 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
    }
}

// dZ = dH * GELU'(z), GELU'(z) = Phi(z) + z phi(z).
__global__ void mlp_gelu_grad_kernel(size_t n, const float *__restrict__ z,
                                     const float *__restrict__ dh, float *__restrict__ dz)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float x = z[i];
        const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
        const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
        dz[i] = dh[i] * (cdf + x * pdf);
    }
}

// dY = 2 (Y - T) / (rows * cols); per-row partial sums of (Y - T)^2.
__global__ void mlp_mse_grad_kernel(int rows, int cols, const float *__restrict__ y,
                                    const float *__restrict__ t, float *__restrict__ dy,
                                    float *__restrict__ row_sq)
{
    const int r = blockIdx.x;
    const float scale = 2.0f / ((float)rows * (float)cols);
    float s = 0.0f;
    for (int c = threadIdx.x; c < cols; c += blockDim.x) {
        const size_t i = (size_t)r * cols + c;
        const float d = y[i] - t[i];
        dy[i] = scale * d;
        s += d * d;
    }
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        row_sq[r] = red[0];
}

// out[c] = sum over rows of m[r][c], rows in order.
__global__ void mlp_colsum_kernel(int rows, int cols, const float *__restrict__ m,
                                  float *__restrict__ out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols)
        return;
    float s = 0.0f;
    for (int r = 0; r < rows; ++r)
        s += m[(size_t)r * cols + c];
    out[c] = s;
}

// Adam (bias-corrected), FP32 master weights.
__global__ void mlp_adam_kernel(size_t n, float *__restrict__ w, const float *__restrict__ g,
                                float *__restrict__ m, float *__restrict__ v, float lr,
                                float b1, float b2, float eps, float c1, float c2)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Gathers rows idx[0..rows) of src [*, cols] into dst [rows, cols].
__global__ void mlp_gather_kernel(int rows, int cols, const int32_t *__restrict__ idx,
                                  const float *__restrict__ src, float *__restrict__ dst)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * cols)
        return;
    const int r = (int)(i / cols), c = (int)(i - (size_t)r * cols);
    dst[i] = src[(size_t)idx[r] * cols + c];
}

}  // namespace mbs
