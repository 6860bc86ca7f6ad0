// qbx_mlp.hip -- BF16 matrix-core kernels for the learned SCF initial guess
// (a pair-feature MLP): an MFMA GEMM with FP32 accumulation and the
// elementwise pieces of its forward and backward passes.  All pointers are
// device pointers; every call is enqueued on the given stream (NULL: the
// default stream) and returns without synchronizing.
//
// Storage: matrices row-major; BF16 values as raw 16-bit patterns (the upper
// half of an IEEE float, round-to-nearest-even on conversion).
#include "qbx.h"
#include "qbx_common.h"

#include <hip/hip_runtime.h>
#include <cstdint>

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ unsigned short f2bf(float x)
{
    unsigned u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);   // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

struct GemmArgs {
    const unsigned short* A;
    const unsigned short* B;
    float* C;
    const float* bias;
    unsigned short* Cb;   // BF16 copy of the output (epilogue EPI_BIAS_RELU_BF16)
    int M, N, K, lda, ldb, ldc;
    float alpha, beta;
};

enum { EPI_STORE = 0, EPI_BIAS_RELU_BF16 = 1, EPI_BIAS = 2 };

// One wavefront computes one 32x32 tile of C = alpha op(A) op(B) + beta C with
// v_mfma_f32_32x32x8_bf16: per instruction lane l supplies A[m0 + l%32][k0 +
// 4(l/32) + j] and B[k0 + 4(l/32) + j][n0 + l%32] (j = 0..3) and holds the
// accumulators D[m0 + 8(i/4) + 4(l/32) + i%4][n0 + l%32], i = 0..15.
// op(A) = A (M x K, row stride lda) or A^T (A stored K x M); likewise B.
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs g)
{
    const int lane = (int)threadIdx.x & 63;
    const int wave = (int)blockIdx.x * 4 + ((int)threadIdx.x >> 6);
    const int tiles_n = (g.N + 31) / 32;
    const int tm = wave / tiles_n, tn = wave % tiles_n;
    if (tm * 32 >= g.M) return;
    const int r = lane & 31, kq = lane >> 5;
    const int m = tm * 32 + r, n = tn * 32 + r;
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
    for (int k0 = 0; k0 < g.K; k0 += 8) {
        bf16x4 af, bfr;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = k0 + 4 * kq + j;
            unsigned short av = 0, bv = 0;
            if (m < g.M && k < g.K) av = TA ? g.A[(size_t)k * g.lda + m] : g.A[(size_t)m * g.lda + k];
            if (n < g.N && k < g.K) bv = TB ? g.B[(size_t)n * g.ldb + k] : g.B[(size_t)k * g.ldb + n];
            af[j] = (short)av;
            bfr[j] = (short)bv;
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(af, bfr, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int row = tm * 32 + 8 * (i / 4) + 4 * kq + (i % 4);
        if (row >= g.M || n >= g.N) continue;
        const size_t o = (size_t)row * g.ldc + n;
        float v = g.alpha * acc[i];
        if (EPI == EPI_STORE) {
            if (g.beta != 0.0f) v += g.beta * g.C[o];
            g.C[o] = v;
        } else {
            v += g.bias[n];
            if (EPI == EPI_BIAS_RELU_BF16) {
                v = v > 0.0f ? v : 0.0f;
                g.Cb[o] = f2bf(v);
            }
            g.C[o] = v;
        }
    }
}

// elementwise operations
__global__ void to_bf16_kernel(const float* x, unsigned short* y, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// dZ = dH * [H > 0] (ReLU backward), as BF16 for the next GEMMs
__global__ void relu_back_kernel(const float* dH, const unsigned short* H, unsigned short* dZb, float* dZ, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = bf2f(H[i]) > 0.0f ? dH[i] : 0.0f;
    dZ[i] = v;
    dZb[i] = f2bf(v);
}

// weighted squared error: loss_i = w_r (Y - T)^2, dY = 2 w_r (Y - T) * scale, with the
// per-row weight w_r (0 for padding rows and masked columns via the mask)
__global__ void mse_kernel(const float* Y, const float* T, const float* rw, const float* cmask, int rows, int cols,
                           float scale, float* dY, unsigned short* dYb, float* loss)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)rows * cols) return;
    const int rr = (int)(i / cols), cc = (int)(i % cols);
    const float w = rw[rr] * cmask[cc];
    const float d = Y[i] - T[i];
    const float gval = 2.0f * w * d * scale;
    dY[i] = gval;
    dYb[i] = f2bf(gval);
    atomicAdd(loss, w * d * d * scale);
}

// column sums of an FP32 rows x cols matrix (bias gradients)
__global__ void colsum_kernel(const float* X, int rows, int cols, float* out)
{
    const int c = (int)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.0f;
    for (int r = 0; r < rows; ++r) s += X[(size_t)r * cols + c];
    out[c] = s;
}

// Adam (Kingma & Ba, 2015) on FP32 master weights; writes the BF16 copy
__global__ void adam_kernel(float* w, float* mo, float* ve, const float* gr, unsigned short* wb, int64_t n, float lr,
                            float b1, float b2, float eps, float c1, float c2)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = gr[i];
    const float m1 = b1 * mo[i] + (1.0f - b1) * gi;
    const float v1 = b2 * ve[i] + (1.0f - b2) * gi * gi;
    mo[i] = m1;
    ve[i] = v1;
    const float wi = w[i] - lr * (m1 / c1) / (sqrtf(v1 / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

extern "C" {

int qbx_gemm_bf16(int32_t trans_a, int32_t trans_b, int32_t M, int32_t N, int32_t K, float alpha, const void* A,
                  int32_t lda, const void* B, int32_t ldb, float beta, float* C, int32_t ldc, int32_t epilogue,
                  const float* bias, void* Cb, void* stream)
{
    return qbx::guarded([&] {
        if (M < 0 || N < 0 || K < 0) throw qbx::QbxError(QBX_EINVAL, "negative GEMM size");
        if (!A || !B || !C) throw qbx::QbxError(QBX_EINVAL, "null GEMM operand");
        if (epilogue < 0 || epilogue > 2) throw qbx::QbxError(QBX_EINVAL, "unknown GEMM epilogue");
        if (epilogue != EPI_STORE && !bias) throw qbx::QbxError(QBX_EINVAL, "the epilogue needs a bias");
        if (epilogue == EPI_BIAS_RELU_BF16 && !Cb) throw qbx::QbxError(QBX_EINVAL, "the epilogue needs a BF16 output");
        if ((trans_a ? lda < M : lda < K) || (trans_b ? ldb < K : ldb < N) || ldc < N)
            throw qbx::QbxError(QBX_EINVAL, "leading dimension too small");
        if (M == 0 || N == 0) return;
        GemmArgs g;
        g.A = (const unsigned short*)A;
        g.B = (const unsigned short*)B;
        g.C = C;
        g.bias = bias;
        g.Cb = (unsigned short*)Cb;
        g.M = M;
        g.N = N;
        g.K = K;
        g.lda = lda;
        g.ldb = ldb;
        g.ldc = ldc;
        g.alpha = alpha;
        g.beta = beta;
        const int64_t waves = (int64_t)((M + 31) / 32) * ((N + 31) / 32);
        const dim3 grid((unsigned)((waves + 3) / 4)), block(256);
        hipStream_t st = (hipStream_t)stream;
#define QBX_GEMM_CASE(TA, TB, E) \
    if (!!trans_a == TA && !!trans_b == TB && epilogue == E) \
        hipLaunchKernelGGL((gemm_bf16_kernel<TA, TB, E>), grid, block, 0, st, g);
        QBX_GEMM_CASE(false, false, 0) QBX_GEMM_CASE(false, false, 1) QBX_GEMM_CASE(false, false, 2)
        QBX_GEMM_CASE(true, false, 0) QBX_GEMM_CASE(false, true, 0)
        QBX_GEMM_CASE(true, true, 0)
#undef QBX_GEMM_CASE
        if (epilogue != EPI_STORE && (trans_a || trans_b))
            throw qbx::QbxError(QBX_EUNSUP, "bias epilogues take untransposed operands");
        QBX_HIP(hipGetLastError());
    });
}

int qbx_mlp_elementwise(int32_t op, int64_t n, const void* x, const void* y, void* out0, void* out1, void* stream)
{
    // op 0: out0(bf16) = bf16(x f32)
    // op 1: ReLU backward: x = dH (f32), y = H (bf16) -> out1 = dZ (f32), out0 = dZ (bf16)
    return qbx::guarded([&] {
        if (n < 0) throw qbx::QbxError(QBX_EINVAL, "negative size");
        if (n == 0) return;
        hipStream_t st = (hipStream_t)stream;
        if (op == 0) {
            hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(n)), dim3(256), 0, st, (const float*)x,
                               (unsigned short*)out0, n);
        } else if (op == 1) {
            hipLaunchKernelGGL(relu_back_kernel, dim3(blocks_for(n)), dim3(256), 0, st, (const float*)x,
                               (const unsigned short*)y, (unsigned short*)out0, (float*)out1, n);
        } else {
            throw qbx::QbxError(QBX_EINVAL, "unknown elementwise op");
        }
        QBX_HIP(hipGetLastError());
    });
}

int qbx_mlp_mse(int32_t rows, int32_t cols, const float* Y, const float* T, const float* row_w, const float* col_mask,
                float scale, float* dY, void* dYb, float* loss, void* stream)
{
    return qbx::guarded([&] {
        if (rows < 0 || cols < 0) throw qbx::QbxError(QBX_EINVAL, "negative size");
        const int64_t n = (int64_t)rows * cols;
        if (n == 0) return;
        hipLaunchKernelGGL(mse_kernel, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, Y, T, row_w, col_mask,
                           rows, cols, scale, dY, (unsigned short*)dYb, loss);
        QBX_HIP(hipGetLastError());
    });
}

int qbx_mlp_colsum(int32_t rows, int32_t cols, const float* X, float* out, void* stream)
{
    return qbx::guarded([&] {
        if (rows < 0 || cols < 0) throw qbx::QbxError(QBX_EINVAL, "negative size");
        if (cols == 0) return;
        hipLaunchKernelGGL(colsum_kernel, dim3(blocks_for(cols)), dim3(256), 0, (hipStream_t)stream, X, rows, cols,
                           out);
        QBX_HIP(hipGetLastError());
    });
}

int qbx_mlp_adam(int64_t n, float* w, float* m, float* v, const float* g, void* wb, float lr, float b1, float b2,
                 float eps, int32_t step, void* stream)
{
    return qbx::guarded([&] {
        if (n < 0 || step < 1) throw qbx::QbxError(QBX_EINVAL, "bad Adam arguments");
        if (n == 0) return;
        const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
        hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, w, m, v, g,
                           (unsigned short*)wb, n, lr, b1, b2, eps, c1, c2);
        QBX_HIP(hipGetLastError());
    });
}

}  // extern "C"
