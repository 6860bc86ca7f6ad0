// FP64 GEMM on the CDNA4 matrix cores (v_mfma_f64_16x16x4f64) and the
// small kernels of trace-correcting purification.
//
// MFMA f64 16x16x4 operand layout (one wave): lane l holds A[l % 16][l / 16]
// and B[l / 16][l % 16]; the four result registers r hold
// D[4 r + l / 16][l % 16].
#include <cstdint>

#include "qbx_common.h"
#include "qbx_linalg.h"

namespace qbx {
namespace {

constexpr int GT = 32;   // block tile (GT x GT outputs, four 16 x 16 wave tiles)
constexpr int GK = 16;   // K step (four MFMA steps)

typedef double dbl4 __attribute__((ext_vector_type(4)));

struct GemmF64 {
    const double* A;
    const double* B;
    const double* Z;
    double* C;
    int M, N, K, lda, ldb, ldz, ldc;
    double alpha, beta;
};

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_f64_kernel(GemmF64 g)
{
    __shared__ double As[GK][GT + 1];   // As[k][i] = op(A)[row0 + i][k0 + k]
    __shared__ double Bs[GK][GT + 1];   // Bs[k][j] = op(B)[k0 + k][col0 + j]
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;
    const int wr = w >> 1, wc = w & 1;
    // 1-D grid, the column tiles of one row panel adjacent: a tall operand
    // (M up to 2^31 rows) is read from HBM once and reused from L2
    const int ntile = (g.N + GT - 1) / GT;
    const int row0 = (int)(blockIdx.x / ntile) * GT, col0 = (int)(blockIdx.x % ntile) * GT;
    dbl4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < g.K; k0 += GK) {
        // 512 elements of each tile, two per thread, coalesced along the
        // operand's contiguous index
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int e = t + 256 * h;
            int i, k;
            if (TA) { k = e / GT; i = e % GT; } else { i = e / GK; k = e % GK; }
            const int gi = row0 + i, gk = k0 + k;
            double v = 0.0;
            if (gi < g.M && gk < g.K) v = TA ? g.A[(size_t)gk * g.lda + gi] : g.A[(size_t)gi * g.lda + gk];
            As[k][i] = v;
            int j;
            if (TB) { j = e / GK; k = e % GK; } else { k = e / GT; j = e % GT; }
            const int gj = col0 + j;
            const int gk2 = k0 + k;
            double u = 0.0;
            if (gj < g.N && gk2 < g.K) u = TB ? g.B[(size_t)gj * g.ldb + gk2] : g.B[(size_t)gk2 * g.ldb + gj];
            Bs[k][j] = u;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GK / 4; ++kk) {
            const double a = As[4 * kk + (lane >> 4)][wr * 16 + (lane & 15)];
            const double b = Bs[4 * kk + (lane >> 4)][wc * 16 + (lane & 15)];
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    const int j = col0 + wc * 16 + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = row0 + wr * 16 + 4 * r + (lane >> 4);
        if (i < g.M && j < g.N) {
            double v = g.alpha * acc[r];
            if (g.beta != 0.0) v += g.beta * g.Z[(size_t)i * g.ldz + j];
            g.C[(size_t)i * g.ldc + j] = v;
        }
    }
}

// out[(a nao + b) nmo + m] = C[a][m] C[b][m]: the orbital-pair columns of the
// opposite-spin Coulomb matrix
__global__ void pair_columns_kernel(const double* C, int nao, int nmo, size_t tot, double* out)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= tot) return;
    const int m = (int)(e % nmo);
    const size_t ab = e / nmo;
    const int a = (int)(ab / nao), b = (int)(ab % nao);
    out[e] = C[(size_t)a * nmo + m] * C[(size_t)b * nmo + m];
}

__global__ void gershgorin_kernel(const double* P, int n, double* lohi)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double r = 0.0;
    for (int j = 0; j < n; ++j)
        if (j != i) r += fabs(P[(size_t)i * n + j]);
    const double c = P[(size_t)i * n + i];
    lohi[2 * i] = c - r;
    lohi[2 * i + 1] = c + r;
}

__global__ void sp2_start_kernel(const double* F, int n, double lo, double hi, double* P)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * n) return;
    const size_t i = e / n, j = e % n;
    P[e] = ((i == j ? hi : 0.0) - F[e]) / (hi - lo);
}

__global__ __launch_bounds__(1024) void trace_frob_kernel(const double* P, int n, double* out)
{
    __shared__ double s_tr[16], s_fr[16];
    const size_t nn = (size_t)n * n;
    double tr = 0.0, fr = 0.0;
    for (size_t e = threadIdx.x; e < nn; e += 1024) {
        const double v = P[e];
        fr += v * v;
    }
    for (int i = threadIdx.x; i < n; i += 1024) tr += P[(size_t)i * n + i];
    for (int o = 32; o > 0; o >>= 1) {
        tr += __shfl_xor(tr, o);
        fr += __shfl_xor(fr, o);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_tr[w] = tr;
        s_fr[w] = fr;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < 16; ++k) {
            a += s_tr[k];
            b += s_fr[k];
        }
        out[0] = a;
        out[1] = b;
    }
}

}  // namespace

void gemm_f64(bool trans_a, bool trans_b, int M, int N, int K, double alpha, const double* A, int lda,
              const double* B, int ldb, double beta, const double* Z, int ldz, double* C, int ldc, hipStream_t s)
{
    if (M <= 0 || N <= 0) return;
    if (beta != 0.0 && !Z) throw QbxError(QBX_EINVAL, "GEMM addend missing");
    GemmF64 g{A, B, Z, C, M, N, K, lda, ldb, ldz, ldc, alpha, beta};
    const size_t tiles = (size_t)((N + GT - 1) / GT) * (size_t)((M + GT - 1) / GT);
    if (tiles > 0x7fffffffu) throw QbxError(QBX_EUNSUP, "GEMM too large for one grid");
    const dim3 grid((unsigned)tiles), block(256);
    if (!trans_a && !trans_b) hipLaunchKernelGGL((gemm_f64_kernel<false, false>), grid, block, 0, s, g);
    else if (trans_a && !trans_b) hipLaunchKernelGGL((gemm_f64_kernel<true, false>), grid, block, 0, s, g);
    else if (!trans_a && trans_b) hipLaunchKernelGGL((gemm_f64_kernel<false, true>), grid, block, 0, s, g);
    else hipLaunchKernelGGL((gemm_f64_kernel<true, true>), grid, block, 0, s, g);
    QBX_HIP(hipGetLastError());
}

void pair_columns(const double* C, int nao, int nmo, double* out, hipStream_t s)
{
    const size_t tot = (size_t)nao * nao * nmo;
    if (tot == 0) return;
    hipLaunchKernelGGL(pair_columns_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, C, nao, nmo, tot,
                       out);
    QBX_HIP(hipGetLastError());
}

void gershgorin(const double* P, int n, double* lohi, hipStream_t s)
{
    hipLaunchKernelGGL(gershgorin_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, P, n, lohi);
    QBX_HIP(hipGetLastError());
}

void sp2_start(const double* F, int n, double lo, double hi, double* P, hipStream_t s)
{
    const size_t nn = (size_t)n * n;
    hipLaunchKernelGGL(sp2_start_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, F, n, lo, hi, P);
    QBX_HIP(hipGetLastError());
}

void trace_and_frobenius(const double* P, int n, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(trace_frob_kernel, dim3(1), dim3(1024), 0, s, P, n, out);
    QBX_HIP(hipGetLastError());
}

}  // namespace qbx

extern "C" {

int qbx_gemm_f64(int32_t trans_a, int32_t trans_b, int32_t M, int32_t N, int32_t K, double alpha, const double* A,
                 int32_t lda, const double* B, int32_t ldb, double beta, double* C, int32_t ldc, void* stream)
{
    return qbx::guarded([&] {
        if (M < 0 || N < 0 || K < 0) throw qbx::QbxError(QBX_EINVAL, "negative GEMM size");
        if (!A || !B || !C) throw qbx::QbxError(QBX_EINVAL, "null GEMM operand");
        if ((trans_a ? lda < M : lda < K) || (trans_b ? ldb < K : ldb < N) || ldc < N)
            throw qbx::QbxError(QBX_EINVAL, "leading dimension too small");
        qbx::gemm_f64(trans_a != 0, trans_b != 0, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, C, ldc,
                      (hipStream_t)stream);
    });
}

}  // extern "C"
