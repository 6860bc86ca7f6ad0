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

// Block of four waves in a 2 x 2 arrangement; each wave owns WT x WT tiles of
// 16 x 16 outputs, so the block tile is 32 WT square.  WT = 2 reads half the
// LDS bytes per MFMA of WT = 1 and is taken where the grid still has several
// blocks per CU.
template <bool TA, bool TB, int WT>
__global__ __launch_bounds__(256) void gemm_f64_kernel(GemmF64 g)
{
    constexpr int BT = 32 * WT;
    constexpr int LOADS = GK * BT / 256;
    __shared__ double As[GK][BT + 1];   // As[k][i] = op(A)[row0 + i][k0 + k]
    __shared__ double Bs[GK][BT + 1];   // Bs[k][j] = op(B)[k0 + k][col0 + j]
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;
    const int wr = w >> 1, wc = w & 1;
    // 1-D grid, the column tiles of one row panel adjacent: a tall operand
    // (M up to 2^31 rows) is read from HBM once and reused from L2
    const int ntile = (g.N + BT - 1) / BT;
    const int row0 = (int)(blockIdx.x / ntile) * BT, col0 = (int)(blockIdx.x % ntile) * BT;
    dbl4 acc[WT][WT];
#pragma unroll
    for (int a = 0; a < WT; ++a)
#pragma unroll
        for (int b = 0; b < WT; ++b) acc[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
    // tile loads: GK x BT elements of each operand, LOADS per thread,
    // coalesced along the operand's contiguous index; the next K step's loads
    // are issued before the current step's MFMAs
    double ra[LOADS], rb[LOADS];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int h = 0; h < LOADS; ++h) {
            const int e = t + 256 * h;
            int i, k;
            if (TA) { k = e / BT; i = e % BT; } else { i = e / GK; k = e % GK; }
            const int gi = row0 + i, gk = k0 + k;
            ra[h] = (gi < g.M && gk < g.K) ? (TA ? g.A[(size_t)gk * g.lda + gi] : g.A[(size_t)gi * g.lda + gk]) : 0.0;
            int j;
            if (TB) { j = e / GK; k = e % GK; } else { k = e / BT; j = e % BT; }
            const int gj = col0 + j, gk2 = k0 + k;
            rb[h] = (gj < g.N && gk2 < g.K) ? (TB ? g.B[(size_t)gj * g.ldb + gk2] : g.B[(size_t)gk2 * g.ldb + gj]) : 0.0;
        }
    };
    fetch(0);
    for (int k0 = 0; k0 < g.K; k0 += GK) {
#pragma unroll
        for (int h = 0; h < LOADS; ++h) {
            const int e = t + 256 * h;
            int i, k;
            if (TA) { k = e / BT; i = e % BT; } else { i = e / GK; k = e % GK; }
            As[k][i] = ra[h];
            int j;
            if (TB) { j = e / GK; k = e % GK; } else { k = e / BT; j = e % BT; }
            Bs[k][j] = rb[h];
        }
        __syncthreads();
        if (k0 + GK < g.K) fetch(k0 + GK);
#pragma unroll
        for (int kk = 0; kk < GK / 4; ++kk) {
            double a[WT], b[WT];
#pragma unroll
            for (int m = 0; m < WT; ++m) {
                a[m] = As[4 * kk + (lane >> 4)][(wr * WT + m) * 16 + (lane & 15)];
                b[m] = Bs[4 * kk + (lane >> 4)][(wc * WT + m) * 16 + (lane & 15)];
            }
#pragma unroll
            for (int m = 0; m < WT; ++m)
#pragma unroll
                for (int n = 0; n < WT; ++n) acc[m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[m], b[n], acc[m][n], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < WT; ++m)
#pragma unroll
        for (int n = 0; n < WT; ++n) {
            const int j = col0 + (wc * WT + n) * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = row0 + (wr * WT + m) * 16 + 4 * r + (lane >> 4);
                if (i < g.M && j < g.N) {
                    double v = g.alpha * acc[m][n][r];
                    if (g.beta != 0.0) v += g.beta * g.Z[(size_t)i * g.ldz + j];
                    g.C[(size_t)i * g.ldc + j] = v;
                }
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

// P <- (P + P^T) / 2 in place, one thread per (i, j) with i < j
__global__ void symmetrize_square_kernel(double* P, int n)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * n) return;
    const size_t i = e / n, j = e % n;
    if (i >= j) return;
    const double v = 0.5 * (P[i * n + j] + P[j * n + i]);
    P[i * n + j] = v;
    P[j * n + i] = v;
}

// column signs: sgn[j] = -1 where C[0][j] < 0, else 1
__global__ void column_sign_kernel(const double* C, int n, double* sgn)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) sgn[j] = C[j] < 0.0 ? -1.0 : 1.0;
}

__global__ void scale_columns_kernel(double* C, int n, const double* sgn)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * n) return;
    C[e] *= sgn[e % n];
}

// out[0] = min_i lohi[2i], out[1] = max_i lohi[2i+1]: the spectral interval
// from the rows' Gershgorin intervals (one block; min and max are exact in
// any order)
__global__ __launch_bounds__(1024) void interval_kernel(const double* lohi, int n, double* out)
{
    __shared__ double s_lo[16], s_hi[16];
    double lo = lohi[0], hi = lohi[1];
    for (int i = threadIdx.x; i < n; i += 1024) {
        lo = fmin(lo, lohi[2 * i]);
        hi = fmax(hi, lohi[2 * i + 1]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, o));
        hi = fmax(hi, __shfl_xor(hi, o));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_lo[w] = lo;
        s_hi[w] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 16; ++k) {
            lo = fmin(lo, s_lo[k]);
            hi = fmax(hi, s_hi[k]);
        }
        out[0] = fmin(lo, s_lo[0]);
        out[1] = fmax(hi, s_hi[0]);
    }
}

}  // namespace

void gemm_f64(bool trans_a, bool trans_b, int M, int N, int K, double alpha, const double* A, int lda,
              const double* B, int ldb, double beta, const double* Z, int ldz, double* C, int ldc, hipStream_t s)
{
    if (M <= 0 || N <= 0) return;
    if (beta != 0.0 && !Z) throw QbxError(QBX_EINVAL, "GEMM addend missing");
    GemmF64 g{A, B, Z, C, M, N, K, lda, ldb, ldz, ldc, alpha, beta};
    // 64 x 64 blocks where they still give 1024 blocks or more (four per CU),
    // else 32 x 32 blocks
    const size_t big = (size_t)((N + 63) / 64) * (size_t)((M + 63) / 64);
    const int bt = big >= 1024 ? 64 : 32;
    const size_t tiles = (size_t)((N + bt - 1) / bt) * (size_t)((M + bt - 1) / bt);
    if (tiles > 0x7fffffffu) throw QbxError(QBX_EUNSUP, "GEMM too large for one grid");
    const dim3 grid((unsigned)tiles), block(256);
    auto launch = [&](void (*kern)(GemmF64)) {
        void* args[] = {&g};
        QBX_HIP(hipLaunchKernel((const void*)kern, grid, block, args, 0, s));
    };
    if (bt == 64) {
        if (!trans_a && !trans_b) launch(gemm_f64_kernel<false, false, 2>);
        else if (trans_a && !trans_b) launch(gemm_f64_kernel<true, false, 2>);
        else if (!trans_a && trans_b) launch(gemm_f64_kernel<false, true, 2>);
        else launch(gemm_f64_kernel<true, true, 2>);
    } else {
        if (!trans_a && !trans_b) launch(gemm_f64_kernel<false, false, 1>);
        else if (trans_a && !trans_b) launch(gemm_f64_kernel<true, false, 1>);
        else if (!trans_a && trans_b) launch(gemm_f64_kernel<false, true, 1>);
        else launch(gemm_f64_kernel<true, true, 1>);
    }
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

void symmetrize_square(double* P, int n, hipStream_t s)
{
    const size_t nn = (size_t)n * n;
    hipLaunchKernelGGL(symmetrize_square_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, P, n);
    QBX_HIP(hipGetLastError());
}

void fix_column_signs(double* C, int n, double* sgn, hipStream_t s)
{
    const size_t nn = (size_t)n * n;
    hipLaunchKernelGGL(column_sign_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, C, n, sgn);
    QBX_HIP(hipGetLastError());
    hipLaunchKernelGGL(scale_columns_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, C, n, sgn);
    QBX_HIP(hipGetLastError());
}

void spectral_interval(const double* lohi, int n, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(interval_kernel, dim3(1), dim3(1024), 0, s, lohi, n, out);
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
