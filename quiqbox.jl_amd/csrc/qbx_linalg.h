// Dense FP64 linear algebra on the device for the SCF step: a GEMM on the
// FP64 matrix cores and the pieces of trace-correcting purification.
// Row-major matrices, device pointers, launches on the given stream.
#pragma once
#include <hip/hip_runtime.h>

namespace qbx {

// C = alpha * op(A) op(B) + beta * Z   (op = transpose when the flag is set;
// op(A) is M x K, op(B) is K x N; Z may be C itself, or null when beta == 0)
void gemm_f64(bool trans_a, bool trans_b, int M, int N, int K, double alpha, const double* A, int lda,
              const double* B, int ldb, double beta, const double* Z, int ldz, double* C, int ldc, hipStream_t s);

// out[(a nao + b) nmo + m] = C[a][m] C[b][m] for the nao x nmo matrix C
void pair_columns(const double* C, int nao, int nmo, double* out, hipStream_t s);

// Gershgorin interval of each row of the symmetric n x n matrix P:
// lohi[2i] = P_ii - sum_j!=i |P_ij|, lohi[2i+1] = P_ii + sum_j!=i |P_ij|
void gershgorin(const double* P, int n, double* lohi, hipStream_t s);

// P <- (P + P^T) / 2 for the n x n matrix P
void symmetrize_square(double* P, int n, hipStream_t s);

// flip the columns j of the n x n matrix C whose first-row entry C[0][j] is
// negative (sgn: n doubles of work space)
void fix_column_signs(double* C, int n, double* sgn, hipStream_t s);

// out[0] = min of lohi[2i], out[1] = max of lohi[2i+1] over the n rows
void spectral_interval(const double* lohi, int n, double* out, hipStream_t s);

// P = (hi I - F) / (hi - lo)
void sp2_start(const double* F, int n, double lo, double hi, double* P, hipStream_t s);

// out[0] = trace(P), out[1] = sum_ij P_ij^2 (= trace(P^2) for symmetric P);
// one block, fixed summation order
void trace_and_frobenius(const double* P, int n, double* out, hipStream_t s);

}  // namespace qbx
