/*
 * qbx.h -- C-ABI of libqbx_hip.so, the MI355X (gfx950) Gaussian-integral and
 * Fock-build engine that sits under Quiqbox.jl's integral entry points and
 * runHartreeFock.
 *
 * Plain C types only.  Every call is synchronous on return (one HIP stream per
 * context) unless its name ends in _device, which enqueues on the context's
 * stream, or on the caller's stream when one is given, and returns
 * immediately.  Return value: 0 on success, a negative QBX_E* code on failure;
 * the message is available from qbx_last_error() (per calling thread).
 * Indices are 0-based.  Square matrices are n_bf*n_bf and symmetric, so
 * row- and column-major layouts coincide.  The ERI tensor is n_bf^4 with
 * eri[((i*n+j)*n+k)*n+l] = (ij|kl) in chemists' notation; being 8-fold
 * symmetric it reads the same in Julia's column-major order.
 *
 * Reference interfaces cited as file:line of Quiqbox.jl v0.6.3.
 */
#ifndef QBX_H
#define QBX_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QBX_OK 0
#define QBX_EINVAL (-1)   /* bad argument (Julia: DomainError / DimensionMismatch) */
#define QBX_EHIP (-2)     /* HIP runtime failure                                */
#define QBX_ENOMEM (-3)   /* allocation failure                                 */
#define QBX_EUNSUP (-4)   /* unsupported basis (angular momentum above the build) */
#define QBX_EINTERNAL (-5)
#define QBX_ENOCONV (-6)  /* an iteration did not settle (qbx_density_purify)  */

/* Shells up to g (L = 4): s, p and d quartets run the generated class
   kernels, quartets with an f or g shell a generic kernel (always FP64).
   QBX_SHELL_MAXC: Cartesian components of the largest shell. */
#define QBX_MAX_L 4
#define QBX_SHELL_MAXC 15

#define QBX_PREC_FP64 0   /* every quartet in FP64                              */
#define QBX_PREC_MIXED 1  /* FP32 primitive arithmetic for quartets whose Schwarz
                             bound is below the FP64 cut, FP64 for the rest     */

typedef struct qbx_ctx qbx_ctx;
typedef struct qbx_basis qbx_basis;

/* Context: one GPU, one stream.  No reference counterpart (the reference has no
   device state); the Julia glue keeps one per process. */
int  qbx_init(int32_t device, qbx_ctx** out);
void qbx_finalize(qbx_ctx* ctx);

/* Flattened basis set.  Replaces the walk from MultiOrbitalData
   (OrbitalBases.jl:439-468) through getOrbCorePointers (Framework.jl:784-817)
   that every integral pipeline performs in computeOrbDataIntegral
   (Framework.jl:895-916).
   prim_cen  3*n_prim  primitive centres, xyz interleaved
   prim_xpn  n_prim    exponents
   prim_ang  3*n_prim  Cartesian exponents (i,j,k)
   bf_ptr    n_bf+1    CSR row pointer of basis functions over primitives
   bf_prim   bf_ptr[n_bf]  primitive index of each term
   bf_w      bf_ptr[n_bf]  final contraction weight of each term (coefficient x
                           normalization, Framework.jl:748-773)
   All primitives of one basis function must share centre and (i,j,k). */
int  qbx_basis_create(qbx_ctx* ctx, int32_t n_prim, const double* prim_cen,
                      const double* prim_xpn, const int32_t* prim_ang,
                      int32_t n_bf, const int32_t* bf_ptr, const int32_t* bf_prim,
                      const double* bf_w, qbx_basis** out);
void qbx_basis_destroy(qbx_basis* basis);
/* shell / shell-pair counts of a basis (diagnostics) */
int  qbx_basis_info(const qbx_basis* basis, int32_t* n_bf, int32_t* n_shell,
                    int32_t* n_pair, int32_t* max_l);

/* overlaps(bs)            Interface.jl:46-54   -> computeOrbDataIntegral(OneBody, Overlap) */
int  qbx_overlap(qbx_ctx* ctx, const qbx_basis* basis, double* S);
/* elecKinetics(bs)        Interface.jl:143-153 (kinetic sampler Samplers.jl:81-107) */
int  qbx_kinetic(qbx_ctx* ctx, const qbx_basis* basis, double* T);
/* nucAttractions(nuc, coords, bs)  Interface.jl:236-256; charges Z (n_nuc), R 3*n_nuc;
   test charge -1 as CoulombMultiPointPotential (Samplers.jl:56-61) */
int  qbx_nuclear(qbx_ctx* ctx, const qbx_basis* basis, int32_t n_nuc, const double* Z,
                 const double* R, double* V);
/* Own-nucleus attraction (no reference counterpart; the learned guess's pair
   features): V[i][j] = <i| -Z_n / |r - R_n| |j> for the nucleus n at the centre
   of function i alone, or 0 where no nucleus sits there (centres match within
   1e-10 bohr).  V is not symmetric: V^T holds the column function's nucleus.
   One call replaces one qbx_nuclear call per nucleus. */
int  qbx_nuclear_own(qbx_ctx* ctx, const qbx_basis* basis, int32_t n_nuc, const double* Z,
                     const double* R, double* V);
/* coreHamiltonian(nuc, coords, bs)  Interface.jl:288-312 = kinetic + nuclear */
int  qbx_core_hamiltonian(qbx_ctx* ctx, const qbx_basis* basis, int32_t n_nuc,
                          const double* Z, const double* R, double* H);
/* multipoleMoments(center, degree, bs)  Interface.jl:71-105 */
int  qbx_multipole(qbx_ctx* ctx, const qbx_basis* basis, const double* center,
                   const int32_t* degree, double* M);
/* elecRepulsions(bs)  Interface.jl:356-365 -> computeOrbDataIntegral(TwoBody, Coulomb)
   (Framework.jl:907-916) dense branch (Framework.jl:640-666).  Limited to
   n_bf <= 256 (the tensor is n_bf^4 doubles). */
int  qbx_eri_dense(qbx_ctx* ctx, const qbx_basis* basis, double* eri);

/* changeOrbitalBasis(DbodyInt, C)  Interface.jl:376-380: one- or two-body
   integrals in the orbital basis of the n_ao x n_mo coefficient matrix C,
     rank 2: out[i][j]       = sum_ab   ints[a][b] C[a][i] C[b][j]
     rank 4: out[i][j][k][l] = sum_abcd ints[a][b][c][d] C[a][i] C[b][j] C[c][k] C[d][l]
   as `rank` successive quarter transformations, each one FP64 matrix-core
   GEMM over the device copy (two device buffers of the largest intermediate,
   n_ao^3 max(n_ao, n_mo) doubles for rank 4).  Row-major host arrays; the
   tensor needs no index symmetry. */
int  qbx_mo_transform(qbx_ctx* ctx, int32_t rank, int32_t n_ao, int32_t n_mo, const double* ints,
                      const double* C, double* out);
/* getJab  Interface.jl:382-393 (third result of changeOrbitalBasis with two
   coefficient matrices, Interface.jl:405-407): the Coulomb matrix between
   orbitals of different spin,
     Jab[m][n] = sum_abcd eri[a][b][c][d] C1[a][m] C1[b][m] C2[c][n] C2[d][n],
   C1 n_ao x n1, C2 n_ao x n2, as two GEMMs over the orbital-pair columns
   C1[a][m] C1[b][m] and C2[c][n] C2[d][n]. */
int  qbx_mo_coulomb_ab(qbx_ctx* ctx, int32_t n_ao, int32_t n1, int32_t n2, const double* eri,
                       const double* C1, const double* C2, double* Jab);

/* Fock two-electron part: getGcore (HartreeFock.jl:305-319) for n_set density
   pairs at once.  For each set s:
     G[s] = J[DJ[s]] - K[DK[s]],
     J[D]_{mn} = sum_{ls} D_{sl} (mn|ls),  K[D]_{mn} = sum_{ls} D_{ls} (ml|sn).
   RHF passes (2D, D), UHF passes (Da+Db, Da) and (Da+Db, Db) (HartreeFock.jl:322-327).
   The ERIs are evaluated directly (never stored), shell quartets with
   Q_ab*Q_cd < schwarz_thr skipped (Q = sqrt of the largest diagonal integral of
   the shell pair).  DJ, DK, G: n_set*n_bf*n_bf host arrays. */
int  qbx_jk(qbx_ctx* ctx, qbx_basis* basis, int32_t n_set, const double* DJ,
            const double* DK, double* G, double schwarz_thr, int32_t precision_mode);
/* Same with device pointers; enqueued on `stream` (a hipStream_t, NULL for the
   context stream) and returning without synchronizing.  Builds of one basis
   share its work buffers: a call on another stream than the basis's last
   build waits on the host for that build to finish first. */
int  qbx_jk_device(qbx_ctx* ctx, qbx_basis* basis, int32_t n_set, const double* dDJ,
                   const double* dDK, double* dG, double schwarz_thr,
                   int32_t precision_mode, void* stream);

/* Work statistics of qbx_jk at a threshold (for measurement): per ERI class c
   (order of qbx_class_table) the number of screened shell quartets, contracted
   Cartesian quartets and primitive Cartesian quartets, plus the flops that the
   generated kernels perform per primitive shell quartet and per contracted
   shell quartet.  Arrays have qbx_n_classes() entries; NULL entries skipped. */
int32_t qbx_n_classes(void);
int  qbx_class_table(int32_t* l4 /* 4*n_classes */);
int  qbx_jk_stats(qbx_ctx* ctx, qbx_basis* basis, double schwarz_thr,
                  double fp64_cut, int64_t* shell_quartets, int64_t* cart_quartets,
                  int64_t* prim_quartets, int64_t* fp64_shell_quartets);
/* Per-class kernel time of qbx_jk / qbx_jk_device, from HIP events on the
   stream each class is launched on.  mode 1: the class kernels of a build run
   one after another on one stream, and qbx_last_timing writes the last
   build's n_classes milliseconds and its total.  mode 2: the build runs
   concurrently as it does untimed (classes spread over the auxiliary
   streams); every build since the last read (or since qbx_set_timing) is
   recorded, and qbx_last_timing writes each class's average duration in the
   concurrent builds and the average build time, then starts a new average.
   mode 0: off. */
int  qbx_set_timing(qbx_ctx* ctx, int32_t mode);
int  qbx_last_timing(qbx_ctx* ctx, float* ms_per_class, float* ms_total);

/* The Boys function the ERI kernels use, evaluated on its own for tests (no
   reference counterpart; the reference's is computeBoysSequence,
   BoysFunction.jl:67-77): F_0..F_order at the n arguments T (host arrays,
   0 <= T <= 1e6), from the same order-`order` table rows staged in LDS, into F
   (n x (order+1), row-major).  order in [0, 8]. */
int  qbx_boys_device(qbx_ctx* ctx, int32_t order, int32_t n, const double* T, double* F);

/* One-electron integrals of n primitive pairs (unnormalized primitives
   x_A^i y_A^j z_A^k exp(-a r_A^2), Cartesian exponents up to a total of 10 per
   primitive), one value per pair: the reference's per-primitive engines
   computePGTOrbOverlap (GaussianOrbitals.jl:147-179), the kinetic energy
   -1/2 <a|nabla^2|b> (computePGTOrbCoordDiff, :295-363, with Samplers.jl's
   -1/2), the one-body repulsion int a b / |r - C| without a charge
   (computePGTOrbOneBodyRepulsion!, :478-522) and the multipole
   <a|(x-C_x)^d_x (y-C_y)^d_y (z-C_z)^d_z|b> (:184-247; degree per axis in
   [0, 4]).  A, B: [n][3] centres; a, b: [n] exponents; la, lb: [n][3]
   Cartesian exponents; C: the point (kinds COULOMB and MULTIPOLE, else may be
   NULL); degree: [3] (MULTIPOLE, else may be NULL). */
#define QBX_PRIM_OVERLAP 0
#define QBX_PRIM_KINETIC 1
#define QBX_PRIM_COULOMB 2
#define QBX_PRIM_MULTIPOLE 3
int  qbx_prim_one_e(qbx_ctx* ctx, int32_t kind, int32_t n, const double* A, const double* a, const int32_t* la,
                    const double* B, const double* b, const int32_t* lb, const double* C, const int32_t* degree,
                    double* out);

/* Electron-repulsion integrals (ab|cd) of n primitive quartets with unit
   coefficients, one value each: the reference's per-primitive engine
   computePGTOrbTwoBodyRepulsion! (GaussianOrbitals.jl:594-663).  cen: [n][4][3]
   centres, xpn: [n][4] exponents, ang: [n][4][3] Cartesian exponents, in the
   order a, b, c, d; at most 10 per primitive and 24 per quartet. */
int  qbx_prim_eri(qbx_ctx* ctx, int32_t n, const double* cen, const double* xpn, const int32_t* ang, double* out);

/* Mixed-precision cut: quartets whose Schwarz bound exceeds `cut` run in FP64
   under QBX_PREC_MIXED (default 1e-5). */
int  qbx_set_mixed_cut(qbx_ctx* ctx, double cut);

/* Density screening for qbx_jk / qbx_jk_device (0 = off, the default): shell
   quartets with Q_ab Q_cd max|D| < thr are skipped, where max|D| is the largest
   element of the six density blocks (ab, cd, ac, ad, bc, bd) over all DJ/DK sets
   of the call (Haser & Ahlrichs, J. Comput. Chem. 10, 104 (1989)).  Meant for
   incremental builds G(D_n) = G(D_{n-1}) + G(D_n - D_{n-1}) in an SCF loop.
   No reference counterpart: the reference contracts a stored tensor. */
int  qbx_set_density_screening(qbx_ctx* ctx, double thr);

/* Shell table of a basis (the shells found by qbx_basis_create: functions with a
   common centre and exponent list whose weights are proportional).  Shell s has
   angular momentum shell_L[s] and shell_nprim[s] radial primitives
   radial_xpn/radial_coef[shell_poff[s] + k] (the coefficients of its first
   component's weights); component c (Quiqbox order) is basis function
   shell_bf[QBX_SHELL_MAXC s + c] (-1 if absent) with weight ratio
   shell_sc[QBX_SHELL_MAXC s + c].
   Arrays: n_shell entries (x6 for bf/sc, x3 for the centres), n_radial for the
   radial arrays; NULL entries skipped.  No reference counterpart: it exposes
   how the basis maps onto the kernels' shells, which the gradient reports in. */
int  qbx_basis_shells(const qbx_basis* basis, int32_t* n_shell, int32_t* n_radial, int32_t* shell_L,
                      int32_t* shell_nprim, int32_t* shell_poff, int32_t* shell_bf, double* shell_sc,
                      double* shell_cen, double* radial_xpn, double* radial_coef);

/* Gradient of the RHF energy with respect to the basis parameters, s and p
   shells with at most 8 radial primitives (QBX_EUNSUP otherwise).  Inputs: the
   converged total density Dt = 2 C_occ C_occ^T and the energy-weighted density
   W = 2 C_occ diag(eps_occ) C_occ^T (n_bf x n_bf), the nuclei (charges Z, R
   3*n_nuc, fixed), the Schwarz threshold of the ERI task lists.  Outputs, from
   E = sum D (T+V) - sum W S + 1/2 sum Gamma (mu nu|lam sig) with every basis
   function's weights held fixed:
     g_xpn[n_radial]  dE/d(exponent) of each radial primitive,
     g_lin[n_radial]  c dE/dc for its coefficient c (all components scaled alike),
     g_cen[3*n_shell] dE/d(centre) of each shell.
   The reference has no gradient (Quiqbox v0.6.3 offers the parameter boxes of
   Parameters.jl:14-43 but no derivative of the energy); the host chains these
   through the normalization to the user's parameters. */
int  qbx_rhf_gradient(qbx_ctx* ctx, qbx_basis* basis, const double* Dt, const double* W, int32_t n_nuc,
                      const double* Z, const double* R, double schwarz_thr, double* g_xpn, double* g_lin,
                      double* g_cen);

/* The same gradient for a UHF state (the reference's UOHartreeFock,
   HartreeFock.jl:63; its G, :322-327): Dt = D_alpha + D_beta, the spin densities
   D_s = C_s,occ C_s,occ^T, and W = sum_s C_s,occ diag(eps_s,occ) C_s,occ^T;
   the two-electron term is sum (1/2 Dt_mn Dt_ls - 1/2 sum_s D_s,ml D_s,ns)
   (mn|ls).  With D_alpha = D_beta = Dt/2 it equals qbx_rhf_gradient. */
int  qbx_uhf_gradient(qbx_ctx* ctx, qbx_basis* basis, const double* Dt, const double* D_alpha,
                      const double* D_beta, const double* W, int32_t n_nuc, const double* Z, const double* R,
                      double schwarz_thr, double* g_xpn, double* g_lin, double* g_cen);

/* ---- Density from the Fock matrix without diagonalization ----------------
   The reference forms each SCF step's density from eigen(X' F X)
   (solveFockMatrix/getC/getD, HartreeFock.jl:46-59, 296-302); this entry point
   gives the same matrix, D = C[:, :nocc] C[:, :nocc]^T, as the spectral
   projector of X' F X onto its nocc lowest eigenvalues, by trace-correcting
   purification (SP2: P <- P^2 or 2P - P^2 from P0 = (hi I - X'FX)/(hi - lo),
   Gershgorin bounds lo, hi), every product an FP64 GEMM on the matrix cores.
   Host arrays, n x n row-major; X is the orthonormalization
   (HartreeFock.jl:39-42).  Needs a gap between eigenvalues nocc and nocc+1:
   QBX_ENOCONV when the trace or the idempotency has not settled after
   max_iter products (the caller then diagonalizes).  iterations (products
   taken) and idempotency (|tr P - tr P^2| at the end) may be NULL. */
int  qbx_density_purify(qbx_ctx* ctx, int32_t n, int32_t nocc, const double* X, const double* F, double* D,
                        int32_t max_iter, int32_t* iterations, double* idempotency);

/* The same density without allocating: no counterpart in the reference, for
   callers that cannot allocate device memory (code under HIP graph capture,
   tight serving loops, harnesses that forbid hipMalloc).  X, F (inputs) and D
   (output) are n x n row-major DEVICE arrays; the work matrices live in a
   fixed device array of the library (3 n^2 + 2 n + 4 doubles for n up to
   QBX_PURIFY_NOALLOC_MAX_N: 25 MB per GPU, held until the process ends), so
   n above the maximum is QBX_EINVAL.  Enqueued on `stream` (NULL: the
   context stream), which it synchronizes to read the traces; returns with D
   written.  Every call on one GPU shares that array: calls that overlap on
   one GPU race.  Nothing in it needs clearing between calls (every entry is
   written before it is read). */
#define QBX_PURIFY_NOALLOC_MAX_N 1024
int  qbx_density_purify_noalloc(qbx_ctx* ctx, int32_t n, int32_t nocc, const double* dX, const double* dF,
                                double* dD, int32_t max_iter, int32_t* iterations, double* idempotency,
                                void* stream);

/* One SCF step's orbitals and density on the GPU: getC + getD
   (solveFockMatrix, HartreeFock.jl:46-59; getD :296-302).  eigen(X' F X)
   ascending, C = X C' with each column's sign fixed so that C[0][j] >= 0,
   D = C[:, :nocc] C[:, :nocc]^T; X' F X, X C' and D are FP64 matrix-core
   GEMMs and the symmetric eigenproblem is rocSOLVER's divide and conquer
   (dsyevd).  Host arrays, n x n row-major (X and F symmetric); eps (n) and
   D may be NULL.  QBX_ENOCONV if the eigensolver does not converge. */
int  qbx_fock_solve(qbx_ctx* ctx, int32_t n, int32_t nocc, const double* X, const double* F, double* C,
                    double* eps, double* D);

/* C[M x N] = alpha op(A) op(B) + beta C in FP64 on the matrix cores
   (v_mfma_f64_16x16x4f64), row-major device matrices, op as for
   qbx_gemm_bf16; enqueued on `stream` (NULL: default stream), no sync.  The
   GEMM of the purification above, exported for its tests. */
int  qbx_gemm_f64(int32_t trans_a, int32_t trans_b, int32_t M, int32_t N, int32_t K, double alpha,
                  const double* A, int32_t lda, const double* B, int32_t ldb, double beta, double* C,
                  int32_t ldc, void* stream);

/* ---- BF16 matrix-core kernels of the learned initial guess --------------
   No reference counterpart: a pair-feature MLP that predicts the initial SCF
   density (the reference's guesses are CoreH/GWH/SAD, HartreeFock.jl:215-293).
   Device pointers, enqueued on `stream` (NULL: default stream), no sync.
   Matrices row-major; BF16 as raw 16-bit patterns.

   C[M x N] (FP32, row stride ldc) = alpha op(A) op(B) + beta C, op(A) = A
   (M x K, stride lda) or A^T (A stored K x M), likewise B; FP32 accumulation
   on the MFMA units.  epilogue 0: plain; 1: C = relu(alpha AB + bias[n]) and
   its BF16 copy Cb; 2: C = alpha AB + bias[n].  Epilogues 1-2 take
   untransposed operands. */
int  qbx_gemm_bf16(int32_t trans_a, int32_t trans_b, int32_t M, int32_t N, int32_t K, float alpha,
                   const void* A, int32_t lda, const void* B, int32_t ldb, float beta, float* C, int32_t ldc,
                   int32_t epilogue, const float* bias, void* Cb, void* stream);
/* op 0: out0 (BF16) = x (FP32) converted; op 1: ReLU backward, out1 (FP32) and
   out0 (BF16) = x (FP32 dH) where y (BF16 H) > 0, else 0. */
int  qbx_mlp_elementwise(int32_t op, int64_t n, const void* x, const void* y, void* out0, void* out1, void* stream);
/* weighted squared error: loss += sum w_r m_c (Y - T)^2 scale, dY = 2 w_r m_c (Y - T) scale (FP32 and BF16) */
int  qbx_mlp_mse(int32_t rows, int32_t cols, const float* Y, const float* T, const float* row_w,
                 const float* col_mask, float scale, float* dY, void* dYb, float* loss, void* stream);
/* out[c] = sum_r X[r][c] */
int  qbx_mlp_colsum(int32_t rows, int32_t cols, const float* X, float* out, void* stream);
/* Adam step `step` (1-based) on FP32 weights w with moments m, v and gradient g; BF16 copy to wb if non-NULL */
int  qbx_mlp_adam(int64_t n, float* w, float* m, float* v, const float* g, void* wb, float lr, float b1, float b2,
                  float eps, int32_t step, void* stream);

/* Last error message of the calling thread; returns its full length. */
int  qbx_last_error(char* buf, size_t len);

/* Digest (16 hex digits of a sha256) of the sources this library was built
   from, as quiqbox.jl_amd/build.py computes it ("unknown" for a build without
   it).  No reference counterpart: build provenance, so that a run can tell a
   library built from the sources beside it from a stale one. */
const char* qbx_source_digest(void);

#ifdef __cplusplus
}
#endif
#endif
