/*
 * qbx_oracle.h -- CPU FP64 restatement of Quiqbox's Gaussian-integral and
 * Fock-build path.  TEST INFRASTRUCTURE ONLY: this library is the checker
 * that the HIP path is compared against (tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py).  Product code never links or calls it.
 *
 * Every entry point names the reference function it restates
 * (paths relative to the Quiqbox.jl v0.6.3 source tree).
 *
 * Conventions: FP64, 0-based indices, Cartesian angular momentum (i,j,k),
 * basis functions given in CSR form over deduplicated primitives with final
 * contraction weights (the normalized weights of Framework.jl:748-773).
 * Square matrices are N*N, symmetric, so row/column major do not matter.
 * The ERI tensor is N^4 with eri[((i*N+j)*N+k)*N+l] = (ij|kl); the real
 * 8-fold symmetric tensor reads the same in Julia's column-major order.
 */
#ifndef QBX_ORACLE_H
#define QBX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* --- Boys function (src/Integration/Engines/BoysFunction.jl) ------------ */
double qor_boys(double x, int n);                     /* computeBoysFunc :62   */
void   qor_boys_seq(double x, int n, double* out);    /* computeBoysSequence :67 */
double qor_gamma_p(double a, double x);               /* regularized lower incomplete gamma */

/* --- primitive integrals (src/Integration/Engines/GaussianOrbitals.jl) -- */
/* A primitive is (center[3], exponent, ang[3]); unnormalized x^i y^j z^k e^{-a r^2}. */
double qor_prim_overlap(const double* A, double a, const int32_t* la,
                        const double* B, double b, const int32_t* lb);            /* :158 */
double qor_prim_kinetic(const double* A, double a, const int32_t* la,
                        const double* B, double b, const int32_t* lb);            /* :321,:686 */
/* positive Coulomb-field value of the pair at point C (no charge factor)        :478 */
double qor_prim_coulomb1(const double* A, double a, const int32_t* la,
                         const double* B, double b, const int32_t* lb, const double* C);
/* primitive ERI (ab|cd), chemists' notation                                     :594 */
double qor_prim_eri(const double* A, double a, const int32_t* la,
                    const double* B, double b, const int32_t* lb,
                    const double* C, double c, const int32_t* lc,
                    const double* D, double d, const int32_t* ld);
/* primitive multipole moment <a| prod (x-M)^n |b>                               :206 */
double qor_prim_multipole(const double* A, double a, const int32_t* la,
                          const double* B, double b, const int32_t* lb,
                          const double* M, const int32_t* deg);

/* --- contracted basis (src/Integration/Framework.jl) ---------------------- */
typedef struct {
    int32_t n_prim;
    const double*  cen;      /* 3*n_prim */
    const double*  xpn;      /* n_prim   */
    const int32_t* ang;      /* 3*n_prim */
    int32_t n_bf;
    const int32_t* bf_ptr;   /* n_bf+1   */
    const int32_t* bf_prim;  /* bf_ptr[n_bf] */
    const double*  bf_w;     /* bf_ptr[n_bf] */
} qor_basis;

void   qor_overlap_matrix(const qor_basis* b, double* S);                         /* :581 */
void   qor_kinetic_matrix(const qor_basis* b, double* T);
void   qor_nuclear_matrix(const qor_basis* b, int32_t n_nuc, const double* Z,
                          const double* R, double* V);                             /* Samplers.jl:56 */
void   qor_multipole_matrix(const qor_basis* b, const double* M, const int32_t* deg,
                            double* out);
/* 1: evaluate primitive quartets in the reference's canonical orientation;
   0 (default): larger exponent first in each pair, smaller exponent sum on electron 1 */
void   qor_set_reference_orientation(int32_t on);
double qor_eri_quartet(const qor_basis* b, int32_t i, int32_t j, int32_t k, int32_t l); /* :526 */
void   qor_eri_dense(const qor_basis* b, double* eri);                             /* :640 */
/* evaluate a list of contracted quartets (4*n indices); threads<=0 -> 1 */
void   qor_eri_list(const qor_basis* b, int64_t n, const int32_t* ijkl, double* out,
                    int32_t threads);

/* --- Fock build (src/HartreeFock.jl:305-319) ------------------------------ */
void   qor_gcore(int32_t n, const double* eri, const double* DJ, const double* DK,
                 double* G);

#ifdef __cplusplus
}
#endif
#endif
