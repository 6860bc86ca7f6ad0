// qbx_device.h -- device-side building blocks of libqbx_hip.so (gfx950).
//
// Boys function F_m(T), one-electron kernels: tabulated on T_i = i/10
// (0 <= T < QBX_BOYS_TMAX) for m = 0..QBX_BOYS_NCOL-1 together with exp(-T_i)
// (global memory; the ERI kernels use the per-order LDS rows below).  For T below the cut, the
// highest needed order comes from an 8-term Taylor expansion about the nearest
// grid point (|T - T_i| <= 0.05, truncation < 1e-15 relative) and the lower
// orders from the stable downward recurrence
//   F_{m-1}(T) = (2T F_m(T) + e^{-T}) / (2m - 1);
// above the cut, F_0 = sqrt(pi/T)/2 and the upward recurrence without the
// exponentially small e^{-T} term.  Everything here is FP64, also in the
// mixed-precision kernels.
#pragma once
#include <hip/hip_runtime.h>

#define QBX_BOYS_DT 0.1
#define QBX_BOYS_INVDT 10.0
#define QBX_BOYS_TMAX 60.0
#define QBX_BOYS_ROWS 602
#define QBX_BOYS_NCOL 33     // F_0 .. F_32: highest order 25 (g quartets need 16, primitive quartets up to 24) plus 7 Taylor orders above it
#define QBX_BOYS_TAYLOR 8
#define QBX_BOYS_NMAX (QBX_BOYS_NCOL - QBX_BOYS_TAYLOR)
// Per-order rows for the ERI kernels (staged in LDS, 23 KB): on the grid
// T_i = i/4 (T < QBX_BOYS_TMAX), for total angular momentum L, row i holds
// F_{L+k}(T_i) / k! for k = 0..10 and exp(-T_i) (96-byte rows).  With
// |T - T_i| <= 1/8 the 11-term Taylor series truncates below 3e-18 relative.
#define QBX_BOYSL_DT 0.25
#define QBX_BOYSL_INVDT 4.0
#define QBX_BOYSL_ROWS 242
#define QBX_BOYSL_TAYLOR 11
#define QBX_BOYSL_STRIDE 12
#define QBX_BOYSL_LMAX 9

#define QBX_MAXL 4           // largest shell angular momentum (g); above 2 the ERIs take the generic kernel
#define QBX_MAXC 15          // Cartesian components of the largest shell
#define QBX_GRAD_MAXP 8      // radial primitives per shell in the gradient kernels

struct DevShell {
    double x, y, z;
    int L;
    int nprim;
    int bf[QBX_MAXC];        // basis-function index of each component, -1 absent
    double sc[QBX_MAXC];     // component scale (weight ratio to the shell's radial coefficients)
};

// Primitive-pair record of a shell pair (A,B): p = a+b, P the product centre,
// PA = P - A, Kp = c_a c_b exp(-ab/p |A-B|^2) / p.
struct PrimPair {
    double p, inv2p, Px, Py, Pz, PAx, PAy, PAz, Kp, pad;
};

__host__ __device__ constexpr int qbx_ncart(int L) { return (L + 1) * (L + 2) / 2; }

#define QBX_MAX_XCD 8        // J/K copies of the per-XCD accumulation

// XCD of the calling wave: hardware register XCC_ID (gfx940 and later),
// s_getreg_b32 hwreg(HW_REG_XCC_ID = 20, offset 0, size 4).
__device__ __forceinline__ int qbx_xcc_id() { return __builtin_amdgcn_s_getreg((3 << 11) | 20) & (QBX_MAX_XCD - 1); }

template <int L>
__device__ __forceinline__ void qbx_boys(double T, double* __restrict__ F,
                                         const double* __restrict__ tab,
                                         const double* __restrict__ etab)
{
    if (T < QBX_BOYS_TMAX) {
        const int i = __double2int_rn(T * QBX_BOYS_INVDT);
        const double md = (double)i * QBX_BOYS_DT - T;   // -(T - T_i)
        const double* __restrict__ row = tab + i * QBX_BOYS_NCOL + L;
        double acc = row[QBX_BOYS_TAYLOR - 1];
#pragma unroll
        for (int k = QBX_BOYS_TAYLOR - 1; k >= 1; --k) acc = row[k - 1] + (md * (1.0 / k)) * acc;
        F[L] = acc;
        if (L > 0) {
            double ex = 1.0;
#pragma unroll
            for (int k = 8; k >= 1; --k) ex = 1.0 + (md * (1.0 / k)) * ex;
            const double e = etab[i] * ex;
            const double t2 = 2.0 * T;
#pragma unroll
            for (int m = L - 1; m >= 0; --m) F[m] = (t2 * F[m + 1] + e) * (1.0 / (2 * m + 1));
        }
    } else {
        const double rt = rsqrt(T);
        F[0] = 0.88622692545275801365 * rt;
        const double h = 0.5 * rt * rt;
#pragma unroll
        for (int m = 0; m < L; ++m) F[m + 1] = F[m] * ((2 * m + 1) * h);
    }
}

// Runtime-order Boys function (1e kernels and the generic ERI kernel):
// F_0..F_n for n <= QBX_BOYS_NMAX (the top order's Taylor series reads table
// columns n .. n + QBX_BOYS_TAYLOR - 1).
__device__ __forceinline__ void qbx_boys_rt(int n, double T, double* __restrict__ F,
                                            const double* __restrict__ tab,
                                            const double* __restrict__ etab)
{
    if (T < QBX_BOYS_TMAX) {
        const int i = __double2int_rn(T * QBX_BOYS_INVDT);
        const double md = (double)i * QBX_BOYS_DT - T;
        const double* __restrict__ row = tab + i * QBX_BOYS_NCOL + n;
        double acc = row[QBX_BOYS_TAYLOR - 1];
        for (int k = QBX_BOYS_TAYLOR - 1; k >= 1; --k) acc = row[k - 1] + (md * (1.0 / k)) * acc;
        F[n] = acc;
        double ex = 1.0;
        for (int k = 8; k >= 1; --k) ex = 1.0 + (md * (1.0 / k)) * ex;
        const double e = etab[i] * ex;
        for (int m = n - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + e) / (2 * m + 1);
    } else {
        const double rt = rsqrt(T);
        F[0] = 0.88622692545275801365 * rt;
        const double h = 0.5 * rt * rt;
        for (int m = 0; m < n; ++m) F[m + 1] = F[m] * ((2 * m + 1) * h);
    }
}

// Cartesian component (i,j,k) of index c in shell L (order of Lexicons.jl:17-36).
__host__ __device__ inline void qbx_cart(int L, int c, int* ijk)
{
    int n = 0;
    for (int i = L; i >= 0; --i)
        for (int j = L - i; j >= 0; --j) {
            if (n == c) { ijk[0] = i; ijk[1] = j; ijk[2] = L - i - j; return; }
            ++n;
        }
    ijk[0] = ijk[1] = ijk[2] = 0;
}

// 1D overlap recurrence (Obara-Saika): table S[i][j] of
// int (x-A)^i (x-B)^j exp(-p (x-P)^2) dx / sqrt(pi/p),  i <= ni, j <= nj < M.
#define QBX_OV_MAX 9
template <int M>
__device__ inline void qbx_ov1d_t(double PA, double PB, double inv2p, int ni, int nj, double S[M][M])
{
    S[0][0] = 1.0;
    for (int i = 0; i <= ni; ++i) {
        if (i > 0) {
            double v = PA * S[i - 1][0];
            if (i > 1) v += inv2p * (i - 1) * S[i - 2][0];
            S[i][0] = v;
        }
        for (int j = 1; j <= nj; ++j) {
            double v = PB * S[i][j - 1];
            if (i > 0) v += inv2p * i * S[i - 1][j - 1];
            if (j > 1) v += inv2p * (j - 1) * S[i][j - 2];
            S[i][j] = v;
        }
    }
}

__device__ inline void qbx_ov1d(double PA, double PB, double inv2p, int ni, int nj, double S[QBX_OV_MAX][QBX_OV_MAX])
{
    qbx_ov1d_t<QBX_OV_MAX>(PA, PB, inv2p, ni, nj, S);
}
