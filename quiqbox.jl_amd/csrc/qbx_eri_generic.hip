// qbx_eri_generic.hip -- two-electron integrals of the shell quartets that
// hold an f or g shell (angular momentum above the generated classes), in
// every mode: J/K digest, dense tensor, Schwarz bounds.
//
// Scheme: the Obara-Saika recurrences axis by axis for one Cartesian
// component quartet (Obara & Saika, J. Chem. Phys. 84, 3963 (1986)), the
// Boys auxiliary index carried from axis to axis: on each axis the vertical
// recurrence on electron 1 up to the axis's total power, the electron
// transfer to electron 2, then the horizontal transfers on the ket and the
// bra.  This is the per-component form of the recurrences that Quiqbox
// evaluates (GaussianOrbitals.jl:525-663); here every primitive quartet is
// first put in the orientation where the transfer steps do not amplify
// rounding: within each pair the larger exponent first, and the electron
// whose pair has the smaller exponent sum first (the transfer multiplies by
// p/q <= 1).  Its arithmetic does not share intermediates between component
// quartets, so it is kept for the rare high-L quartets; s/p/d quartets take
// the generated class kernels.
//
// Work mapping: one wavefront per shell quartet, its lanes over the
// quartet's Cartesian component quartets (strides of 64), each lane summing
// its component over all primitive quartets; a quartet with fewer than 32
// components gives each component several lanes, which split its primitive
// quartets.  Per-lane recurrence tables of (NM+1)^2 doubles live in scratch.
#include "qbx_eri.h"

namespace qbx {
namespace {

constexpr double kTwoPi52 = 34.986836655249725693;   // 2 pi^{5/2}

struct PrimGeom {
    double p, Px, Py, Pz;     // exponent sum, product centre
    double PA[3];             // P - A (first centre of the pair)
    double AB[3];             // A - B
    double ea, eb;            // exponents of the pair's first and second primitive
    double K;                 // c_a c_b exp(-ab/p |A-B|^2) / p
};

// One axis: vals[0..nUp] hold the values carried from the previous axes
// (index = Boys auxiliary order); on return vals[0..nUp-S] hold this axis
// processed, S = a + b + c + d its total power.
template <int NM>
__device__ void axis_pass(double* vals, int& nUp, int a, int b, int c, int d, double xPA, double xPQ, double xAB,
                          double xCD, double fac, double i2p, double iq, double poq, double shiftq)
{
    const int S = a + b + c + d;
    if (S == 0) return;
    double T[NM + 1][NM + 1];
    for (int j = 0; j <= nUp; ++j) T[0][j] = vals[j];
    // vertical recurrence on electron 1:
    // [i+1]^(j) = PA [i]^(j) - (rho/p) PQ [i]^(j+1) + i/(2p) ([i-1]^(j) - (rho/p) [i-1]^(j+1))
    for (int i = 0; i < S; ++i)
        for (int j = 0; j + i < nUp; ++j) {
            double v = xPA * T[i][j] - fac * xPQ * T[i][j + 1];
            if (i > 0) v += (i * i2p) * (T[i - 1][j] - fac * T[i - 1][j + 1]);
            T[i + 1][j] = v;
        }
    const int nOut = nUp - S, cd = c + d;
    const double i2q = 0.5 * iq;
    for (int m = 0; m <= nOut; ++m) {
        // electron transfer: [i, o+1] = (i [i-1, o] + o [i, o-1]) / (2q)
        //                               - (b X_AB + d X_CD) / q [i, o] - p/q [i+1, o]
        double E[NM + 1][NM + 1];
        for (int i = 0; i <= S; ++i) E[i][0] = T[i][m];
        for (int o = 0; o < cd; ++o)
            for (int i = 0; i + o < S; ++i) {
                double t1 = 0.0;
                if (i > 0) t1 += i * E[i - 1][o];
                if (o > 0) t1 += o * E[i][o - 1];
                E[i][o + 1] = t1 * i2q - shiftq * E[i][o] - poq * E[i + 1][o];
            }
        // horizontal transfers: (k, t+1) = (k+1, t) + X (k, t), ket then bra
        double g[NM + 1];
        for (int i = a; i <= a + b; ++i) {
            double h[NM + 1];
            for (int o = c; o <= cd; ++o) h[o] = E[i][o];
            for (int s = 0; s < d; ++s)
                for (int k = c; k < cd - s; ++k) h[k] = h[k + 1] + xCD * h[k];
            g[i] = h[c];
        }
        for (int s = 0; s < b; ++s)
            for (int k = a; k < a + b - s; ++k) g[k] = g[k + 1] + xAB * g[k];
        vals[m] = g[a];
    }
    nUp = nOut;
}

__device__ inline void flip_pair(PrimGeom& g, int* la, int* lb)
{
    for (int t = 0; t < 3; ++t) {
        g.PA[t] += g.AB[t];       // P - B = (P - A) + (A - B)
        g.AB[t] = -g.AB[t];
        const int x = la[t];
        la[t] = lb[t];
        lb[t] = x;
    }
    const double e = g.ea;
    g.ea = g.eb;
    g.eb = e;
}

// One primitive quartet of one Cartesian component quartet, weights included.
template <int NM>
__device__ double prim_quartet(PrimGeom X, PrimGeom Y, int la[3], int lb[3], int lc[3], int ld[3],
                               const double* boys, const double* bexp)
{
    if (X.eb > X.ea) flip_pair(X, la, lb);
    if (Y.eb > Y.ea) flip_pair(Y, lc, ld);
    int* a = la;
    int* b = lb;
    int* c = lc;
    int* d = ld;
    if (X.p > Y.p) {
        const PrimGeom t = X;
        X = Y;
        Y = t;
        a = lc;
        b = ld;
        c = la;
        d = lb;
    }
    const double p = X.p, q = Y.p, pq = p + q;
    const double PQ[3] = {X.Px - Y.Px, X.Py - Y.Py, X.Pz - Y.Pz};
    const double rho = p * q / pq;
    const double T = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
    const int Lt = a[0] + a[1] + a[2] + b[0] + b[1] + b[2] + c[0] + c[1] + c[2] + d[0] + d[1] + d[2];
    double vals[NM + 1];
    qbx_boys_rt(Lt, T, vals, boys, bexp);
    const double fac = q / pq, i2p = 0.5 / p, iq = 1.0 / q, poq = p / q;
    int nUp = Lt;
    for (int t = 0; t < 3; ++t) {
        const double shiftq = (X.eb * X.AB[t] + Y.eb * Y.AB[t]) * iq;
        axis_pass<NM>(vals, nUp, a[t], b[t], c[t], d[t], X.PA[t], PQ[t], X.AB[t], Y.AB[t], fac, i2p, iq, poq,
                      shiftq);
    }
    return kTwoPi52 * X.K * Y.K / sqrt(pq) * vals[0];
}

__device__ inline PrimGeom load_geom(const GenArgs& a, const PrimPair& P, int k, int sA, int sB, const double AB[3])
{
    PrimGeom g;
    g.p = P.p;
    g.Px = P.Px;
    g.Py = P.Py;
    g.Pz = P.Pz;
    g.PA[0] = P.PAx;
    g.PA[1] = P.PAy;
    g.PA[2] = P.PAz;
    g.AB[0] = AB[0];
    g.AB[1] = AB[1];
    g.AB[2] = AB[2];
    const int kk = a.ppk[k];
    g.ea = a.pexp[a.poff[sA] + (kk & 0xffff)];
    g.eb = a.pexp[a.poff[sB] + (kk >> 16)];
    g.K = P.Kp;
    return g;
}

__device__ inline void atomic_max_nonneg(unsigned long long* p, double v)
{
    // non-negative doubles order as their bit patterns
    atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

template <int NM, int MODE>
__global__ __launch_bounds__(256) void eri_generic_kernel(GenArgs a)
{
    const int lane = (int)threadIdx.x & 63;
    const int64_t task = (int64_t)blockIdx.x * 4 + ((int)threadIdx.x >> 6);
    if (task >= a.total) return;
    int lo = 0, hi = a.nbra;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.prefix[mid] <= task) lo = mid; else hi = mid;
    }
    const int bi = lo;
    const int k = (int)(task - a.prefix[bi]);
    const int sA = a.hA[bi], sB = a.hB[bi];
    const int bo = a.ho[bi], bn = a.hn[bi];
    int sC, sD, ko, kn;
    double qk;
    bool same = false;
    if (MODE == MODE_SCHWARZ) {
        sC = sA;
        sD = sB;
        ko = bo;
        kn = bn;
        qk = 0.0;
        same = true;
    } else if (k < a.nlow[bi]) {
        sC = a.lA[k];
        sD = a.lB[k];
        ko = a.lo[k];
        kn = a.ln[k];
        qk = a.lQ[k];
    } else {
        const int j = k - a.nlow[bi];
        sC = a.hA[j];
        sD = a.hB[j];
        ko = a.ho[j];
        kn = a.hn[j];
        qk = a.hQ[j];
        same = j == bi;
    }
    if (MODE == MODE_JK && a.pdmax) {
        const int ns = a.nshell;
        const double* dm = a.pdmax;
        double dmax = fmax(fmax(dm[sA * ns + sB], dm[sC * ns + sD]), fmax(dm[sA * ns + sC], dm[sA * ns + sD]));
        dmax = fmax(dmax, fmax(dm[sB * ns + sC], dm[sB * ns + sD]));
        if (a.hQ[bi] * qk * dmax < a.dthr) return;
    }
    const DevShell* SA = a.sh + sA;
    const DevShell* SB = a.sh + sB;
    const DevShell* SC = a.sh + sC;
    const DevShell* SD = a.sh + sD;
    const int NA = qbx_ncart(SA->L), NB = qbx_ncart(SB->L), NC = qbx_ncart(SC->L), ND = qbx_ncart(SD->L);
    const double AB[3] = {SA->x - SB->x, SA->y - SB->y, SA->z - SB->z};
    const double CD[3] = {SC->x - SD->x, SC->y - SD->y, SC->z - SD->z};
    const int ncomp = MODE == MODE_SCHWARZ ? NA * NB : NA * NB * NC * ND;
    double f = 1.0;
    if (sA == sB) f *= 0.5;
    if (sC == sD) f *= 0.5;
    if (same) f *= 0.5;
    const size_t N = (size_t)a.N;
    EriArgs ea;
    ea.J = a.J;
    ea.K = a.K;
    ea.xcd_local = a.xcd_local;
    ea.jk_stride = a.jk_stride;
    const JKSink sink(ea);
    // Lanes per component quartet: a quartet with fewer than 32 Cartesian
    // component quartets (an f shell with s shells, say) would leave most of
    // the wave idle, so lpc lanes (a power of two, lpc * ncomp <= 64) share
    // one component's primitive quartets and sum them with shuffles.
    int lpc = 1;
    while (2 * lpc * ncomp <= 64) lpc *= 2;
    const int sub = lane & (lpc - 1);
    const int npq = bn * kn;
    for (int c0 = 0; c0 < ncomp; c0 += 64 / lpc) {
        const int cidx = c0 + lane / lpc;
        int ia = 0, ib = 0, ic = 0, id = 0;
        bool live = cidx < ncomp;
        if (live) {
            if (MODE == MODE_SCHWARZ) {
                ia = cidx / NB;
                ib = cidx - ia * NB;
                ic = ia;
                id = ib;
            } else {
                int r = cidx;
                id = r % ND;
                r /= ND;
                ic = r % NC;
                r /= NC;
                ib = r % NB;
                ia = r / NB;
            }
        }
        const int fa = live ? SA->bf[ia] : -1, fb = live ? SB->bf[ib] : -1;
        const int fc = live ? SC->bf[ic] : -1, fd = live ? SD->bf[id] : -1;
        live = live && fa >= 0 && fb >= 0 && fc >= 0 && fd >= 0;
        double acc = 0.0;
        if (live) {
            int ga[3], gb[3], gc[3], gd[3];
            qbx_cart(SA->L, ia, ga);
            qbx_cart(SB->L, ib, gb);
            qbx_cart(SC->L, ic, gc);
            qbx_cart(SD->L, id, gd);
            if (lpc == 1) {
                for (int u = 0; u < bn; ++u) {
                    const PrimGeom X = load_geom(a, a.pp[bo + u], bo + u, sA, sB, AB);
                    for (int v = 0; v < kn; ++v) {
                        const PrimGeom Y = load_geom(a, a.pp[ko + v], ko + v, sC, sD, CD);
                        int la[3] = {ga[0], ga[1], ga[2]}, lb[3] = {gb[0], gb[1], gb[2]};
                        int lc[3] = {gc[0], gc[1], gc[2]}, ld[3] = {gd[0], gd[1], gd[2]};
                        acc += prim_quartet<NM>(X, Y, la, lb, lc, ld, a.boys, a.bexp);
                    }
                }
            } else {
                for (int t = sub; t < npq; t += lpc) {
                    const int u = t / kn, v = t - u * kn;
                    const PrimGeom X = load_geom(a, a.pp[bo + u], bo + u, sA, sB, AB);
                    const PrimGeom Y = load_geom(a, a.pp[ko + v], ko + v, sC, sD, CD);
                    int la[3] = {ga[0], ga[1], ga[2]}, lb[3] = {gb[0], gb[1], gb[2]};
                    int lc[3] = {gc[0], gc[1], gc[2]}, ld[3] = {gd[0], gd[1], gd[2]};
                    acc += prim_quartet<NM>(X, Y, la, lb, lc, ld, a.boys, a.bexp);
                }
            }
        }
        // every lane of the wave takes part in the shuffles (lpc is uniform)
        for (int off = lpc >> 1; off > 0; off >>= 1) acc += __shfl_down(acc, off, lpc);
        if (!live || sub != 0) continue;
        const double val = acc * (SA->sc[ia] * SB->sc[ib]) * (SC->sc[ic] * SD->sc[id]);
        if (MODE == MODE_SCHWARZ) {
            atomic_max_nonneg(a.qbits + bi, fabs(val));
        } else if (MODE == MODE_DENSE) {
            const size_t i = fa, j = fb, kk = fc, l = fd;
            a.eri[((i * N + j) * N + kk) * N + l] = val;
            a.eri[((j * N + i) * N + kk) * N + l] = val;
            a.eri[((i * N + j) * N + l) * N + kk] = val;
            a.eri[((j * N + i) * N + l) * N + kk] = val;
            a.eri[((kk * N + l) * N + i) * N + j] = val;
            a.eri[((l * N + kk) * N + i) * N + j] = val;
            a.eri[((kk * N + l) * N + j) * N + i] = val;
            a.eri[((l * N + kk) * N + j) * N + i] = val;
        } else {
            // J_ab += 4f v D_cd, J_cd += 4f v D_ab, K_ac/K_ad/K_bd/K_bc += 2f v D
            // (the symmetrization G = (J + J^T)/2 - (K + K^T)/2 completes the 8-fold sum)
            const double fj = 4.0 * f * val, fk = 2.0 * f * val;
            for (int s = 0; s < a.nset; ++s) {
                const double* DJ = a.D + (size_t)s * N * N;
                const double* DK = a.D + (size_t)(a.nset + s) * N * N;
                double* Js = sink.J + (size_t)s * N * N;
                double* Ks = sink.K + (size_t)s * N * N;
                sink.add(Js + fa * N + fb, fj * DJ[fc * N + fd]);
                sink.add(Js + fc * N + fd, fj * DJ[fa * N + fb]);
                sink.add(Ks + fa * N + fc, fk * DK[fb * N + fd]);
                sink.add(Ks + fa * N + fd, fk * DK[fb * N + fc]);
                sink.add(Ks + fb * N + fd, fk * DK[fa * N + fc]);
                sink.add(Ks + fb * N + fc, fk * DK[fa * N + fd]);
            }
        }
    }
}

template <int NM>
void launch_nm(const GenArgs& a, int mode, hipStream_t s)
{
    if (a.total <= 0) return;
    const int64_t blocks = (a.total + 3) / 4;
    if (blocks > 0x7fffffffLL) throw QbxError(QBX_EUNSUP, "too many high angular momentum quartets for one grid");
    switch (mode) {
    case MODE_JK:
        hipLaunchKernelGGL((eri_generic_kernel<NM, MODE_JK>), dim3((unsigned)blocks), dim3(256), 0, s, a);
        break;
    case MODE_DENSE:
        hipLaunchKernelGGL((eri_generic_kernel<NM, MODE_DENSE>), dim3((unsigned)blocks), dim3(256), 0, s, a);
        break;
    case MODE_SCHWARZ:
        hipLaunchKernelGGL((eri_generic_kernel<NM, MODE_SCHWARZ>), dim3((unsigned)blocks), dim3(256), 0, s, a);
        break;
    default: throw QbxError(QBX_EINTERNAL, "unknown ERI mode");
    }
    QBX_HIP(hipGetLastError());
}

// Primitive quartets one thread each (unit coefficients): the reference's
// per-primitive engine computePGTOrbTwoBodyRepulsion! (GaussianOrbitals.jl:594-663).
__global__ __launch_bounds__(64) void prim_eri_kernel(int n, const double* __restrict__ cen,
                                                      const double* __restrict__ xpn, const int* __restrict__ ang,
                                                      const double* boys, const double* bexp, double* __restrict__ out)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    int l[4][3];
    double c[4][3], e[4];
    for (int k = 0; k < 4; ++k) {
        e[k] = xpn[4 * i + k];
        for (int t = 0; t < 3; ++t) {
            c[k][t] = cen[12 * i + 3 * k + t];
            l[k][t] = ang[12 * i + 3 * k + t];
        }
    }
    PrimGeom G[2];
    for (int h = 0; h < 2; ++h) {
        const int ka = 2 * h, kb = 2 * h + 1;
        PrimGeom& g = G[h];
        g.ea = e[ka];
        g.eb = e[kb];
        g.p = e[ka] + e[kb];
        double r2 = 0.0, P[3];
        for (int t = 0; t < 3; ++t) {
            g.AB[t] = c[ka][t] - c[kb][t];
            r2 += g.AB[t] * g.AB[t];
            P[t] = (e[ka] * c[ka][t] + e[kb] * c[kb][t]) / g.p;
            g.PA[t] = P[t] - c[ka][t];
        }
        g.Px = P[0];
        g.Py = P[1];
        g.Pz = P[2];
        g.K = exp(-e[ka] * e[kb] / g.p * r2) / g.p;
    }
    out[i] = prim_quartet<QBX_PRIM_ERI_LTOT>(G[0], G[1], l[0], l[1], l[2], l[3], boys, bexp);
}

}  // namespace

void prim_eri(int n, const double* cen, const double* xpn, const int* ang, const double* boys, const double* bexp,
              double* out, hipStream_t s)
{
    static_assert(QBX_PRIM_ERI_LTOT <= QBX_BOYS_NMAX, "Boys table orders for the primitive ERIs");
    if (n <= 0) return;
    hipLaunchKernelGGL(prim_eri_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, n, cen, xpn, ang, boys, bexp,
                       out);
    QBX_HIP(hipGetLastError());
}

void launch_eri_generic(const GenArgs& a, int nm, int mode, hipStream_t s)
{
    static_assert(4 * QBX_MAXL <= 16 && 4 * QBX_MAXL <= QBX_BOYS_NMAX, "generic kernel sizes");
    if (nm <= 12) launch_nm<12>(a, mode, s);
    else if (nm <= 16) launch_nm<16>(a, mode, s);
    else throw QbxError(QBX_EUNSUP, "angular momentum above the generic kernel");
}

}  // namespace qbx
