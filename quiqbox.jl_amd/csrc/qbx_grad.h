// qbx_grad.h -- adjoint ERI kernels of libqbx_hip.so (gfx950): the two-electron
// part of the RHF energy gradient with respect to the basis parameters.
//
// For every unique contracted shell quartet (the screened task lists of the
// Fock build, one wave per bra pair and 64 of its kets) the kernel
//   1. forms the seeds G = f (4 D_ab D_cd - D_ac D_bd - D_ad D_bc) s_a s_b s_c s_d
//      on the contracted (ab|cd) (D the total density, f the 8-fold degeneracy
//      of the unique quartet, s the component scales), so that the sum of
//      G (ab|cd) over all unique quartets is the two-electron energy
//      sum (1/2 D_mn D_ls - 1/4 D_ml D_ns) (mn|ls) over all index quartets;
//      for UHF (nset = 2: D holds D, D_alpha, D_beta) the exchange part is
//      2 sum_s (D^s_ac D^s_bd + D^s_ad D^s_bc), i.e. the energy
//      sum (1/2 D_mn D_ls - 1/2 sum_s D^s_ml D^s_ns) (mn|ls), which is the
//      RHF form when D_alpha = D_beta = D/2;
//   2. runs the forward primitive loop (contracted [e0|f0], as the Fock build);
//   3. transposes the horizontal recurrence (EriAdj::hrr_adj): seeds on the
//      contracted [e0|f0] and the derivatives with respect to A-B and C-D;
//   4. runs the primitive loop again, recomputing every primitive pair quantity
//      (nothing is stored between the passes), the vertical recurrence forward
//      and then in reverse (EriAdj::vrr_adj), and the adjoint of the geometry,
//      Boys function (dF_m/dT = -F_{m+1}) and prefactor, down to the exponents
//      a, b, c, d and the centres A, B, C, D of the primitive quartet.
// Per radial primitive k of each shell it accumulates dE/d(alpha_k) at fixed
// weights and c_k dE/dc_k (the seeded value of the primitive quartets that
// contain k), per shell dE/d(centre); the bra's sums are reduced across the
// wave, the kets' go out per lane, all as FP64 atomics once per quartet.
#pragma once
#include "qbx_eri.h"

template <int LA, int LB, int LC, int LD> struct EriAdj;

namespace qbx {

constexpr int GRAD_MAXP = QBX_GRAD_MAXP;

struct GeomAdj {
    double da, db, dc, dd;
    double dA[3], dB[3], dC[3], dD[3];
};

// Adjoint of the primitive-quartet set-up.  Forward: p = a + b, P = (aA + bB)/p,
// PA = P - A, K_p = c_a c_b exp(-ab/p |A-B|^2)/p (likewise q, Q, QC, K_q), s =
// (p+q)^{-1/2}, PQ = P - Q, rho = pq/(p+q), T = rho |PQ|^2, pref = 2 pi^{5/2}
// K_p K_q s, F_m = pref B_m(T), WP = -q/(p+q) PQ, WQ = p/(p+q) PQ, inv2p =
// 1/2p, inv2q = 1/2q, inv2pq = 1/(2(p+q)), rho_p = q/(p+q), rho_q = p/(p+q).
// Given Fb[m] = dL/dF_m and bin[] = dL/d(PA, WP, QC, WQ, inv2p, inv2q, inv2pq,
// rho_p, rho_q), returns dL/d(a, b, c, d, A, B, C, D).  B_m for m <= LT + 1.
template <int LT>
__device__ __forceinline__ GeomAdj geom_adjoint(double a, double b, double c, double d, const double* PA,
                                                const double* QC, const double* AB, const double* CD,
                                                const double* PQ, double pref, double s, const double* Bm,
                                                const double* Fb, const double* bin)
{
    const double p = a + b, q = c + d;
    const double inv = s * s;
    const double rho = p * q * inv, rho_p = q * inv, rho_q = p * inv;
    const double R2 = PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2];
    double bpref = 0.0, bT = 0.0;
#pragma unroll
    for (int m = 0; m <= LT; ++m) {
        bpref += Fb[m] * Bm[m];
        bT -= Fb[m] * Bm[m + 1];
    }
    bT *= pref;
    const double blnK = bpref * pref;     // d pref / d ln K_p = d pref / d ln K_q = pref
    const double bs = bpref * pref / s;
    const double brho = bT * R2;
    double bPQ[3];
    double brho_p = bin[15], brho_q = bin[16];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        bPQ[t] = 2.0 * rho * bT * PQ[t] - rho_p * bin[3 + t] + rho_q * bin[9 + t];
        brho_p -= bin[3 + t] * PQ[t];
        brho_q += bin[9 + t] * PQ[t];
    }
    double bp = brho * q * inv + brho_q * inv;
    double bq = brho * p * inv + brho_p * inv;
    const double binv = brho * p * q + brho_p * q + brho_q * p + 0.5 * bin[14] + bs * 0.5 / s;
    bp -= binv * inv * inv + 0.5 * bin[12] / (p * p);
    bq -= binv * inv * inv + 0.5 * bin[13] / (q * q);
    GeomAdj g;
    g.da = bp;
    g.db = bp;
    g.dc = bq;
    g.dd = bq;
    // P = (aA + bB)/p: dP/da = (A - P)/p = -PA/p, dP/db = (B - P)/p = -(AB + PA)/p
    const double ab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
    const double cd2 = CD[0] * CD[0] + CD[1] * CD[1] + CD[2] * CD[2];
    const double mu = a * b / p, nu = c * d / q;
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const double bP = bPQ[t] + bin[t];     // PQ = P - Q, PA = P - A
        const double bQ = -bPQ[t] + bin[6 + t];
        g.da += -bP * PA[t] / p;
        g.db += -bP * (AB[t] + PA[t]) / p;
        g.dc += -bQ * QC[t] / q;
        g.dd += -bQ * (CD[t] + QC[t]) / q;
        // ln K_p = ln c_a c_b - (ab/p)|A-B|^2 - ln p
        g.dA[t] = (a / p) * bP - bin[t] - 2.0 * mu * AB[t] * blnK;
        g.dB[t] = (b / p) * bP + 2.0 * mu * AB[t] * blnK;
        g.dC[t] = (c / q) * bQ - bin[6 + t] - 2.0 * nu * CD[t] * blnK;
        g.dD[t] = (d / q) * bQ + 2.0 * nu * CD[t] * blnK;
    }
    g.da += blnK * (-(b * b) / (p * p) * ab2 - 1.0 / p);
    g.db += blnK * (-(a * a) / (p * p) * ab2 - 1.0 / p);
    g.dc += blnK * (-(d * d) / (q * q) * cd2 - 1.0 / q);
    g.dd += blnK * (-(c * c) / (q * q) * cd2 - 1.0 / q);
    return g;
}

__device__ __forceinline__ void sel_add(double* v, int k, double x)
{
#pragma unroll
    for (int t = 0; t < GRAD_MAXP; ++t) v[t] += (t == k) ? x : 0.0;
}

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(256) void eri_grad_kernel(EriArgs a)
{
    using Cls = EriClass<LA, LB, LC, LD, double>;
    using Adj = EriAdj<LA, LB, LC, LD>;
    constexpr int NA = qbx_ncart(LA), NB = qbx_ncart(LB), NC = qbx_ncart(LC), ND = qbx_ncart(LD);
    constexpr int LT = LA + LB + LC + LD;
    constexpr int NACC = Adj::NACC;
    constexpr int NINT = Adj::NINT;
    __shared__ double2 sboys[BOYS_HEAD2 + boys_rows(LT + 1) * QBX_BOYSL_STRIDE / 2];
    stage_boys<256, boys_rows(LT + 1)>(sboys, a.boysl);    // rows of order LT + 1
    __syncthreads();
    const double* sb = (const double*)sboys + BOYS_HEAD;

    const int lane = (int)threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)blockIdx.x * 4 + ((int)threadIdx.x >> 6));
    if (wave >= a.nwave) return;
    const cptr<int> cpre = cst(a.cpre);
    int lo = 0, hi = a.nbra;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (cpre[mid] <= wave) lo = mid; else hi = mid;
    }
    const int tb = lo, bi = tb / a.nbk;
    const int nk = (int)(cst(a.prefix)[tb + 1] - cst(a.prefix)[tb]);
    const int pos = (wave - cpre[tb]) * 64 + lane;
    const bool inrange = pos < nk;
    const int ki = cst(a.koff)[tb] + (inrange ? pos : 0);
    const int sA = cst(a.bra.shA)[bi], sB = cst(a.bra.shB)[bi];
    const int sC = a.ket.shA[ki], sD = a.ket.shB[ki];
    const int kn = inrange ? a.ket.ppn[ki] : 0;
    const bool live = kn > 0;
    if (__ballot(live) == 0) return;

    const cptr<DevShell> SA = cst(a.sh + sA);
    const cptr<DevShell> SB = cst(a.sh + sB);
    const DevShell* SC = a.sh + sC;
    const DevShell* SD = a.sh + sD;
    const size_t N = (size_t)a.N;
    const double* __restrict__ D = a.D;

    // ---- seeds on the contracted integrals ----
    double f = live ? 1.0 : 0.0;
    if (sA == sB) f *= 0.5;
    if (sC == sD) f *= 0.5;
    if (a.same && bi == ki) f *= 0.5;
    double G[NINT];
#pragma unroll
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
        for (int ib = 0; ib < NB; ++ib)
#pragma unroll
            for (int ic = 0; ic < NC; ++ic)
#pragma unroll
                for (int id = 0; id < ND; ++id) {
                    const int m = max(SA->bf[ia], 0), n = max(SB->bf[ib], 0);
                    const int l = max(SC->bf[ic], 0), s = max(SD->bf[id], 0);
                    const double sc = SA->sc[ia] * SB->sc[ib] * SC->sc[ic] * SD->sc[id];
                    double gam;
                    if (a.nset == 2) {
                        const double* Da = D + N * N;
                        const double* Db = D + 2 * N * N;
                        gam = 4.0 * D[m * N + n] * D[l * N + s] -
                              2.0 * (Da[m * N + l] * Da[n * N + s] + Db[m * N + l] * Db[n * N + s] +
                                     Da[m * N + s] * Da[n * N + l] + Db[m * N + s] * Db[n * N + l]);
                    } else {
                        gam = 4.0 * D[m * N + n] * D[l * N + s] - D[m * N + l] * D[n * N + s] -
                              D[m * N + s] * D[n * N + l];
                    }
                    G[((ia * NB + ib) * NC + ic) * ND + id] = f * sc * gam;
                }

    // ---- forward contracted pass ----
    const cptr<PrimPair> BP = cst(a.pp + cst(a.bra.ppo)[bi]);
    const int bn = cst(a.bra.ppn)[bi];
    const PrimPair* KP = a.pp + a.ket.ppo[ki];
    double acc[NACC];
#pragma unroll
    for (int n = 0; n < NACC; ++n) acc[n] = 0.0;
    for (int v = 0; v < kn; ++v) {
        const KetPrim Q = load_ket(KP + v);
        for (int u = 0; u < bn; ++u) {
            const double p = BP[u].p;
            const double PQx = BP[u].Px - Q.Qx, PQy = BP[u].Py - Q.Qy, PQz = BP[u].Pz - Q.Qz;
            const double s = rsqrt(p + Q.q), inv = s * s;
            const double T = p * Q.q * inv * (PQx * PQx + PQy * PQy + PQz * PQz);
            const double pref = 34.986836655249725693 * BP[u].Kp * Q.Kq * s;
            double Bm[LT + 2];
            boys_lds<LT + 1>(T, Bm, sb);
            double F[LT + 1];
#pragma unroll
            for (int m = 0; m <= LT; ++m) F[m] = Bm[m] * pref;
            const double rho_p = Q.q * inv, rho_q = p * inv;
            Cls::vrr(F, acc, BP[u].PAx, BP[u].PAy, BP[u].PAz, -rho_p * PQx, -rho_p * PQy, -rho_p * PQz,
                     Q.QCx, Q.QCy, Q.QCz, rho_q * PQx, rho_q * PQy, rho_q * PQz, BP[u].inv2p, Q.inv2q, 0.5 * inv,
                     rho_p, rho_q);
        }
    }

    // ---- transpose of the horizontal recurrence ----
    const cptr<double> ABp = cst(a.bra.AB + 3 * bi);
    const double AB[3] = {ABp[0], ABp[1], ABp[2]};
    const double CD[3] = {a.ket.AB[3 * ki], a.ket.AB[3 * ki + 1], a.ket.AB[3 * ki + 2]};
    double accb[NACC], bAB[3], bCD[3];
    Adj::hrr_adj(acc, G, accb, bAB, bCD, AB[0], AB[1], AB[2], CD[0], CD[1], CD[2]);

    double gxa[GRAD_MAXP], gxb[GRAD_MAXP], gxc[GRAD_MAXP], gxd[GRAD_MAXP];
    double gla[GRAD_MAXP], glb[GRAD_MAXP], glc[GRAD_MAXP], gld[GRAD_MAXP];
#pragma unroll
    for (int t = 0; t < GRAD_MAXP; ++t) {
        gxa[t] = gxb[t] = gxc[t] = gxd[t] = 0.0;
        gla[t] = glb[t] = glc[t] = gld[t] = 0.0;
    }
    double gA[3], gB[3], gC[3], gD[3];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        gA[t] = bAB[t];
        gB[t] = -bAB[t];
        gC[t] = bCD[t];
        gD[t] = -bCD[t];
    }

    // ---- adjoint primitive pass (pair quantities recomputed) ----
    const cptr<int> BK = cst(a.ppk + cst(a.bra.ppo)[bi]);
    const int* KK = a.ppk + a.ket.ppo[ki];
    for (int v = 0; v < kn; ++v) {
        const KetPrim Q = load_ket(KP + v);
        const double dx = KP[v].pad;
        const int kk = KK[v];
        const int kc = kk & 0xffff, kd = kk >> 16;
        for (int u = 0; u < bn; ++u) {
            const double p = BP[u].p, bx = BP[u].pad;
            const int kab = BK[u];
            const int ka = kab & 0xffff, kb = kab >> 16;
            const double PA[3] = {BP[u].PAx, BP[u].PAy, BP[u].PAz};
            const double QC[3] = {Q.QCx, Q.QCy, Q.QCz};
            const double PQ[3] = {BP[u].Px - Q.Qx, BP[u].Py - Q.Qy, BP[u].Pz - Q.Qz};
            const double s = rsqrt(p + Q.q), inv = s * s;
            const double T = p * Q.q * inv * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
            const double pref = 34.986836655249725693 * BP[u].Kp * Q.Kq * s;
            double Bm[LT + 2];
            boys_lds<LT + 1>(T, Bm, sb);
            double F[LT + 1];
#pragma unroll
            for (int m = 0; m <= LT; ++m) F[m] = Bm[m] * pref;
            const double rho_p = Q.q * inv, rho_q = p * inv;
            double Fb[LT + 1], bin[17];
            const double Lp = Adj::vrr_adj(F, accb, PA[0], PA[1], PA[2], -rho_p * PQ[0], -rho_p * PQ[1],
                                           -rho_p * PQ[2], QC[0], QC[1], QC[2], rho_q * PQ[0], rho_q * PQ[1],
                                           rho_q * PQ[2], BP[u].inv2p, Q.inv2q, 0.5 * inv, rho_p, rho_q, Fb, bin);
            const GeomAdj g = geom_adjoint<LT>(p - bx, bx, Q.q - dx, dx, PA, QC, AB, CD, PQ, pref, s, Bm, Fb, bin);
            sel_add(gxa, ka, g.da);
            sel_add(gxb, kb, g.db);
            sel_add(gxc, kc, g.dc);
            sel_add(gxd, kd, g.dd);
            sel_add(gla, ka, Lp);
            sel_add(glb, kb, Lp);
            sel_add(glc, kc, Lp);
            sel_add(gld, kd, Lp);
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                gA[t] += g.dA[t];
                gB[t] += g.dB[t];
                gC[t] += g.dC[t];
                gD[t] += g.dD[t];
            }
        }
    }

    // ---- flush: bra sums across the wave, ket sums per lane ----
    const int pA = cst(a.shell_poff)[sA], pB = cst(a.shell_poff)[sB];
    const int nA = SA->nprim, nB = SB->nprim;
#pragma unroll
    for (int t = 0; t < GRAD_MAXP; ++t) {
        const double x1 = wave_sum(gxa[t]), x2 = wave_sum(gla[t]);
        const double x3 = wave_sum(gxb[t]), x4 = wave_sum(glb[t]);
        if (lane == 0) {
            if (t < nA) {
                atomicAdd(a.g_xpn + pA + t, x1);
                atomicAdd(a.g_lin + pA + t, x2);
            }
            if (t < nB) {
                atomicAdd(a.g_xpn + pB + t, x3);
                atomicAdd(a.g_lin + pB + t, x4);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const double x1 = wave_sum(gA[t]), x2 = wave_sum(gB[t]);
        if (lane == 0) {
            atomicAdd(a.g_cen + 3 * sA + t, x1);
            atomicAdd(a.g_cen + 3 * sB + t, x2);
        }
    }
    if (live) {
        const int pC = a.shell_poff[sC], pD = a.shell_poff[sD];
        const int nC = SC->nprim, nD = SD->nprim;
#pragma unroll
        for (int t = 0; t < GRAD_MAXP; ++t) {
            if (t < nC) {
                atomicAdd(a.g_xpn + pC + t, gxc[t]);
                atomicAdd(a.g_lin + pC + t, glc[t]);
            }
            if (t < nD) {
                atomicAdd(a.g_xpn + pD + t, gxd[t]);
                atomicAdd(a.g_lin + pD + t, gld[t]);
            }
        }
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            atomicAdd(a.g_cen + 3 * sC + t, gC[t]);
            atomicAdd(a.g_cen + 3 * sD + t, gD[t]);
        }
    }
}

template <int LA, int LB, int LC, int LD>
void launch_eri_grad(const EriArgs& a, hipStream_t s)
{
    if (a.nwave <= 0) return;
    const int64_t blocks = ((int64_t)a.nwave + 3) / 4;
    hipLaunchKernelGGL((eri_grad_kernel<LA, LB, LC, LD>), dim3((unsigned)blocks), dim3(256), 0, s, a);
    QBX_HIP(hipGetLastError());
}

}  // namespace qbx
