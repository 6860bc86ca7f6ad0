// qbx_eri.h -- two-electron kernels of libqbx_hip.so (gfx950).
//
// One thread evaluates one contracted shell quartet of class (LA LB|LC LD):
// the primitive loop runs the generated vertical recurrence (EriClass::vrr)
// and sums [e0|f0] over primitive quartets; the generated horizontal
// recurrence (EriClass::hrr) then yields (ab|cd).  The epilogue depends on
// MODE: Fock digest (J/K with FP64 atomics), dense tensor, or Schwarz bound.
// Each class is instantiated in its own generated translation unit
// (gen/eri_<class>.hip) so that the classes compile in parallel.
#pragma once
#include "qbx.h"
#include "qbx_common.h"
#include "qbx_device.h"
#include "qbx_eri_api.h"

#include <cstdint>

template <int LA, int LB, int LC, int LD, typename R> struct EriClass;

namespace qbx {
template <int LA, int LB, int LC, int LD, typename R, int MODE>
__global__ __launch_bounds__(256) void eri_kernel(EriArgs a)
{
    using Cls = EriClass<LA, LB, LC, LD, R>;
    constexpr int NA = qbx_ncart(LA), NB = qbx_ncart(LB), NC = qbx_ncart(LC), ND = qbx_ncart(LD);
    constexpr int LT = LA + LB + LC + LD;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.total) return;
    int bi, ki;
    if (MODE == MODE_SCHWARZ) {
        bi = (int)t;
        ki = (int)t;
    } else {
        int lo = 0, hi = a.nbra;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (a.prefix[mid] <= t) lo = mid; else hi = mid;
        }
        bi = lo;
        ki = a.koff[lo] + (int)(t - a.prefix[lo]);
    }

    if (MODE == MODE_JK && a.pdmax) {
        // Haser-Ahlrichs bound with the densities of this build
        const int xA = a.bra.shA[bi], xB = a.bra.shB[bi], xC = a.ket.shA[ki], xD = a.ket.shB[ki];
        const double* dm = a.pdmax;
        const int ns = a.nshell;
        double d = fmax(fmax(dm[xA * ns + xB], dm[xC * ns + xD]), fmax(dm[xA * ns + xC], dm[xA * ns + xD]));
        d = fmax(d, fmax(dm[xB * ns + xC], dm[xB * ns + xD]));
        if (a.bra.Q[bi] * a.ket.Q[ki] * d < a.dthr) return;
    }

    R acc[Cls::NACC];
#pragma unroll
    for (int n = 0; n < Cls::NACC; ++n) acc[n] = (R)0;

    const int bo = a.bra.ppo[bi], bn = a.bra.ppn[bi];
    const int ko = a.ket.ppo[ki], kn = a.ket.ppn[ki];
    for (int u = 0; u < bn; ++u) {
        const PrimPair P = a.pp[bo + u];
        for (int v = 0; v < kn; ++v) {
            const PrimPair Q = a.pp[ko + v];
            const double pq = P.p + Q.p;
            const double s = rsqrt(pq);
            const double inv = s * s;
            const double PQx = P.Px - Q.Px, PQy = P.Py - Q.Py, PQz = P.Pz - Q.Pz;
            const double rho = P.p * Q.p * inv;
            const double T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
            // 2 pi^{5/2} (c c K / p)_{ab} (c c K / q)_{cd} / sqrt(p+q)
            const double pref = 34.986836655249725693 * P.Kp * Q.Kp * s;
            double Fd[LT + 1];
            qbx_boys<LT>(T, Fd, a.boys, a.bexp);
            R F[LT + 1];
#pragma unroll
            for (int m = 0; m <= LT; ++m) F[m] = (R)(Fd[m] * pref);
            const double rho_p = Q.p * inv, rho_q = P.p * inv;
            Cls::vrr(F, acc, (R)P.PAx, (R)P.PAy, (R)P.PAz,
                     (R)(-rho_p * PQx), (R)(-rho_p * PQy), (R)(-rho_p * PQz),
                     (R)Q.PAx, (R)Q.PAy, (R)Q.PAz,
                     (R)(rho_q * PQx), (R)(rho_q * PQy), (R)(rho_q * PQz),
                     (R)P.inv2p, (R)Q.inv2p, (R)(0.5 * inv), (R)rho_p, (R)rho_q);
        }
    }

    constexpr int NCD = NC * ND;
    const double* ABp = a.bra.AB + 3 * bi;
    const double* CDp = a.ket.AB + 3 * ki;
    R kacc[Cls::NE * NCD];
    Cls::hrr_ket(acc, kacc, (R)CDp[0], (R)CDp[1], (R)CDp[2]);
    const R ABx = (R)ABp[0], ABy = (R)ABp[1], ABz = (R)ABp[2];

    const int sA = a.bra.shA[bi], sB = a.bra.shB[bi];
    const int sC = a.ket.shA[ki], sD = a.ket.shB[ki];
    const DevShell& SA = a.sh[sA];
    const DevShell& SB = a.sh[sB];
    const DevShell& SC = a.sh[sC];
    const DevShell& SD = a.sh[sD];
    int fa[NA], fb[NB], fc[NC], fd[ND];
#pragma unroll
    for (int n = 0; n < NA; ++n) fa[n] = SA.bf[n];
#pragma unroll
    for (int n = 0; n < NB; ++n) fb[n] = SB.bf[n];
#pragma unroll
    for (int n = 0; n < NC; ++n) fc[n] = SC.bf[n];
#pragma unroll
    for (int n = 0; n < ND; ++n) fd[n] = SD.bf[n];
    // product of the ket component scales (weight ratios inside a shell)
    double scd[NCD];
#pragma unroll
    for (int ic = 0; ic < NC; ++ic)
#pragma unroll
        for (int id = 0; id < ND; ++id) scd[ic * ND + id] = SC.sc[ic] * SD.sc[id];

    if (MODE == MODE_SCHWARZ) {
        // bra == ket: the diagonal (ab|ab) elements of each slice a
        double mx = 0.0;
#pragma unroll
        for (int ia = 0; ia < NA; ++ia) {
            R out[NB * NCD];
            Cls::hrr_bra(ia, kacc, out, ABx, ABy, ABz);
#pragma unroll
            for (int ib = 0; ib < NB; ++ib) {
                const double w = SA.sc[ia] * SB.sc[ib];
                const double v = fabs((double)out[ib * NCD + ia * ND + ib] * w * w);
                mx = v > mx ? v : mx;
            }
        }
        a.qout[bi] = sqrt(mx);
        return;
    }

    const size_t N = (size_t)a.N;
    if (MODE == MODE_DENSE) {
#pragma unroll
        for (int ia = 0; ia < NA; ++ia) {
            R out[NB * NCD];
            Cls::hrr_bra(ia, kacc, out, ABx, ABy, ABz);
#pragma unroll
            for (int ib = 0; ib < NB; ++ib)
#pragma unroll
                for (int ic = 0; ic < NC; ++ic)
#pragma unroll
                    for (int id = 0; id < ND; ++id) {
                        if (fa[ia] < 0 || fb[ib] < 0 || fc[ic] < 0 || fd[id] < 0) continue;
                        const size_t i = fa[ia], j = fb[ib], k = fc[ic], l = fd[id];
                        const double v = (double)out[ib * NCD + ic * ND + id] * (SA.sc[ia] * SB.sc[ib]) *
                                         scd[ic * ND + id];
                        a.eri[((i * N + j) * N + k) * N + l] = v;
                        a.eri[((j * N + i) * N + k) * N + l] = v;
                        a.eri[((i * N + j) * N + l) * N + k] = v;
                        a.eri[((j * N + i) * N + l) * N + k] = v;
                        a.eri[((k * N + l) * N + i) * N + j] = v;
                        a.eri[((l * N + k) * N + i) * N + j] = v;
                        a.eri[((k * N + l) * N + j) * N + i] = v;
                        a.eri[((l * N + k) * N + j) * N + i] = v;
                    }
        }
        return;
    }

    // ---- MODE_JK: digest, one bra component a at a time, with the
    //      degeneracy of the unique shell quartet ----
    double f = 1.0;
    if (sA == sB) f *= 0.5;
    if (sC == sD) f *= 0.5;
    if (a.same && bi == ki) f *= 0.5;
    const double fj = 4.0 * f, fk = 2.0 * f;
    // missing components carry scale 0 and point at function 0
#pragma unroll
    for (int n = 0; n < NA; ++n) fa[n] = fa[n] < 0 ? 0 : fa[n];
#pragma unroll
    for (int n = 0; n < NB; ++n) fb[n] = fb[n] < 0 ? 0 : fb[n];
#pragma unroll
    for (int n = 0; n < NC; ++n) fc[n] = fc[n] < 0 ? 0 : fc[n];
#pragma unroll
    for (int n = 0; n < ND; ++n) fd[n] = fd[n] < 0 ? 0 : fd[n];

    for (int s = 0; s < a.nset; ++s) {
        const double* __restrict__ DJ = a.D + (size_t)s * N * N;
        const double* __restrict__ DK = a.D + (size_t)(a.nset + s) * N * N;
        double* Js = a.J + (size_t)s * N * N;
        double* Ks = a.K + (size_t)s * N * N;
        // blocks that every slice reads
        double dcd[NCD], dbc[NB * NC], dbd[NB * ND];
#pragma unroll
        for (int ic = 0; ic < NC; ++ic)
#pragma unroll
            for (int id = 0; id < ND; ++id) dcd[ic * ND + id] = DJ[fc[ic] * N + fd[id]];
#pragma unroll
        for (int ib = 0; ib < NB; ++ib)
#pragma unroll
            for (int ic = 0; ic < NC; ++ic) dbc[ib * NC + ic] = DK[fb[ib] * N + fc[ic]];
#pragma unroll
        for (int ib = 0; ib < NB; ++ib)
#pragma unroll
            for (int id = 0; id < ND; ++id) dbd[ib * ND + id] = DK[fb[ib] * N + fd[id]];
        // sums over all slices
        double jcd[NCD], kbd[NB * ND], kbc[NB * NC];
#pragma unroll
        for (int n = 0; n < NCD; ++n) jcd[n] = 0.0;
#pragma unroll
        for (int n = 0; n < NB * ND; ++n) kbd[n] = 0.0;
#pragma unroll
        for (int n = 0; n < NB * NC; ++n) kbc[n] = 0.0;
#pragma unroll
        for (int ia = 0; ia < NA; ++ia) {
            R out[NB * NCD];
            Cls::hrr_bra(ia, kacc, out, ABx, ABy, ABz);
            double dac[NC], dad[ND], kac[NC], kad[ND];
#pragma unroll
            for (int ic = 0; ic < NC; ++ic) { dac[ic] = DK[fa[ia] * N + fc[ic]]; kac[ic] = 0.0; }
#pragma unroll
            for (int id = 0; id < ND; ++id) { dad[id] = DK[fa[ia] * N + fd[id]]; kad[id] = 0.0; }
#pragma unroll
            for (int ib = 0; ib < NB; ++ib) {
                const double sab = SA.sc[ia] * SB.sc[ib];
                const double dab = DJ[fa[ia] * N + fb[ib]];
                double jab = 0.0;
#pragma unroll
                for (int ic = 0; ic < NC; ++ic)
#pragma unroll
                    for (int id = 0; id < ND; ++id) {
                        const int cd = ic * ND + id;
                        const double v = (double)out[ib * NCD + cd] * (sab * scd[cd]);
                        jab += v * dcd[cd];
                        jcd[cd] += v * dab;
                        kad[id] += v * dbc[ib * NC + ic];
                        kac[ic] += v * dbd[ib * ND + id];
                        kbd[ib * ND + id] += v * dac[ic];
                        kbc[ib * NC + ic] += v * dad[id];
                    }
                atomicAdd(Js + fa[ia] * N + fb[ib], fj * jab);
            }
#pragma unroll
            for (int id = 0; id < ND; ++id) atomicAdd(Ks + fa[ia] * N + fd[id], fk * kad[id]);
#pragma unroll
            for (int ic = 0; ic < NC; ++ic) atomicAdd(Ks + fa[ia] * N + fc[ic], fk * kac[ic]);
        }
#pragma unroll
        for (int ic = 0; ic < NC; ++ic)
#pragma unroll
            for (int id = 0; id < ND; ++id) atomicAdd(Js + fc[ic] * N + fd[id], fj * jcd[ic * ND + id]);
#pragma unroll
        for (int ib = 0; ib < NB; ++ib)
#pragma unroll
            for (int id = 0; id < ND; ++id) atomicAdd(Ks + fb[ib] * N + fd[id], fk * kbd[ib * ND + id]);
#pragma unroll
        for (int ib = 0; ib < NB; ++ib)
#pragma unroll
            for (int ic = 0; ic < NC; ++ic) atomicAdd(Ks + fb[ib] * N + fc[ic], fk * kbc[ib * NC + ic]);
    }
}


template <int LA, int LB, int LC, int LD, typename R, int MODE>
void launch_eri(const EriArgs& a, hipStream_t s)
{
    if (a.total <= 0) return;
    // Kernels that spill to scratch run on the caller's main stream only: each
    // extra hardware queue would need its own scratch allocation.
    static const size_t scratch = [] {
        hipFuncAttributes fa;
        if (hipFuncGetAttributes(&fa, (const void*)eri_kernel<LA, LB, LC, LD, R, MODE>) != hipSuccess) return (size_t)1;
        return (size_t)fa.localSizeBytes;
    }();
    hipStream_t use = (scratch > 0 && a.scratch_stream) ? (hipStream_t)a.scratch_stream : s;
    const int threads = 256;
    const int64_t blocks = (a.total + threads - 1) / threads;
    hipLaunchKernelGGL((eri_kernel<LA, LB, LC, LD, R, MODE>), dim3((unsigned)blocks), dim3(threads), 0, use, a);
    QBX_HIP(hipGetLastError());
}

}  // namespace qbx
