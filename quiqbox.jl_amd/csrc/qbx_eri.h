// qbx_eri.h -- two-electron kernels of libqbx_hip.so (gfx950).
//
// Work mapping (Fock digest and dense modes): one wavefront takes one bra
// shell pair and up to 64 of its screened kets, one ket per lane.  The bra is
// therefore wave-uniform: its primitive pairs, shells and density elements
// come through scalar loads (constant address space) and sit in SGPRs, and
// the J_ab contribution is summed across the wave before a single atomic.
// Per lane, the primitive loop runs over the ket's primitive pairs (outer,
// loaded once, next one prefetched) and the bra's (inner, scalar); each
// primitive quartet evaluates the Boys function from the class's table rows
// staged in LDS and runs the generated vertical recurrence (EriClass::vrr),
// summing [e0|f0] over primitive quartets; the generated horizontal
// recurrence then yields (ab|cd).  The epilogue depends on MODE: Fock digest
// (J/K with FP64 atomics), dense tensor, or Schwarz bound (one thread per
// diagonal pair, its own kernel).  Each class is instantiated in its own
// generated translation unit (gen/eri_<class>.hip) so that the classes
// compile in parallel.
#pragma once
#include "qbx.h"
#include "qbx_common.h"
#include "qbx_device.h"
#include "qbx_eri_api.h"

#include <cstdint>
#include <type_traits>
#include <utility>

template <int LA, int LB, int LC, int LD, typename R> struct EriClass;

namespace qbx {

// Loads through the constant address space become scalar (SMEM) loads when the
// address is wave-uniform.  Only read-only data of the launch goes through it.
template <typename T> using cptr = const __attribute__((address_space(4))) T*;
template <typename T> __device__ __forceinline__ cptr<T> cst(const T* p) { return (cptr<T>)p; }

// J/K accumulation.  With per-XCD copies every writer of a copy runs on one
// XCD, so the add only has to be atomic within that XCD's L2: workgroup scope
// keeps it there instead of sending each 8-byte add out to memory.
struct JKSink {
    double* J;
    double* K;
    bool local;
    __device__ __forceinline__ JKSink(const EriArgs& a)
    {
        local = a.xcd_local != 0;
        const size_t off = local ? (size_t)qbx_xcc_id() * a.jk_stride : 0;
        J = a.J + off;
        K = a.K + off;
    }
    __device__ __forceinline__ void add(double* p, double v) const
    {
        if (local) __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        else atomicAdd(p, v);
    }
};

// Threads per block: classes with few accumulators (and so few registers) run
// 8 or 16 waves per block, so that one LDS copy of the Boys rows (23 KB, six
// copies fit a CU) serves more waves.
__host__ __device__ constexpr int eri_tpb(int nacc) { return nacc <= 3 ? 1024 : nacc <= 18 ? 512 : 256; }

// Boys table extent per order.  Beyond T_L = 36, 40, 44, 47, 50, 52, 55, 57
// and 60 (L = 0..8) the asymptotic forms F_0 = sqrt(pi/4T), F_{m+1} =
// F_m (2m+1)/2T are exact to double precision for every order m <= L
// (relative error below 4e-17 from the first grid point past 35.5, 39.75,
// 43.25, 46.25, 49, 51.75, 54.25, 56.75 and 59 against 40-digit values,
// falling as e^{-T} beyond), so a kernel of total angular momentum L stages
// only the rows below T_L: 14 KB of LDS instead of 23 KB at L = 0, more
// resident blocks to hide the latency of the lookup.
constexpr double kBoysTmax[QBX_BOYSL_LMAX + 1] = {36.0, 40.0, 44.0, 47.0, 50.0, 52.0, 55.0, 57.0, 60.0, 60.0};
constexpr double boys_tmax(int L) { return kBoysTmax[L]; }
constexpr int boys_rows(int L) { return (int)(kBoysTmax[L] * QBX_BOYSL_INVDT) + 2; }
static_assert(kBoysTmax[QBX_BOYSL_LMAX] <= QBX_BOYS_TMAX && boys_rows(QBX_BOYSL_LMAX) <= QBX_BOYSL_ROWS,
              "the per-order extent must stay inside the table");

// LDS copy of the Boys table: BOYS_HEAD doubles of constants (1/k! for
// k = 11..0, the Taylor coefficients of e^x), then the rows; kernels hold
// sb = table + BOYS_HEAD.  The constants come from LDS rather than from the
// code: as literals the compiler kept them in VGPRs and copied one into the
// accumulator before each Horner step of e^x (v_mov_b64 + v_fmac_f64, 22
// VALU instructions per primitive quartet instead of 11), while a loaded value
// is consumed in place.
constexpr int BOYS_HEAD = 12;
constexpr int BOYS_HEAD2 = BOYS_HEAD / 2;    // the head in double2

// Boys F_0..F_L from the LDS rows of order L (see qbx_device.h for the table):
// an 11-term Taylor expansion of F_L about the nearest grid point, e^{-T} from
// e^{-T_i} times the Taylor series of e^{T_i - T} (|T_i - T| <= 1/8, terms to
// 11th order), the downward recurrence for the lower orders, and the
// asymptotic form above the table.
template <int L>
__device__ __forceinline__ void boys_lds(double T, double* __restrict__ F, const double* __restrict__ sb)
{
    if (T < boys_tmax(L)) {
        const int i = __double2int_rn(T * QBX_BOYSL_INVDT);
        const double md = (double)i * QBX_BOYSL_DT - T;
        const double2* row = (const double2*)(sb + i * QBX_BOYSL_STRIDE);
        const double2 c01 = row[0], c23 = row[1], c45 = row[2], c67 = row[3], c89 = row[4], c10e = row[5];
        double acc = fma(md, c10e.x, c89.y);
        acc = fma(md, acc, c89.x);
        acc = fma(md, acc, c67.y);
        acc = fma(md, acc, c67.x);
        acc = fma(md, acc, c45.y);
        acc = fma(md, acc, c45.x);
        acc = fma(md, acc, c23.y);
        acc = fma(md, acc, c23.x);
        acc = fma(md, acc, c01.y);
        acc = fma(md, acc, c01.x);
        F[L] = acc;
        if (L > 0) {
            const double2* kh = (const double2*)(sb - BOYS_HEAD);
            const double2 k01 = kh[0], k23 = kh[1], k45 = kh[2], k67 = kh[3], k89 = kh[4], kab = kh[5];
            double ex = fma(md, k01.x, k01.y);
            ex = fma(md, ex, k23.x);
            ex = fma(md, ex, k23.y);
            ex = fma(md, ex, k45.x);
            ex = fma(md, ex, k45.y);
            ex = fma(md, ex, k67.x);
            ex = fma(md, ex, k67.y);
            ex = fma(md, ex, k89.x);
            ex = fma(md, ex, k89.y);
            ex = fma(md, ex, kab.x);
            ex = fma(md, ex, kab.y);
            const double e = c10e.y * ex;
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

// Copy the head constants and the Boys rows of the class's total angular
// momentum into LDS (dst holds BOYS_HEAD2 + ROWS * QBX_BOYSL_STRIDE / 2 double2).
template <int TPB, int ROWS = QBX_BOYSL_ROWS>
__device__ __forceinline__ void stage_boys(double2* __restrict__ dst, const double* __restrict__ src)
{
    if (threadIdx.x < BOYS_HEAD2) {
        // 1/k! for k = 11 - 2t and 10 - 2t
        const int k = 11 - 2 * (int)threadIdx.x;
        double f = 1.0;
        for (int j = 2; j < k; ++j) f *= j;
        dst[threadIdx.x] = make_double2(1.0 / (f * k), 1.0 / f);
    }
    const double2* s = (const double2*)src;
    for (int i = threadIdx.x; i < ROWS * QBX_BOYSL_STRIDE / 2; i += TPB) dst[BOYS_HEAD2 + i] = s[i];
}

// 1/sqrt(x) for finite x > 0 (p + q of a primitive quartet): the hardware
// estimate r refined by the second-order term of the binomial series,
// 1/sqrt(x) = r (1 - e)^(-1/2) = r (1 + e/2 + 3e^2/8 + ...) with e = 1 - x r^2.
// The library rsqrt() applies the same correction and then re-checks the
// estimate for zero, infinite and NaN inputs, which cannot occur here.
__device__ __forceinline__ double rsqrt_pos(double x)
{
    const double r = __builtin_amdgcn_rsq(x);
    const double e = fma(r * -x, r, 1.0);
    return fma(r * e, fma(e, 0.375, 0.5), r);
}

// The ket-pair quantities one primitive quartet reads.
struct KetPrim {
    double q, inv2q, Qx, Qy, Qz, QCx, QCy, QCz, Kq;
};

__device__ __forceinline__ KetPrim load_ket(const PrimPair* __restrict__ p)
{
    KetPrim k;
    k.q = p->p;
    k.inv2q = p->inv2p;
    k.Qx = p->Px;
    k.Qy = p->Py;
    k.Qz = p->Pz;
    k.QCx = p->PAx;
    k.QCy = p->PAy;
    k.QCz = p->PAz;
    k.Kq = p->Kp;
    return k;
}

// Sum over the primitive quartets of one contracted shell quartet: bra pairs
// BP[0..bn), ket pairs KP[0..kn).  BraPtr is a constant-address-space pointer
// when the bra is wave-uniform (scalar loads), a plain pointer otherwise.
// GRP < 0 runs the full recurrence, GRP >= 0 that group of a staged class.
template <class Cls, int LT, typename R, bool PREFETCH, int GRP, typename BraPtr>
__device__ __forceinline__ void prim_loop(R* __restrict__ acc, BraPtr BP, int bn,
                                          const PrimPair* __restrict__ KP, int kn,
                                          const double* __restrict__ sb)
{
    KetPrim nxt;
    if (PREFETCH && kn > 0) nxt = load_ket(KP);
    for (int v = 0; v < kn; ++v) {
        KetPrim Q;
        if (PREFETCH) {
            Q = nxt;
            if (v + 1 < kn) nxt = load_ket(KP + v + 1);
        } else {
            Q = load_ket(KP + v);
        }
        for (int u = 0; u < bn; ++u) {
            const double p = BP[u].p, inv2p = BP[u].inv2p, Px = BP[u].Px, Py = BP[u].Py, Pz = BP[u].Pz;
            const double PAx = BP[u].PAx, PAy = BP[u].PAy, PAz = BP[u].PAz, Kp = BP[u].Kp;
            const double pq = p + Q.q;
            const double s = rsqrt_pos(pq);
            const double inv = s * s;
            const double PQx = Px - Q.Qx, PQy = Py - Q.Qy, PQz = Pz - Q.Qz;
            const double rho = p * Q.q * inv;
            const double T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
            // 2 pi^{5/2} (c c K / p)_{ab} (c c K / q)_{cd} / sqrt(p+q)
            const double pref = 34.986836655249725693 * Kp * Q.Kq * s;
            double Fd[LT + 1];
            boys_lds<LT>(T, Fd, sb);
            R F[LT + 1];
#pragma unroll
            for (int m = 0; m <= LT; ++m) F[m] = (R)(Fd[m] * pref);
            const double rho_p = Q.q * inv, rho_q = p * inv;
            if constexpr (GRP < 0)
                Cls::vrr(F, acc, (R)PAx, (R)PAy, (R)PAz,
                         (R)(-rho_p * PQx), (R)(-rho_p * PQy), (R)(-rho_p * PQz),
                         (R)Q.QCx, (R)Q.QCy, (R)Q.QCz,
                         (R)(rho_q * PQx), (R)(rho_q * PQy), (R)(rho_q * PQz),
                         (R)inv2p, (R)Q.inv2q, (R)(0.5 * inv), (R)rho_p, (R)rho_q);
            else
                Cls::template vrr_grp<GRP>(F, acc, (R)PAx, (R)PAy, (R)PAz,
                         (R)(-rho_p * PQx), (R)(-rho_p * PQy), (R)(-rho_p * PQz),
                         (R)Q.QCx, (R)Q.QCy, (R)Q.QCz,
                         (R)(rho_q * PQx), (R)(rho_q * PQy), (R)(rho_q * PQz),
                         (R)inv2p, (R)Q.inv2q, (R)(0.5 * inv), (R)rho_p, (R)rho_q);
        }
    }
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// J/K digest and dense tensor: one wave per (bra, chunk of 64 kets).
template <int LA, int LB, int LC, int LD, typename R, int MODE>
__global__ __launch_bounds__(eri_tpb(EriClass<LA, LB, LC, LD, R>::NACC)) void eri_kernel(EriArgs a)
{
    using Cls = EriClass<LA, LB, LC, LD, R>;
    constexpr int TPB = eri_tpb(Cls::NACC);
    constexpr int NA = qbx_ncart(LA), NB = qbx_ncart(LB), NC = qbx_ncart(LC), ND = qbx_ncart(LD);
    constexpr int LT = LA + LB + LC + LD;
    constexpr int NCD = NC * ND;
    constexpr bool PREFETCH = Cls::NACC <= 18;
    __shared__ double2 sboys[BOYS_HEAD2 + boys_rows(LT) * QBX_BOYSL_STRIDE / 2];
    stage_boys<TPB, boys_rows(LT)>(sboys, a.boysl);
    __syncthreads();
    const double* sb = (const double*)sboys + BOYS_HEAD;

    const int lane = (int)threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)blockIdx.x * (TPB / 64) + ((int)threadIdx.x >> 6));
    if (wave >= a.nwave) return;
    // bra: the last one whose first wave is <= this wave (bras without kets own no wave)
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
    int kn = inrange ? a.ket.ppn[ki] : 0;
    if (MODE == MODE_JK && a.pdmax) {
        // Haser-Ahlrichs bound with the densities of this build
        const double* dm = a.pdmax;
        const int ns = a.nshell;
        const double dab = cst(dm)[sA * ns + sB];
        double d = fmax(fmax(dab, dm[sC * ns + sD]), fmax(dm[sA * ns + sC], dm[sA * ns + sD]));
        d = fmax(d, fmax(dm[sB * ns + sC], dm[sB * ns + sD]));
        if (cst(a.bra.Q)[bi] * a.ket.Q[ki] * d < a.dthr) kn = 0;
    }
    if (__ballot(kn > 0) == 0) return;
    const bool live = kn > 0;

    R acc[Cls::NACC];
#pragma unroll
    for (int n = 0; n < Cls::NACC; ++n) acc[n] = (R)0;
    prim_loop<Cls, LT, R, PREFETCH, -1>(acc, cst(a.pp + cst(a.bra.ppo)[bi]), cst(a.bra.ppn)[bi],
                                    a.pp + a.ket.ppo[ki], kn, sb);

    const cptr<double> ABp = cst(a.bra.AB + 3 * bi);
    const double* CDp = a.ket.AB + 3 * ki;
    R kacc[Cls::NE * NCD];
    Cls::hrr_ket(acc, kacc, (R)CDp[0], (R)CDp[1], (R)CDp[2]);
    const R ABx = (R)ABp[0], ABy = (R)ABp[1], ABz = (R)ABp[2];

    const cptr<DevShell> SA = cst(a.sh + sA);
    const cptr<DevShell> SB = cst(a.sh + sB);
    const DevShell* SC = a.sh + sC;
    const DevShell* SD = a.sh + sD;
    int fc[NC], fd[ND];
#pragma unroll
    for (int n = 0; n < NC; ++n) fc[n] = SC->bf[n];
#pragma unroll
    for (int n = 0; n < ND; ++n) fd[n] = SD->bf[n];
    // product of the ket component scales (weight ratios inside a shell)
    double scd[NCD];
#pragma unroll
    for (int ic = 0; ic < NC; ++ic)
#pragma unroll
        for (int id = 0; id < ND; ++id) scd[ic * ND + id] = SC->sc[ic] * SD->sc[id];

    const size_t N = (size_t)a.N;
    if (MODE == MODE_DENSE) {
        if (!live) return;
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
                        if (SA->bf[ia] < 0 || SB->bf[ib] < 0 || fc[ic] < 0 || fd[id] < 0) continue;
                        const size_t i = SA->bf[ia], j = SB->bf[ib], k = fc[ic], l = fd[id];
                        const double v = (double)out[ib * NCD + ic * ND + id] * (SA->sc[ia] * SB->sc[ib]) *
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
    double f = live ? 1.0 : 0.0;
    if (sA == sB) f *= 0.5;
    if (sC == sD) f *= 0.5;
    if (a.same && bi == ki) f *= 0.5;
    const double fj = 4.0 * f, fk = 2.0 * f;
    // missing components carry scale 0 and point at function 0
    int fa[NA], fb[NB];
#pragma unroll
    for (int n = 0; n < NA; ++n) fa[n] = SA->bf[n] < 0 ? 0 : SA->bf[n];
#pragma unroll
    for (int n = 0; n < NB; ++n) fb[n] = SB->bf[n] < 0 ? 0 : SB->bf[n];
#pragma unroll
    for (int n = 0; n < NC; ++n) fc[n] = fc[n] < 0 ? 0 : fc[n];
#pragma unroll
    for (int n = 0; n < ND; ++n) fd[n] = fd[n] < 0 ? 0 : fd[n];

    const JKSink sink(a);
    for (int s = 0; s < a.nset; ++s) {
        const double* __restrict__ DJ = a.D + (size_t)s * N * N;
        const double* __restrict__ DK = a.D + (size_t)(a.nset + s) * N * N;
        double* Js = sink.J + (size_t)s * N * N;
        double* Ks = sink.K + (size_t)s * N * N;
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
                const double sab = SA->sc[ia] * SB->sc[ib];
                const double dab = cst(DJ)[fa[ia] * N + fb[ib]];
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
                // the bra is wave-uniform: one atomic per wave for J_ab
                jab = wave_sum(fj * jab);
                if (lane == 0) sink.add(Js + fa[ia] * N + fb[ib], jab);
            }
            if (live) {
#pragma unroll
                for (int id = 0; id < ND; ++id) sink.add(Ks + fa[ia] * N + fd[id], fk * kad[id]);
#pragma unroll
                for (int ic = 0; ic < NC; ++ic) sink.add(Ks + fa[ia] * N + fc[ic], fk * kac[ic]);
            }
        }
        if (live) {
#pragma unroll
            for (int ic = 0; ic < NC; ++ic)
#pragma unroll
                for (int id = 0; id < ND; ++id) sink.add(Js + fc[ic] * N + fd[id], fj * jcd[ic * ND + id]);
#pragma unroll
            for (int ib = 0; ib < NB; ++ib)
#pragma unroll
                for (int id = 0; id < ND; ++id) sink.add(Ks + fb[ib] * N + fd[id], fk * kbd[ib * ND + id]);
#pragma unroll
            for (int ib = 0; ib < NB; ++ib)
#pragma unroll
                for (int ic = 0; ic < NC; ++ic) sink.add(Ks + fb[ib] * N + fc[ic], fk * kbc[ib * NC + ic]);
        }
    }
}


// K destinations of the staged digest: straight to the (per-XCD) K matrix,
// or to the block's LDS rows of the bra's functions (row = a component, or
// NA + b component).
struct KToGlobal {
    const JKSink& sink;
    size_t N;
    __device__ __forceinline__ void add(int s, int fn, int, int col, double v) const
    {
        sink.add(sink.K + (size_t)s * N * N + (size_t)fn * N + col, v);
    }
};
template <int NROW> struct KToRows {
    double* rows;
    int Ni;
    __device__ __forceinline__ void add(int s, int, int lrow, int col, double v) const
    {
        atomicAdd(rows + (s * NROW + lrow) * Ni + col, v);
    }
};

// Flush of a block's K rows ([nset][NA + NB][N] in LDS) to K: each nonzero
// element once; with `clear` the rows are zeroed for the block's next bra.
template <int NA, int TPB>
__device__ __forceinline__ void flush_rows(double* krow, int nrowel, int nrow, int Ni, const int* bfA,
                                           const int* bfB, const JKSink& sink, bool clear)
{
    const size_t N = (size_t)Ni;
    for (int i = threadIdx.x; i < nrowel; i += TPB) {
        const double v = krow[i];
        if (v == 0.0) continue;
        if (clear) krow[i] = 0.0;
        const int s = i / (nrow * Ni), rem = i - s * nrow * Ni;
        const int row = rem / Ni, col = rem - row * Ni;
        const int fr = row < NA ? max(bfA[row], 0) : max(bfB[row - NA], 0);
        sink.add(sink.K + (size_t)s * N * N + (size_t)fr * N + col, v);
    }
}

// J/K digest of one wave's contracted quartets for a wave-uniform bra:
// kacc holds [e|cd) after the ket transfer.  J_ab is reduced across the wave,
// J_cd goes out per lane through the sink, and the K updates (each with one
// index on the bra) go to the LDS rows rows0 + set * rstride, laid out [A and
// B functions][N].  f is the quartet's degeneracy weight (0 for dead lanes).
template <int LA, int LB, int LC, int LD, typename R>
__device__ __forceinline__ void digest_rows(const EriArgs& a, const R* kacc, const int* fa, const int* fb,
                                            cptr<DevShell> SA, cptr<DevShell> SB, R ABx, R ABy, R ABz, int sC,
                                            int sD, double f, bool live, double* rows0, int rstride,
                                            const JKSink& sink, int lane)
{
    using Cls = EriClass<LA, LB, LC, LD, R>;
    constexpr int NA = qbx_ncart(LA), NB = qbx_ncart(LB), NC = qbx_ncart(LC), ND = qbx_ncart(LD);
    constexpr int NCD = NC * ND;
    const int Ni = a.N;
    const size_t N = (size_t)a.N;
    const DevShell* SC = a.sh + sC;
    const DevShell* SD = a.sh + sD;
    int fc[NC], fd[ND];
    double scd[NCD];
#pragma unroll
    for (int n = 0; n < NC; ++n) fc[n] = SC->bf[n] < 0 ? 0 : SC->bf[n];
#pragma unroll
    for (int n = 0; n < ND; ++n) fd[n] = SD->bf[n] < 0 ? 0 : SD->bf[n];
#pragma unroll
    for (int ic = 0; ic < NC; ++ic)
#pragma unroll
        for (int id = 0; id < ND; ++id) scd[ic * ND + id] = SC->sc[ic] * SD->sc[id];
    const double fj = 4.0 * f, fk = 2.0 * f;

    for (int s = 0; s < a.nset; ++s) {
        const double* __restrict__ DJ = a.D + (size_t)s * N * N;
        const double* __restrict__ DK = a.D + (size_t)(a.nset + s) * N * N;
        double* Js = sink.J + (size_t)s * N * N;
        double* rows = rows0 + s * rstride;
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
                const double sab = SA->sc[ia] * SB->sc[ib];
                const double dab = cst(DJ)[fa[ia] * N + fb[ib]];
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
                jab = wave_sum(fj * jab);
                if (lane == 0) sink.add(Js + fa[ia] * N + fb[ib], jab);
            }
            if (live) {
#pragma unroll
                for (int id = 0; id < ND; ++id) atomicAdd(rows + ia * Ni + fd[id], fk * kad[id]);
#pragma unroll
                for (int ic = 0; ic < NC; ++ic) atomicAdd(rows + ia * Ni + fc[ic], fk * kac[ic]);
            }
        }
        if (live) {
#pragma unroll
            for (int ic = 0; ic < NC; ++ic)
#pragma unroll
                for (int id = 0; id < ND; ++id) sink.add(Js + fc[ic] * N + fd[id], fj * jcd[ic * ND + id]);
#pragma unroll
            for (int ib = 0; ib < NB; ++ib)
#pragma unroll
                for (int id = 0; id < ND; ++id) atomicAdd(rows + (NA + ib) * Ni + fd[id], fk * kbd[ib * ND + id]);
#pragma unroll
            for (int ib = 0; ib < NB; ++ib)
#pragma unroll
                for (int ic = 0; ic < NC; ++ic) atomicAdd(rows + (NA + ib) * Ni + fc[ic], fk * kbc[ib * NC + ic]);
        }
    }
}

// ---------------------------------------------------------------------------
// J/K with the bra's K rows in LDS.  A block of ROW_WAVES waves takes one bra
// and up to ROW_CHUNKS of its 64-ket chunks.  Every K update of a quartet has
// one index on the bra (K_ac, K_ad on the rows of A's functions; K_bc, K_bd on
// B's), so the block sums them in LDS rows [set][A and B functions][N] with
// LDS atomics and sends each nonzero row element to J/K once, instead of one
// global atomic per quartet and element.  J_ab is reduced across the wave and
// J_cd goes out per lane, as in eri_kernel.
// ---------------------------------------------------------------------------
template <int LA, int LB, int LC, int LD, typename R>
__device__ __forceinline__ void jk_rows_body(const EriArgs& a);

template <int LA, int LB, int LC, int LD, typename R>
__global__ __launch_bounds__(64 * ROW_WAVES) void eri_jk_rows_kernel(EriArgs a)
{
    jk_rows_body<LA, LB, LC, LD, R>(a);
}

// The same kernel held to two waves per SIMD (EriClass::ROWS_TWO_WAVES): 256
// registers a wave, the rest spilled to scratch, where the compiler would
// otherwise run one wave with AGPRs.
template <int LA, int LB, int LC, int LD, typename R>
__global__ __launch_bounds__(64 * ROW_WAVES) __attribute__((amdgpu_waves_per_eu(2, 2)))
void eri_jk_rows2_kernel(EriArgs a)
{
    jk_rows_body<LA, LB, LC, LD, R>(a);
}

// the K-row kernel of a class (only that one is instantiated)
template <int LA, int LB, int LC, int LD, typename R>
constexpr auto rows_kernel()
{
    if constexpr (EriClass<LA, LB, LC, LD, R>::ROWS_TWO_WAVES) return &eri_jk_rows2_kernel<LA, LB, LC, LD, R>;
    else return &eri_jk_rows_kernel<LA, LB, LC, LD, R>;
}

template <int LA, int LB, int LC, int LD, typename R>
__device__ __forceinline__ void jk_rows_body(const EriArgs& a)
{
    using Cls = EriClass<LA, LB, LC, LD, R>;
    constexpr int TPB = 64 * ROW_WAVES;
    constexpr int NA = qbx_ncart(LA), NB = qbx_ncart(LB), NC = qbx_ncart(LC), ND = qbx_ncart(LD);
    constexpr int LT = LA + LB + LC + LD;
    constexpr int NCD = NC * ND;
    constexpr int NROW = NA + NB;
    constexpr bool PREFETCH = Cls::NACC <= 18;
    __shared__ double2 sboys[BOYS_HEAD2 + boys_rows(LT) * QBX_BOYSL_STRIDE / 2];
    extern __shared__ double krow[];    // [nset][NROW][N]
    const int Ni = a.N;
    const size_t N = (size_t)a.N;
    const int nrowel = a.nset * NROW * Ni;
    stage_boys<TPB, boys_rows(LT)>(sboys, a.boysl);
    for (int i = threadIdx.x; i < nrowel; i += TPB) krow[i] = 0.0;
    __syncthreads();
    const double* sb = (const double*)sboys + BOYS_HEAD;

    const int lane = (int)threadIdx.x & 63;
    const int wib = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    // bra and superchunk of this block
    const cptr<int> spre = cst(a.spre);
    const int task = (int)blockIdx.x;
    int lo = 0, hi = a.nbra;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (spre[mid] <= task) lo = mid; else hi = mid;
    }
    const int tb = lo, bi = tb / a.nbk;
    const int sc = task - spre[tb];
    const int nk = (int)(cst(a.prefix)[tb + 1] - cst(a.prefix)[tb]);
    const int k0 = cst(a.koff)[tb];
    const int sA = cst(a.bra.shA)[bi], sB = cst(a.bra.shB)[bi];
    const cptr<DevShell> SA = cst(a.sh + sA);
    const cptr<DevShell> SB = cst(a.sh + sB);
    const cptr<double> ABp = cst(a.bra.AB + 3 * bi);
    const R ABx = (R)ABp[0], ABy = (R)ABp[1], ABz = (R)ABp[2];
    int fa[NA], fb[NB];
#pragma unroll
    for (int n = 0; n < NA; ++n) fa[n] = SA->bf[n] < 0 ? 0 : SA->bf[n];
#pragma unroll
    for (int n = 0; n < NB; ++n) fb[n] = SB->bf[n] < 0 ? 0 : SB->bf[n];
    const JKSink sink(a);

    for (int r = 0; r < ROW_CHUNKS / ROW_WAVES; ++r) {
        const int chunk = (sc * (ROW_CHUNKS / ROW_WAVES) + r) * ROW_WAVES + wib;
        if (chunk * 64 >= nk) break;
        const int pos = chunk * 64 + lane;
        const bool inrange = pos < nk;
        const int ki = k0 + (inrange ? pos : 0);
        const int sC = a.ket.shA[ki], sD = a.ket.shB[ki];
        int kn = inrange ? a.ket.ppn[ki] : 0;
        if (a.pdmax) {
            const double* dm = a.pdmax;
            const int ns = a.nshell;
            const double dab = cst(dm)[sA * ns + sB];
            double d = fmax(fmax(dab, dm[sC * ns + sD]), fmax(dm[sA * ns + sC], dm[sA * ns + sD]));
            d = fmax(d, fmax(dm[sB * ns + sC], dm[sB * ns + sD]));
            if (cst(a.bra.Q)[bi] * a.ket.Q[ki] * d < a.dthr) kn = 0;
        }
        if (__ballot(kn > 0) == 0) continue;
        const bool live = kn > 0;

        R acc[Cls::NACC];
#pragma unroll
        for (int n = 0; n < Cls::NACC; ++n) acc[n] = (R)0;
        prim_loop<Cls, LT, R, PREFETCH, -1>(acc, cst(a.pp + cst(a.bra.ppo)[bi]), cst(a.bra.ppn)[bi],
                                            a.pp + a.ket.ppo[ki], kn, sb);
        const double* CDp = a.ket.AB + 3 * ki;
        R kacc[Cls::NE * NCD];
        Cls::hrr_ket(acc, kacc, (R)CDp[0], (R)CDp[1], (R)CDp[2]);

        double f = live ? 1.0 : 0.0;
        if (sA == sB) f *= 0.5;
        if (sC == sD) f *= 0.5;
        if (a.same && bi == ki) f *= 0.5;
        digest_rows<LA, LB, LC, LD, R>(a, kacc, fa, fb, SA, SB, ABx, ABy, ABz, sC, sD, f, live, krow, NROW * Ni,
                                       sink, lane);
    }
    __syncthreads();
    flush_rows<NA, TPB>(krow, nrowel, NROW, Ni, a.sh[sA].bf, a.sh[sB].bf, sink, false);
}

// ---------------------------------------------------------------------------
// J/K for a group of twin bras.  A block of TWIN_WAVES waves takes one group
// and up to ROW_CHUNKS 64-ket chunks of the union of its members' ket ranges
// (each member's range is a prefix of the same sorted ket list, or its FP32
// tail).  The geometry and Boys values of every primitive quartet are computed
// once, from the first member's primitive pairs; each member accumulates them
// with its own c_a c_b factor (Kp) where the ket lies in its range.  Each
// member then goes through the digest with its own shells and LDS K rows
// [set][member][A and B functions][N].
// ---------------------------------------------------------------------------
template <int LA, int LB, int LC, int LD, typename R>
__global__ __launch_bounds__(64 * TWIN_WAVES) void eri_jk_twin_kernel(EriArgs a)
{
    using Cls = EriClass<LA, LB, LC, LD, R>;
    constexpr int TPB = 64 * TWIN_WAVES;
    constexpr int NA = qbx_ncart(LA), NB = qbx_ncart(LB);
    constexpr int LT = LA + LB + LC + LD;
    constexpr int NCD = qbx_ncart(LC) * qbx_ncart(LD);
    constexpr int NROW = NA + NB;
    constexpr int TW = TWIN_MAX;
    static_assert(ROW_CHUNKS % TWIN_WAVES == 0, "chunks per task must split evenly over the waves");
    __shared__ double2 sboys[BOYS_HEAD2 + boys_rows(LT) * QBX_BOYSL_STRIDE / 2];
    extern __shared__ double krow[];    // [nset][twmax][NROW][N]
    const int Ni = a.N;
    const int tw = a.twmax;
    const int nrowel = a.nset * tw * NROW * Ni;
    stage_boys<TPB, boys_rows(LT)>(sboys, a.boysl);
    for (int i = threadIdx.x; i < nrowel; i += TPB) krow[i] = 0.0;
    __syncthreads();
    const double* sb = (const double*)sboys + BOYS_HEAD;

    const int lane = (int)threadIdx.x & 63;
    const int wib = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const cptr<int> gpre = cst(a.gpre);
    const int task = (int)blockIdx.x;
    int lo = 0, hi = a.ngroup;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (gpre[mid] <= task) lo = mid; else hi = mid;
    }
    const int g = lo;
    const int sc = task - gpre[g];
    const cptr<int> gm = cst(a.gmem + 3 * TW * g);
    int bi[TW], klo[TW], khi[TW], sA[TW], sB[TW], ppo[TW];
    int ulo = 0x7fffffff, uhi = 0;
#pragma unroll
    for (int t = 0; t < TW; ++t) {
        bi[t] = t < tw ? gm[3 * t] : -1;
        klo[t] = t < tw ? gm[3 * t + 1] : 0;
        khi[t] = t < tw ? gm[3 * t + 2] : 0;
        const int b = bi[t] < 0 ? bi[0] : bi[t];
        sA[t] = cst(a.bra.shA)[b];
        sB[t] = cst(a.bra.shB)[b];
        ppo[t] = cst(a.bra.ppo)[b];
        if (bi[t] >= 0 && khi[t] > klo[t]) {
            ulo = min(ulo, klo[t]);
            uhi = max(uhi, khi[t]);
        }
    }
    const int bn = cst(a.bra.ppn)[bi[0]];
    const cptr<PrimPair> BP = cst(a.pp + ppo[0]);
    const cptr<double> ABp = cst(a.bra.AB + 3 * bi[0]);
    const R ABx = (R)ABp[0], ABy = (R)ABp[1], ABz = (R)ABp[2];
    const JKSink sink(a);

    for (int r = 0; r < ROW_CHUNKS / TWIN_WAVES; ++r) {
        const int chunk = (sc * (ROW_CHUNKS / TWIN_WAVES) + r) * TWIN_WAVES + wib;
        if (ulo >= uhi || ulo + chunk * 64 >= uhi) break;
        const int pos = ulo + chunk * 64 + lane;
        const bool inrange = pos < uhi;
        const int ki = inrange ? pos : ulo;
        const int sC = a.ket.shA[ki], sD = a.ket.shB[ki];
        bool live[TW];
        bool any = false;
#pragma unroll
        for (int t = 0; t < TW; ++t) {
            bool l = inrange && bi[t] >= 0 && pos >= klo[t] && pos < khi[t];
            if (l && a.pdmax) {
                const double* dm = a.pdmax;
                const int ns = a.nshell;
                double d = fmax(fmax(dm[sA[t] * ns + sB[t]], dm[sC * ns + sD]),
                                fmax(dm[sA[t] * ns + sC], dm[sA[t] * ns + sD]));
                d = fmax(d, fmax(dm[sB[t] * ns + sC], dm[sB[t] * ns + sD]));
                if (cst(a.bra.Q)[bi[t]] * a.ket.Q[ki] * d < a.dthr) l = false;
            }
            live[t] = l;
            any = any || l;
        }
        const int kn = any ? a.ket.ppn[ki] : 0;
        if (__ballot(kn > 0) == 0) continue;

        R acc[TW][Cls::NACC];
#pragma unroll
        for (int t = 0; t < TW; ++t)
#pragma unroll
            for (int n = 0; n < Cls::NACC; ++n) acc[t][n] = (R)0;
        const PrimPair* KP = a.pp + a.ket.ppo[ki];
        for (int v = 0; v < kn; ++v) {
            const KetPrim Q = load_ket(KP + v);
            for (int u = 0; u < bn; ++u) {
                const double p = BP[u].p, inv2p = BP[u].inv2p;
                const double PQx = BP[u].Px - Q.Qx, PQy = BP[u].Py - Q.Qy, PQz = BP[u].Pz - Q.Qz;
                const double pq = p + Q.q;
                const double s = rsqrt_pos(pq);
                const double inv = s * s;
                const double rho = p * Q.q * inv;
                const double T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
                const double pref0 = 34.986836655249725693 * Q.Kq * s;
                double Fd[LT + 1];
                boys_lds<LT>(T, Fd, sb);
                const double rho_p = Q.q * inv, rho_q = p * inv;
#pragma unroll
                for (int t = 0; t < TW; ++t) {
                    if (bi[t] < 0) continue;
                    const double w = live[t] ? pref0 * cst(a.pp + ppo[t])[u].Kp : 0.0;
                    R F[LT + 1];
#pragma unroll
                    for (int m = 0; m <= LT; ++m) F[m] = (R)(Fd[m] * w);
                    Cls::vrr(F, acc[t], (R)BP[u].PAx, (R)BP[u].PAy, (R)BP[u].PAz,
                             (R)(-rho_p * PQx), (R)(-rho_p * PQy), (R)(-rho_p * PQz),
                             (R)Q.QCx, (R)Q.QCy, (R)Q.QCz,
                             (R)(rho_q * PQx), (R)(rho_q * PQy), (R)(rho_q * PQz),
                             (R)inv2p, (R)Q.inv2q, (R)(0.5 * inv), (R)rho_p, (R)rho_q);
                }
            }
        }
        const double* CDp = a.ket.AB + 3 * ki;
#pragma unroll
        for (int t = 0; t < TW; ++t) {
            if (bi[t] < 0 || __ballot(live[t]) == 0) continue;
            R kacc[Cls::NE * NCD];
            Cls::hrr_ket(acc[t], kacc, (R)CDp[0], (R)CDp[1], (R)CDp[2]);
            const cptr<DevShell> SA = cst(a.sh + sA[t]);
            const cptr<DevShell> SB = cst(a.sh + sB[t]);
            int fa[NA], fb[NB];
#pragma unroll
            for (int n = 0; n < NA; ++n) fa[n] = SA->bf[n] < 0 ? 0 : SA->bf[n];
#pragma unroll
            for (int n = 0; n < NB; ++n) fb[n] = SB->bf[n] < 0 ? 0 : SB->bf[n];
            double f = live[t] ? 1.0 : 0.0;
            if (sA[t] == sB[t]) f *= 0.5;
            if (sC == sD) f *= 0.5;
            if (a.same && bi[t] == ki) f *= 0.5;
            digest_rows<LA, LB, LC, LD, R>(a, kacc, fa, fb, SA, SB, ABx, ABy, ABz, sC, sD, f, live[t],
                                           krow + t * NROW * Ni, tw * NROW * Ni, sink, lane);
        }
    }
    __syncthreads();
    // each nonzero row element to K once
    const size_t N = (size_t)Ni;
    for (int i = threadIdx.x; i < nrowel; i += TPB) {
        const double v = krow[i];
        if (v == 0.0) continue;
        const int s = i / (tw * NROW * Ni);
        int rem = i - s * tw * NROW * Ni;
        const int t = rem / (NROW * Ni);
        rem -= t * NROW * Ni;
        const int row = rem / Ni, col = rem - row * Ni;
        const int b = gm[3 * t];
        const int sh = row < NA ? cst(a.bra.shA)[b] : cst(a.bra.shB)[b];
        const int fr = max(a.sh[sh].bf[row < NA ? row : row - NA], 0);
        sink.add(sink.K + (size_t)s * N * N + (size_t)fr * N + col, v);
    }
}

// ---------------------------------------------------------------------------
// Staged classes: the [e0|f0] set is too large for the register file, so each
// group of bra components e makes its own pass over the primitive quartets
// (the intermediates shared between groups are recomputed), transfers its
// part to the ket and stores [e|cd) to the wave's workspace slot, laid out
// [value][lane] so that every access is one coalesced 512-byte row.  The
// epilogue then forms one (a b| row at a time from the workspace and
// digests it.  The grid is persistent: STAGE_SLOTS waves at most, each
// owning one workspace slot and looping over the wave tasks.
// ---------------------------------------------------------------------------
constexpr int STAGE_WAVES_PER_BLOCK = 4;

// One wave's workspace slot through raw buffer instructions: the descriptor
// (slot base, slot size in bytes) sits in SGPRs, the lane offset in one VGPR,
// and value v's row offset v * 64 * sizeof(R) is a scalar immediate, so no
// per-value 64-bit address is ever formed.  Accesses past the slot are
// dropped by the descriptor's range check.
template <typename R> struct WsSlot {
    __amdgpu_buffer_rsrc_t rsrc;
    int voff;
    __device__ __forceinline__ WsSlot(R* base, int nvalues, int lane)
        : rsrc(__builtin_amdgcn_make_buffer_rsrc(base, 0, nvalues * 64 * (int)sizeof(R), 0x00020000)),
          voff(lane * (int)sizeof(R)) {}
    __device__ __forceinline__ R ld(int v) const
    {
        if constexpr (sizeof(R) == 8)
            return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b64(rsrc, voff, v * 64 * 8, 0));
        else
            return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b32(rsrc, voff, v * 64 * 4, 0));
    }
    __device__ __forceinline__ void st(int v, R x) const
    {
        if constexpr (sizeof(R) == 8)
            __builtin_amdgcn_raw_buffer_store_b64(
                __builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, x), rsrc, voff, v * 64 * 8, 0);
        else
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), rsrc, voff, v * 64 * 4, 0);
    }
};

template <class Cls, int LT, typename R, int G, typename BraPtr>
__device__ __forceinline__ void staged_group(BraPtr BP, int bn, const PrimPair* __restrict__ KP, int kn,
                                             const double* __restrict__ sb, R CDx, R CDy, R CDz,
                                             const WsSlot<R>& ws)
{
    constexpr int NF = Cls::NACC / Cls::NE;
    constexpr int NG = (Cls::ge0(G + 1) - Cls::ge0(G)) * NF;
    R acc[NG];
#pragma unroll
    for (int n = 0; n < NG; ++n) acc[n] = (R)0;
    prim_loop<Cls, LT, R, false, G>(acc, BP, bn, KP, kn, sb);
    Cls::template hrr_ket_grp<G>(acc, ws, CDx, CDy, CDz);
}

// The groups of one wave task; with W > 1 waves sharing the task, wave wib
// (wave-uniform) evaluates groups wib, wib + W, ...
template <class Cls, int LT, typename R, int W, typename BraPtr, int... Gs>
__device__ __forceinline__ void staged_groups(std::integer_sequence<int, Gs...>, int wib, BraPtr BP, int bn,
                                              const PrimPair* __restrict__ KP, int kn,
                                              const double* __restrict__ sb, R CDx, R CDy, R CDz,
                                              const WsSlot<R>& ws)
{
    ((W == 1 || Cls::grp_wave(Gs) == wib ? staged_group<Cls, LT, R, Gs>(BP, bn, KP, kn, sb, CDx, CDy, CDz, ws)
                                          : (void)0), ...);
}

// One wave task of a staged class: bra `bi`, kets chunk*64 .. chunk*64+63 of
// its list.  Evaluates the groups into the workspace slot, then writes the
// dense tensor or digests J/K (J through the sink, K through `kout`).  With
// W > 1 the W waves of a block share the task and its slot (wave wib takes
// every W-th group, then one W-th of the (a b| rows of the digest); every
// wave of the block must call with the same task.
template <int LA, int LB, int LC, int LD, typename R, int MODE, class KOut, int W = 1>
__device__ __forceinline__ void staged_chunk(const EriArgs& a, int tb, int chunk, int lane, const WsSlot<R>& ws,
                                             const double* sb, const KOut& kout, int wib = 0)
{
    using Cls = EriClass<LA, LB, LC, LD, R>;
    constexpr int NA = qbx_ncart(LA), NB = qbx_ncart(LB), NC = qbx_ncart(LC), ND = qbx_ncart(LD);
    constexpr int LT = LA + LB + LC + LD;
    constexpr int NCD = NC * ND;
    const size_t N = (size_t)a.N;
    const int bi = tb / a.nbk;
    const int nk = (int)(cst(a.prefix)[tb + 1] - cst(a.prefix)[tb]);
    const int pos = chunk * 64 + lane;
    const bool inrange = pos < nk;
    const int ki = cst(a.koff)[tb] + (inrange ? pos : 0);
    const int sA = cst(a.bra.shA)[bi], sB = cst(a.bra.shB)[bi];
    const int sC = a.ket.shA[ki], sD = a.ket.shB[ki];
    int kn = inrange ? a.ket.ppn[ki] : 0;
    if (MODE == MODE_JK && a.pdmax) {
        const double* dm = a.pdmax;
        const int ns = a.nshell;
        const double dab = cst(dm)[sA * ns + sB];
        double d = fmax(fmax(dab, dm[sC * ns + sD]), fmax(dm[sA * ns + sC], dm[sA * ns + sD]));
        d = fmax(d, fmax(dm[sB * ns + sC], dm[sB * ns + sD]));
        if (cst(a.bra.Q)[bi] * a.ket.Q[ki] * d < a.dthr) kn = 0;
    }
    if (__ballot(kn > 0) == 0) return;
    const bool live = kn > 0;

    const double* CDp = a.ket.AB + 3 * ki;
    staged_groups<Cls, LT, R, W>(std::make_integer_sequence<int, Cls::NGRP>{}, wib,
                                 cst(a.pp + cst(a.bra.ppo)[bi]), cst(a.bra.ppn)[bi],
                                 a.pp + a.ket.ppo[ki], kn, sb, (R)CDp[0], (R)CDp[1], (R)CDp[2], ws);
    if constexpr (W > 1) {
        static_assert(MODE == MODE_JK, "split staged tasks digest J/K only");
        __syncthreads();    // every group's [e|cd) is in the slot
    }

    const cptr<double> ABp = cst(a.bra.AB + 3 * bi);
    const R ABx = (R)ABp[0], ABy = (R)ABp[1], ABz = (R)ABp[2];
    const cptr<DevShell> SA = cst(a.sh + sA);
    const cptr<DevShell> SB = cst(a.sh + sB);
    const DevShell* SC = a.sh + sC;
    const DevShell* SD = a.sh + sD;
    int fc[NC], fd[ND];
    double scc[NC], scdd[ND];
#pragma unroll
    for (int n = 0; n < NC; ++n) { fc[n] = SC->bf[n]; scc[n] = SC->sc[n]; }
#pragma unroll
    for (int n = 0; n < ND; ++n) { fd[n] = SD->bf[n]; scdd[n] = SD->sc[n]; }

    if (MODE == MODE_DENSE) {
        if (!live) return;
#pragma unroll 1
        for (int iab = 0; iab < NA * NB; ++iab) {
            const int ia = iab / NB, ib = iab - ia * NB;
            R out[NCD];
            Cls::hrr_bra_ab(iab, ws, out, ABx, ABy, ABz);
            const int fa_ = SA->bf[ia], fb_ = SB->bf[ib];
            if (fa_ < 0 || fb_ < 0) continue;
            const double sab = SA->sc[ia] * SB->sc[ib];
            const size_t i = fa_, j = fb_;
#pragma unroll
            for (int ic = 0; ic < NC; ++ic)
#pragma unroll
                for (int id = 0; id < ND; ++id) {
                    if (fc[ic] < 0 || fd[id] < 0) continue;
                    const size_t k = fc[ic], l = fd[id];
                    const double v = (double)out[ic * ND + id] * sab * (scc[ic] * scdd[id]);
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

    // ---- J/K digest from the (a b| rows of [r0, r1), in two passes that each keep
    //      few sums live: A) J_ab, K_ac, K_ad per row; B) J_cd summed over
    //      all rows, K_bd and K_bc summed over a.  The density elements a
    //      lane reads are loaded once per quartet (its D_cd block), per B
    //      component (D_bc, D_bd) or per row (D_ac, D_ad) rather than once per
    //      integral ----
    double f = live ? 1.0 : 0.0;
    if (sA == sB) f *= 0.5;
    if (sC == sD) f *= 0.5;
    if (a.same && bi == ki) f *= 0.5;
    const double fj = 4.0 * f, fk = 2.0 * f;
#pragma unroll
    for (int n = 0; n < NC; ++n) fc[n] = fc[n] < 0 ? 0 : fc[n];
#pragma unroll
    for (int n = 0; n < ND; ++n) fd[n] = fd[n] < 0 ? 0 : fd[n];
    // rows t = ib * NA + ia in [r0, r1): all of them, or this wave's share
    constexpr int NR = NA * NB;
    const int r0 = W == 1 ? 0 : wib * NR / W, r1 = W == 1 ? NR : (wib + 1) * NR / W;
    const JKSink sink(a);
    for (int s = 0; s < a.nset; ++s) {
        const double* __restrict__ DJ = a.D + (size_t)s * N * N;
        const double* __restrict__ DK = a.D + (size_t)(a.nset + s) * N * N;
        double* Js = sink.J + (size_t)s * N * N;
        double dcd[NCD];
#pragma unroll
        for (int ic = 0; ic < NC; ++ic)
#pragma unroll
            for (int id = 0; id < ND; ++id) dcd[ic * ND + id] = DJ[fc[ic] * N + fd[id]];
        double dbc[NC], dbd[ND];
#pragma unroll 1
        for (int t = r0; t < r1; ++t) {
            const int ib = t / NA, ia = t - ib * NA;
            if (t == r0 || ia == 0) {
                const int fb_ = max(SB->bf[ib], 0);
#pragma unroll
                for (int ic = 0; ic < NC; ++ic) dbc[ic] = DK[fb_ * N + fc[ic]];
#pragma unroll
                for (int id = 0; id < ND; ++id) dbd[id] = DK[fb_ * N + fd[id]];
            }
            R out[NCD];
            Cls::hrr_bra_ab(ia * NB + ib, ws, out, ABx, ABy, ABz);
            const int fa_ = max(SA->bf[ia], 0), fb_ = max(SB->bf[ib], 0);
            const double sab = SA->sc[ia] * SB->sc[ib];
            double jab = 0.0, kac[NC], kad[ND];
#pragma unroll
            for (int n = 0; n < NC; ++n) kac[n] = 0.0;
#pragma unroll
            for (int n = 0; n < ND; ++n) kad[n] = 0.0;
#pragma unroll
            for (int ic = 0; ic < NC; ++ic)
#pragma unroll
                for (int id = 0; id < ND; ++id) {
                    const double v = (double)out[ic * ND + id] * sab * (scc[ic] * scdd[id]);
                    jab += v * dcd[ic * ND + id];
                    kac[ic] += v * dbd[id];
                    kad[id] += v * dbc[ic];
                }
            jab = wave_sum(fj * jab);
            if (lane == 0) sink.add(Js + fa_ * N + fb_, jab);
            if (live) {
#pragma unroll
                for (int ic = 0; ic < NC; ++ic) kout.add(s, fa_, ia, fc[ic], fk * kac[ic]);
#pragma unroll
                for (int id = 0; id < ND; ++id) kout.add(s, fa_, ia, fd[id], fk * kad[id]);
            }
        }
        double jcd[NCD], kbd[ND], kbc[NC];
#pragma unroll
        for (int n = 0; n < NCD; ++n) jcd[n] = 0.0;
#pragma unroll
        for (int n = 0; n < ND; ++n) kbd[n] = 0.0;
#pragma unroll
        for (int n = 0; n < NC; ++n) kbc[n] = 0.0;
#pragma unroll 1
        for (int t = r0; t < r1; ++t) {
            const int ib = t / NA, ia = t - ib * NA;
            R out[NCD];
            Cls::hrr_bra_ab(ia * NB + ib, ws, out, ABx, ABy, ABz);
            const int fa_ = max(SA->bf[ia], 0), fb_ = max(SB->bf[ib], 0);
            const double sab = SA->sc[ia] * SB->sc[ib];
            const double dab = cst(DJ)[fa_ * N + fb_];
            double dac[NC], dad[ND];
#pragma unroll
            for (int ic = 0; ic < NC; ++ic) dac[ic] = DK[fa_ * N + fc[ic]];
#pragma unroll
            for (int id = 0; id < ND; ++id) dad[id] = DK[fa_ * N + fd[id]];
#pragma unroll
            for (int ic = 0; ic < NC; ++ic)
#pragma unroll
                for (int id = 0; id < ND; ++id) {
                    const int cd = ic * ND + id;
                    const double v = (double)out[cd] * sab * (scc[ic] * scdd[id]);
                    jcd[cd] += v * dab;
                    kbd[id] += v * dac[ic];
                    kbc[ic] += v * dad[id];
                }
            // K_bd, K_bc are complete for this b once its last row of the range is in
            if (ia == NA - 1 || t == r1 - 1) {
                if (live) {
#pragma unroll
                    for (int id = 0; id < ND; ++id) kout.add(s, fb_, NA + ib, fd[id], fk * kbd[id]);
#pragma unroll
                    for (int ic = 0; ic < NC; ++ic) kout.add(s, fb_, NA + ib, fc[ic], fk * kbc[ic]);
                }
#pragma unroll
                for (int n = 0; n < ND; ++n) kbd[n] = 0.0;
#pragma unroll
                for (int n = 0; n < NC; ++n) kbc[n] = 0.0;
            }
        }
        if (live && r1 > r0) {
#pragma unroll
            for (int ic = 0; ic < NC; ++ic)
#pragma unroll
                for (int id = 0; id < ND; ++id) sink.add(Js + fc[ic] * N + fd[id], fj * jcd[ic * ND + id]);
        }
    }
}

template <int LA, int LB, int LC, int LD, typename R, int MODE>
__global__ __launch_bounds__(64 * STAGE_WAVES_PER_BLOCK) __attribute__((amdgpu_waves_per_eu(EriClass<LA, LB, LC, LD, R>::WAVES_PER_SIMD, EriClass<LA, LB, LC, LD, R>::WAVES_PER_SIMD)))
void eri_staged_kernel(EriArgs a)
{
    using Cls = EriClass<LA, LB, LC, LD, R>;
    constexpr int TPB = 64 * STAGE_WAVES_PER_BLOCK;
    constexpr int NCD = qbx_ncart(LC) * qbx_ncart(LD);
    constexpr int WSV = Cls::NE * NCD;          // workspace values per lane
    __shared__ double2 sboys[BOYS_HEAD2 + QBX_BOYSL_ROWS * QBX_BOYSL_STRIDE / 2];
    stage_boys<TPB>(sboys, a.boysl);
    __syncthreads();
    const double* sb = (const double*)sboys + BOYS_HEAD;

    const int lane = (int)threadIdx.x & 63;
    const int slot = __builtin_amdgcn_readfirstlane((int)blockIdx.x * STAGE_WAVES_PER_BLOCK + ((int)threadIdx.x >> 6));
    const int nslot = (int)gridDim.x * STAGE_WAVES_PER_BLOCK;
    const WsSlot<R> ws((R*)a.ws + (size_t)slot * WSV * 64, WSV, lane);
    const cptr<int> cpre = cst(a.cpre);
    const JKSink sink(a);
    const KToGlobal kout{sink, (size_t)a.N};

    for (int wave = slot; wave < a.nwave; wave += nslot) {
        int lo = 0, hi = a.nbra;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (cpre[mid] <= wave) lo = mid; else hi = mid;
        }
        staged_chunk<LA, LB, LC, LD, R, MODE>(a, lo, wave - cpre[lo], lane, ws, sb, kout);
    }
}

// Staged J/K with each wave task shared by the Cls::SPLIT waves of a block
// (classes with few tasks: their latency, not the throughput of a full chip,
// sets the time).  A persistent grid of blocks, one workspace slot each.
template <int LA, int LB, int LC, int LD, typename R>
__global__ __launch_bounds__((64 * EriClass<LA, LB, LC, LD, R>::SPLIT)) __attribute__((amdgpu_waves_per_eu(EriClass<LA, LB, LC, LD, R>::WAVES_PER_SIMD, EriClass<LA, LB, LC, LD, R>::WAVES_PER_SIMD)))
void eri_staged_split_kernel(EriArgs a)
{
    using Cls = EriClass<LA, LB, LC, LD, R>;
    constexpr int W = Cls::SPLIT;
    constexpr int TPB = 64 * W;
    constexpr int NCD = qbx_ncart(LC) * qbx_ncart(LD);
    constexpr int WSV = Cls::NE * NCD;
    __shared__ double2 sboys[BOYS_HEAD2 + QBX_BOYSL_ROWS * QBX_BOYSL_STRIDE / 2];
    stage_boys<TPB>(sboys, a.boysl);
    __syncthreads();
    const double* sb = (const double*)sboys + BOYS_HEAD;

    const int lane = (int)threadIdx.x & 63;
    const int wib = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const WsSlot<R> ws((R*)a.ws + (size_t)blockIdx.x * WSV * 64, WSV, lane);
    const cptr<int> cpre = cst(a.cpre);
    const JKSink sink(a);
    const KToGlobal kout{sink, (size_t)a.N};

    for (int task = (int)blockIdx.x; task < a.nwave; task += (int)gridDim.x) {
        int lo = 0, hi = a.nbra;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (cpre[mid] <= task) lo = mid; else hi = mid;
        }
        staged_chunk<LA, LB, LC, LD, R, MODE_JK, KToGlobal, W>(a, lo, task - cpre[lo], lane, ws, sb, kout, wib);
        __syncthreads();    // the slot is rewritten by the next task
    }
}

// Staged J/K with the bra's K rows in LDS: a persistent grid of blocks, each
// taking one (bra, up to a.tchunks chunks) task at a time; its waves share
// the task's chunks, K goes to the rows, and the rows are flushed to K once
// per task.
template <int LA, int LB, int LC, int LD, typename R>
__global__ __launch_bounds__(64 * STAGE_WAVES_PER_BLOCK) __attribute__((amdgpu_waves_per_eu(EriClass<LA, LB, LC, LD, R>::WAVES_PER_SIMD, EriClass<LA, LB, LC, LD, R>::WAVES_PER_SIMD)))
void eri_staged_rows_kernel(EriArgs a)
{
    using Cls = EriClass<LA, LB, LC, LD, R>;
    constexpr int W = STAGE_WAVES_PER_BLOCK;
    constexpr int TPB = 64 * W;
    constexpr int NA = qbx_ncart(LA), NB = qbx_ncart(LB);
    constexpr int NROW = NA + NB;
    constexpr int NCD = qbx_ncart(LC) * qbx_ncart(LD);
    constexpr int WSV = Cls::NE * NCD;
    __shared__ double2 sboys[BOYS_HEAD2 + QBX_BOYSL_ROWS * QBX_BOYSL_STRIDE / 2];
    extern __shared__ double krow[];    // [nset][NROW][N]
    const int Ni = a.N;
    const int nrowel = a.nset * NROW * Ni;
    stage_boys<TPB>(sboys, a.boysl);
    for (int i = threadIdx.x; i < nrowel; i += TPB) krow[i] = 0.0;
    __syncthreads();
    const double* sb = (const double*)sboys + BOYS_HEAD;

    const int lane = (int)threadIdx.x & 63;
    const int wib = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const WsSlot<R> ws((R*)a.ws + (size_t)((int)blockIdx.x * W + wib) * WSV * 64, WSV, lane);
    const cptr<int> spre = cst(a.spre);
    const JKSink sink(a);
    const KToRows<NROW> kout{krow, Ni};

    for (int task = (int)blockIdx.x; task < a.nsuper; task += (int)gridDim.x) {
        int lo = 0, hi = a.nbra;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (spre[mid] <= task) lo = mid; else hi = mid;
        }
        const int tb = lo, bi = tb / a.nbk;
        const int sc = task - spre[tb];
        const int nk = (int)(cst(a.prefix)[tb + 1] - cst(a.prefix)[tb]);
        // wave wib takes chunks wib, wib + W, ... of the task's a.tchunks
        for (int j = wib; j < a.tchunks; j += W) {
            const int chunk = sc * a.tchunks + j;
            if (chunk * 64 >= nk) break;
            staged_chunk<LA, LB, LC, LD, R, MODE_JK>(a, tb, chunk, lane, ws, sb, kout);
        }
        __syncthreads();
        const int sA = cst(a.bra.shA)[bi], sB = cst(a.bra.shB)[bi];
        flush_rows<NA, TPB>(krow, nrowel, NROW, Ni, a.sh[sA].bf, a.sh[sB].bf, sink, true);
        __syncthreads();
    }
}

// Staged J/K with the bra's K rows in LDS and each 64-ket chunk shared by
// the Cls::SPLIT waves of a block (groups, then digest rows, spread over the
// waves; one workspace slot per block).  A task is one (bra, up to
// a.tchunks chunks) range, its chunks taken one after another by the
// whole block, so no wave idles when a bra has fewer chunks than the block
// has waves; the rows are flushed to K once per task.
template <int LA, int LB, int LC, int LD, typename R>
__global__ __launch_bounds__((64 * EriClass<LA, LB, LC, LD, R>::SPLIT)) __attribute__((amdgpu_waves_per_eu(EriClass<LA, LB, LC, LD, R>::WAVES_PER_SIMD, EriClass<LA, LB, LC, LD, R>::WAVES_PER_SIMD)))
void eri_staged_split_rows_kernel(EriArgs a)
{
    using Cls = EriClass<LA, LB, LC, LD, R>;
    constexpr int W = Cls::SPLIT;
    constexpr int TPB = 64 * W;
    constexpr int NA = qbx_ncart(LA), NB = qbx_ncart(LB);
    constexpr int NROW = NA + NB;
    constexpr int NCD = qbx_ncart(LC) * qbx_ncart(LD);
    constexpr int WSV = Cls::NE * NCD;
    __shared__ double2 sboys[BOYS_HEAD2 + QBX_BOYSL_ROWS * QBX_BOYSL_STRIDE / 2];
    extern __shared__ double krow[];    // [nset][NROW][N]
    const int Ni = a.N;
    const int nrowel = a.nset * NROW * Ni;
    stage_boys<TPB>(sboys, a.boysl);
    for (int i = threadIdx.x; i < nrowel; i += TPB) krow[i] = 0.0;
    __syncthreads();
    const double* sb = (const double*)sboys + BOYS_HEAD;

    const int lane = (int)threadIdx.x & 63;
    const int wib = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const WsSlot<R> ws((R*)a.ws + (size_t)blockIdx.x * WSV * 64, WSV, lane);
    const cptr<int> spre = cst(a.spre);
    const JKSink sink(a);
    const KToRows<NROW> kout{krow, Ni};

    for (int task = (int)blockIdx.x; task < a.nsuper; task += (int)gridDim.x) {
        int lo = 0, hi = a.nbra;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (spre[mid] <= task) lo = mid; else hi = mid;
        }
        const int tb = lo, bi = tb / a.nbk;
        const int sc = task - spre[tb];
        const int nk = (int)(cst(a.prefix)[tb + 1] - cst(a.prefix)[tb]);
        for (int r = 0; r < a.tchunks; ++r) {
            const int chunk = sc * a.tchunks + r;
            if (chunk * 64 >= nk) break;
            staged_chunk<LA, LB, LC, LD, R, MODE_JK, KToRows<NROW>, W>(a, tb, chunk, lane, ws, sb, kout, wib);
            __syncthreads();    // the slot is rewritten by the next chunk
        }
        const int sA = cst(a.bra.shA)[bi], sB = cst(a.bra.shB)[bi];
        flush_rows<NA, TPB>(krow, nrowel, NROW, Ni, a.sh[sA].bf, a.sh[sB].bf, sink, true);
        __syncthreads();
    }
}

// Schwarz bounds: one thread per diagonal shell pair (ab|ab).
template <int LA, int LB, typename R>
__global__ __launch_bounds__(256) void eri_schwarz_kernel(EriArgs a)
{
    using Cls = EriClass<LA, LB, LA, LB, R>;
    constexpr int NA = qbx_ncart(LA), NB = qbx_ncart(LB);
    constexpr int LT = 2 * (LA + LB);
    constexpr int NCD = NA * NB;
    __shared__ double2 sboys[BOYS_HEAD2 + boys_rows(LT) * QBX_BOYSL_STRIDE / 2];
    stage_boys<256, boys_rows(LT)>(sboys, a.boysl);
    __syncthreads();
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.total) return;
    const int bi = (int)t;
    R acc[Cls::NACC];
#pragma unroll
    for (int n = 0; n < Cls::NACC; ++n) acc[n] = (R)0;
    const PrimPair* P = a.pp + a.bra.ppo[bi];
    const int n = a.bra.ppn[bi];
    prim_loop<Cls, LT, R, false, -1>(acc, P, n, P, n, (const double*)sboys + BOYS_HEAD);
    const double* AB = a.bra.AB + 3 * bi;
    R kacc[Cls::NE * NCD];
    Cls::hrr_ket(acc, kacc, (R)AB[0], (R)AB[1], (R)AB[2]);
    const DevShell& SA = a.sh[a.bra.shA[bi]];
    const DevShell& SB = a.sh[a.bra.shB[bi]];
    double mx = 0.0;
#pragma unroll
    for (int ia = 0; ia < NA; ++ia) {
        R out[NB * NCD];
        Cls::hrr_bra(ia, kacc, out, (R)AB[0], (R)AB[1], (R)AB[2]);
#pragma unroll
        for (int ib = 0; ib < NB; ++ib) {
            const double w = SA.sc[ia] * SB.sc[ib];
            const double v = fabs((double)out[ib * NCD + ia * NB + ib] * w * w);
            mx = v > mx ? v : mx;
        }
    }
    a.qout[bi] = sqrt(mx);
}


template <int LA, int LB, int LC, int LD, typename R, int MODE>
void launch_eri(const EriArgs& a, hipStream_t s)
{
    if constexpr (MODE == MODE_SCHWARZ) {
        static_assert(LA == LC && LB == LD, "Schwarz bounds need a diagonal class");
        if (a.total <= 0) return;
        const int64_t blocks = (a.total + 255) / 256;
        hipLaunchKernelGGL((eri_schwarz_kernel<LA, LB, R>), dim3((unsigned)blocks), dim3(256), 0, s, a);
        QBX_HIP(hipGetLastError());
    } else if constexpr (EriClass<LA, LB, LC, LD, R>::STAGED) {
        if (a.nwave <= 0) return;
        if (!a.ws) throw QbxError(QBX_EINTERNAL, "staged ERI class launched without a workspace");
        // LDS K rows pay where K updates are dense: (NA+NB)(NC+ND) updates
        // per quartet against NA NB NC ND integrals, at least 1 to 5.  The
        // sparser (dd|dp) and (dd|dd) get them with chunks shared by the
        // block's waves (SPLIT > 1); with one wave per chunk they lose more
        // to the block's shared task (waves idle at its end, fewer tasks in
        // flight) than they save.
        constexpr int NA = qbx_ncart(LA), NB = qbx_ncart(LB), NC = qbx_ncart(LC), ND = qbx_ncart(LD);
        constexpr bool DENSE_K = 5 * (NA + NB) * (NC + ND) >= NA * NB * NC * ND;
        if constexpr (MODE == MODE_JK && DENSE_K) {
            const size_t shm = (size_t)a.nset * (NA + NB) * a.N * sizeof(double);
            const size_t boys_lds = QBX_BOYSL_ROWS * QBX_BOYSL_STRIDE * sizeof(double);
            static const bool rows_ok = [] {
                const bool ok = hipFuncSetAttribute((const void*)eri_staged_rows_kernel<LA, LB, LC, LD, R>,
                                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                                    ROW_LDS_MAX) == hipSuccess;
                (void)hipGetLastError();
                return ok;
            }();
            if ((EriClass<LA, LB, LC, LD, R>::SPLIT == 1 || !a.split) && rows_ok && a.spre &&
                shm <= (size_t)ROW_LDS_MAX && shm + boys_lds <= (size_t)a.lds_max) {
                if (a.nsuper <= 0) return;
                const int max_blocks = a.ws_slots / STAGE_WAVES_PER_BLOCK;
                const int blocks = a.nsuper < max_blocks ? a.nsuper : max_blocks;
                if (blocks <= 0) throw QbxError(QBX_EINTERNAL, "staged ERI workspace holds no block");
                hipLaunchKernelGGL((eri_staged_rows_kernel<LA, LB, LC, LD, R>), dim3((unsigned)blocks),
                                   dim3(64 * STAGE_WAVES_PER_BLOCK), shm, s, a);
                QBX_HIP(hipGetLastError());
                return;
            }
        }
        if constexpr (MODE == MODE_JK && EriClass<LA, LB, LC, LD, R>::SPLIT > 1) {
            constexpr int W = EriClass<LA, LB, LC, LD, R>::SPLIT;
            const size_t shm = (size_t)a.nset * (NA + NB) * a.N * sizeof(double);
            const size_t boys_lds = QBX_BOYSL_ROWS * QBX_BOYSL_STRIDE * sizeof(double);
            static const bool srows_ok = [] {
                const bool ok = hipFuncSetAttribute((const void*)eri_staged_split_rows_kernel<LA, LB, LC, LD, R>,
                                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                                    ROW_LDS_MAX) == hipSuccess;
                (void)hipGetLastError();
                return ok;
            }();
            if (a.split && srows_ok && a.spre && shm <= (size_t)ROW_LDS_MAX && shm + boys_lds <= (size_t)a.lds_max) {
                if (a.nsuper <= 0) return;
                // one slot per block
                const int blocks = a.nsuper < a.ws_slots ? a.nsuper : a.ws_slots;
                hipLaunchKernelGGL((eri_staged_split_rows_kernel<LA, LB, LC, LD, R>), dim3((unsigned)blocks),
                                   dim3(64 * W), shm, s, a);
                QBX_HIP(hipGetLastError());
                return;
            }
            if (a.split) {
                // one slot per block
                const int blocks = a.nwave < a.ws_slots ? a.nwave : a.ws_slots;
                hipLaunchKernelGGL((eri_staged_split_kernel<LA, LB, LC, LD, R>), dim3((unsigned)blocks),
                                   dim3(64 * EriClass<LA, LB, LC, LD, R>::SPLIT), 0, s, a);
                QBX_HIP(hipGetLastError());
                return;
            }
        }
        const int slots = a.nwave < a.ws_slots ? a.nwave : a.ws_slots;
        const int blocks = (slots + STAGE_WAVES_PER_BLOCK - 1) / STAGE_WAVES_PER_BLOCK;
        if (blocks * STAGE_WAVES_PER_BLOCK > a.ws_slots)
            throw QbxError(QBX_EINTERNAL, "staged ERI grid exceeds its workspace");
        hipLaunchKernelGGL((eri_staged_kernel<LA, LB, LC, LD, R, MODE>), dim3((unsigned)blocks),
                           dim3(64 * STAGE_WAVES_PER_BLOCK), 0, s, a);
        QBX_HIP(hipGetLastError());
    } else {
        if constexpr (MODE == MODE_JK) {
            if (a.ngroup > 0) {
                if constexpr (EriClass<LA, LB, LC, LD, R>::NACC <= TWIN_MAX_ACC) {
                    const size_t tshm =
                        (size_t)a.nset * a.twmax * (qbx_ncart(LA) + qbx_ncart(LB)) * a.N * sizeof(double);
                    static const bool twin_ok = [] {
                        const bool ok = hipFuncSetAttribute((const void*)eri_jk_twin_kernel<LA, LB, LC, LD, R>,
                                                            hipFuncAttributeMaxDynamicSharedMemorySize,
                                                            ROW_LDS_MAX) == hipSuccess;
                        (void)hipGetLastError();
                        return ok;
                    }();
                    if (!twin_ok || tshm > (size_t)ROW_LDS_MAX ||
                        tshm + QBX_BOYSL_ROWS * QBX_BOYSL_STRIDE * sizeof(double) > (size_t)a.lds_max)
                        throw QbxError(QBX_EINTERNAL, "twin-bra groups given where their LDS rows do not fit");
                    if (a.ntwtask > 0) {
                        hipLaunchKernelGGL((eri_jk_twin_kernel<LA, LB, LC, LD, R>), dim3((unsigned)a.ntwtask),
                                           dim3(64 * TWIN_WAVES), tshm, s, a);
                        QBX_HIP(hipGetLastError());
                    }
                } else {
                    throw QbxError(QBX_EINTERNAL, "twin-bra groups given to a class without a twin kernel");
                }
            }
            const size_t shm = (size_t)a.nset * (qbx_ncart(LA) + qbx_ncart(LB)) * a.N * sizeof(double);
            const size_t boys_lds = QBX_BOYSL_ROWS * QBX_BOYSL_STRIDE * sizeof(double);
            // the attribute call fails where the device cannot give a block
            // that much LDS; the one-wave-per-chunk kernel runs then
            static const bool rows_ok = [] {
                const bool ok = hipFuncSetAttribute((const void*)rows_kernel<LA, LB, LC, LD, R>(),
                                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                                    ROW_LDS_MAX) == hipSuccess;
                (void)hipGetLastError();
                return ok;
            }();
            if (rows_ok && a.spre && shm <= (size_t)ROW_LDS_MAX && shm + boys_lds <= (size_t)a.lds_max) {
                if (a.nsuper <= 0) return;
                static const size_t scratch = [] {
                    hipFuncAttributes fa;
                    if (hipFuncGetAttributes(&fa, (const void*)rows_kernel<LA, LB, LC, LD, R>()) != hipSuccess)
                        return (size_t)1;
                    return (size_t)fa.localSizeBytes;
                }();
                hipStream_t use = (scratch > 0 && a.scratch_stream) ? (hipStream_t)a.scratch_stream : s;
                hipLaunchKernelGGL((rows_kernel<LA, LB, LC, LD, R>()), dim3((unsigned)a.nsuper),
                                   dim3(64 * ROW_WAVES), shm, use, a);
                QBX_HIP(hipGetLastError());
                return;
            }
        }
        if (a.nwave <= 0) return;
        // Kernels that spill to scratch run on the caller's main stream only:
        // each extra hardware queue would need its own scratch allocation.
        static const size_t scratch = [] {
            hipFuncAttributes fa;
            if (hipFuncGetAttributes(&fa, (const void*)eri_kernel<LA, LB, LC, LD, R, MODE>) != hipSuccess)
                return (size_t)1;
            return (size_t)fa.localSizeBytes;
        }();
        hipStream_t use = (scratch > 0 && a.scratch_stream) ? (hipStream_t)a.scratch_stream : s;
        constexpr int TPB = eri_tpb(EriClass<LA, LB, LC, LD, R>::NACC);
        const int64_t blocks = ((int64_t)a.nwave + TPB / 64 - 1) / (TPB / 64);
        hipLaunchKernelGGL((eri_kernel<LA, LB, LC, LD, R, MODE>), dim3((unsigned)blocks), dim3(TPB), 0, use, a);
        QBX_HIP(hipGetLastError());
    }
}

}  // namespace qbx
