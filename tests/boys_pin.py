"""Exact two-centre ERIs for pinning the Boys function inside the ERI kernels.

Bra functions on centre A, ket functions on centre C, all single unweighted
primitives: phi = x_A^i y_A^j z_A^k exp(-alpha r_A^2) on A (likewise gamma on
C).  The bra product is the one-centre Gaussian x^I y^J z^K exp(-p r_A^2),
p = 2 alpha, I = i_a + i_b, so with the Hermite expansion about its own
centre (X_PA = 0; Helgaker, Jorgensen & Olsen sec. 9.5)

  E^{n+1}_t = E^n_{t-1} / (2p) + (t + 1) E^n_{t+1},   E^0_0 = 1,
  (ab|cd) = 2 pi^{5/2} / (p q sqrt(p+q)) sum_{tuv} E^I_t E^J_u E^K_v
            sum_{tau nu phi} (-1)^{tau+nu+phi} E'^{I'}_tau E'^{J'}_nu E'^{K'}_phi
            R_{t+tau, u+nu, v+phi},
  R^n_{t+1,u,v} = t R^{n+1}_{t-1,u,v} + X_PQ R^{n+1}_{tuv}  (likewise u, v),
  R^n_{000} = (-2 rho)^n F_n(rho |P-Q|^2),  rho = p q / (p + q),

every step in mpmath at 40 digits with F_n from mpmath's incomplete gamma
(F_n(T) = gamma(n + 1/2, T) / (2 T^{n+1/2})).  The value depends on the Boys
function at the single argument T = rho |A - C|^2 through F_0 .. F_{L_total},
so placing C at sqrt(T / rho) along x pins the device's F values of every
order at T.
"""
from functools import lru_cache

import mpmath as mp

mp.mp.dps = 40


def boys_mp(n, T):
    T = mp.mpf(T)
    if T == 0:
        return mp.mpf(1) / (2 * n + 1)
    return mp.gammainc(n + mp.mpf(1) / 2, 0, T) / (2 * T ** (n + mp.mpf(1) / 2))


def hermite_one_centre(n, p):
    """E^n_t, t = 0..n, of x^n exp(-p x^2) about its own centre."""
    E = [mp.mpf(1)]
    for k in range(n):
        nxt = [mp.mpf(0)] * (k + 2)
        for t in range(k + 2):
            v = mp.mpf(0)
            if t >= 1:
                v += E[t - 1] / (2 * p)
            if t + 1 <= k:
                v += (t + 1) * E[t + 1]
            nxt[t] = v
        E = nxt
    return E


def eri_two_centre(ang_bra, ang_ket, alpha, gamma, A, C, magnitude=False):
    """(ab|cd) for the bra product x^I y^J z^K (ang_bra = (I, J, K)) at A with
    exponent 2 alpha and the ket product (ang_ket) at C with exponent 2 gamma.
    magnitude=True also returns the sum of the absolute values of all the
    F_m terms the value is made of (the scale of the cancellation it comes
    from)."""
    p = 2 * mp.mpf(alpha)
    q = 2 * mp.mpf(gamma)
    rho = p * q / (p + q)
    PQ = [mp.mpf(A[i]) - mp.mpf(C[i]) for i in range(3)]
    T = rho * (PQ[0] ** 2 + PQ[1] ** 2 + PQ[2] ** 2)
    Ltot = sum(ang_bra) + sum(ang_ket)
    F = [boys_mp(n, T) for n in range(Ltot + 1)]

    @lru_cache(maxsize=None)
    def R(n, t, u, v):
        if t < 0 or u < 0 or v < 0:
            return mp.mpf(0)
        if t == 0 and u == 0 and v == 0:
            return (-2 * rho) ** n * F[n]
        if t > 0:
            return (t - 1) * R(n + 1, t - 2, u, v) + PQ[0] * R(n + 1, t - 1, u, v)
        if u > 0:
            return (u - 1) * R(n + 1, t, u - 2, v) + PQ[1] * R(n + 1, t, u - 1, v)
        return (v - 1) * R(n + 1, t, u, v - 2) + PQ[2] * R(n + 1, t, u, v - 1)

    @lru_cache(maxsize=None)
    def Rabs(n, t, u, v):
        """the same recursion over absolute values: the scale of the F_m terms
        that R sums (its internal cancellation)"""
        if t < 0 or u < 0 or v < 0:
            return mp.mpf(0)
        if t == 0 and u == 0 and v == 0:
            return abs((-2 * rho) ** n * F[n])
        if t > 0:
            return (t - 1) * Rabs(n + 1, t - 2, u, v) + abs(PQ[0]) * Rabs(n + 1, t - 1, u, v)
        if u > 0:
            return (u - 1) * Rabs(n + 1, t, u - 2, v) + abs(PQ[1]) * Rabs(n + 1, t, u - 1, v)
        return (v - 1) * Rabs(n + 1, t, u, v - 2) + abs(PQ[2]) * Rabs(n + 1, t, u, v - 1)

    Eb = [hermite_one_centre(k, p) for k in ang_bra]
    Ek = [hermite_one_centre(k, q) for k in ang_ket]
    s = mp.mpf(0)
    mag = mp.mpf(0)
    for t, et in enumerate(Eb[0]):
        if et == 0:
            continue
        for u, eu in enumerate(Eb[1]):
            if eu == 0:
                continue
            for v, ev in enumerate(Eb[2]):
                if ev == 0:
                    continue
                for tau, ft in enumerate(Ek[0]):
                    if ft == 0:
                        continue
                    for nu, fu in enumerate(Ek[1]):
                        if fu == 0:
                            continue
                        for ph, fv in enumerate(Ek[2]):
                            if fv == 0:
                                continue
                            sgn = -1 if (tau + nu + ph) % 2 else 1
                            term = sgn * et * eu * ev * ft * fu * fv * R(0, t + tau, u + nu, v + ph)
                            s += term
                            mag += abs(et * eu * ev * ft * fu * fv) * Rabs(0, t + tau, u + nu, v + ph)
    pref = 2 * mp.pi ** mp.mpf(2.5) / (p * q * mp.sqrt(p + q))
    if magnitude:
        return float(pref * s), float(pref * mag)
    return float(pref * s)


# asymptotic cutoffs T_L of the device's per-order Boys tables (order L = total
# angular momentum of the ERI class): tests straddle each of them
BOYS_TL = (36.0, 40.0, 44.0, 47.0, 50.0, 52.0, 55.0, 57.0, 60.0)
