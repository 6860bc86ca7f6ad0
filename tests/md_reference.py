"""Independent McMurchie-Davidson evaluation of primitive integrals (numpy +
mpmath Boys), used only to pin the oracle's Obara-Saika restatement: a
dropped term, wrong sign or transposed index in either scheme makes the two
disagree.

Hermite expansion (Helgaker, Jorgensen & Olsen, Molecular Electronic-Structure
Theory, sec. 9.5):
  E^{i+1,j}_t = E^{ij}_{t-1}/(2p) + X_PA E^{ij}_t + (t+1) E^{ij}_{t+1}
  E^{i,j+1}_t = E^{ij}_{t-1}/(2p) + X_PB E^{ij}_t + (t+1) E^{ij}_{t+1}
  R^n_{t+1,u,v} = t R^{n+1}_{t-1,u,v} + X_PQ R^{n+1}_{tuv}   (same for u, v)
  R^n_{000} = (-2 alpha)^n F_n(alpha R_PQ^2)
"""
import math
from functools import lru_cache

import mpmath as mp
import numpy as np

mp.mp.dps = 40


def boys(n, T):
    if T == 0.0:
        return 1.0 / (2 * n + 1)
    t = mp.mpf(T)
    return float(mp.gammainc(n + mp.mpf(1) / 2, 0, t) / (2 * t ** (n + mp.mpf(1) / 2)))


def hermite_E(i, j, a, b, XAB):
    """E^{ij}_t, t = 0..i+j, for exponents a, b and X_AB = A - B."""
    p = a + b
    mu = a * b / p
    XPA = -b * XAB / p
    XPB = a * XAB / p
    E = np.zeros((i + 1, j + 1, i + j + 2))
    E[0, 0, 0] = math.exp(-mu * XAB * XAB)
    for ii in range(i + 1):
        for jj in range(j + 1):
            if ii == 0 and jj == 0:
                continue
            for t in range(ii + jj + 1):
                if ii > 0:
                    v = XPA * E[ii - 1, jj, t] + (t + 1) * E[ii - 1, jj, t + 1]
                    if t > 0:
                        v += E[ii - 1, jj, t - 1] / (2 * p)
                else:
                    v = XPB * E[ii, jj - 1, t] + (t + 1) * E[ii, jj - 1, t + 1]
                    if t > 0:
                        v += E[ii, jj - 1, t - 1] / (2 * p)
                E[ii, jj, t] = v
    return E[i, j, : i + j + 1]


def hermite_R(L, alpha, PQ):
    """R^0_{tuv} for t+u+v <= L."""
    T = alpha * float(np.dot(PQ, PQ))
    F = [boys(n, T) for n in range(L + 1)]
    R = {}

    @lru_cache(maxsize=None)
    def r(t, u, v, n):
        if t < 0 or u < 0 or v < 0:
            return 0.0
        if t == u == v == 0:
            return (-2 * alpha) ** n * F[n]
        if t > 0:
            return (t - 1) * r(t - 2, u, v, n + 1) + PQ[0] * r(t - 1, u, v, n + 1)
        if u > 0:
            return (u - 1) * r(t, u - 2, v, n + 1) + PQ[1] * r(t, u - 1, v, n + 1)
        return (v - 1) * r(t, u, v - 2, n + 1) + PQ[2] * r(t, u, v - 1, n + 1)

    return r


def eri(A, a, la, B, b, lb, C, c, lc, D, d, ld):
    A, B, C, D = (np.asarray(x, dtype=float) for x in (A, B, C, D))
    p, q = a + b, c + d
    P = (a * A + b * B) / p
    Q = (c * C + d * D) / q
    alpha = p * q / (p + q)
    Eab = [hermite_E(la[k], lb[k], a, b, A[k] - B[k]) for k in range(3)]
    Ecd = [hermite_E(lc[k], ld[k], c, d, C[k] - D[k]) for k in range(3)]
    L = sum(la) + sum(lb) + sum(lc) + sum(ld)
    r = hermite_R(L, alpha, P - Q)
    s = 0.0
    for t, et in enumerate(Eab[0]):
        for u, eu in enumerate(Eab[1]):
            for v, ev in enumerate(Eab[2]):
                for tt, ft in enumerate(Ecd[0]):
                    for uu, fu in enumerate(Ecd[1]):
                        for vv, fv in enumerate(Ecd[2]):
                            sign = -1.0 if (tt + uu + vv) % 2 else 1.0
                            s += et * eu * ev * ft * fu * fv * sign * r(t + tt, u + uu, v + vv, 0)
    return 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * s


def nuclear(A, a, la, B, b, lb, Cpt):
    """positive <a|1/|r-C||b>"""
    A, B, Cpt = (np.asarray(x, dtype=float) for x in (A, B, Cpt))
    p = a + b
    P = (a * A + b * B) / p
    E = [hermite_E(la[k], lb[k], a, b, A[k] - B[k]) for k in range(3)]
    r = hermite_R(sum(la) + sum(lb), p, P - Cpt)
    s = 0.0
    for t, et in enumerate(E[0]):
        for u, eu in enumerate(E[1]):
            for v, ev in enumerate(E[2]):
                s += et * eu * ev * r(t, u, v, 0)
    return 2 * math.pi / p * s


def overlap(A, a, la, B, b, lb):
    A, B = np.asarray(A, float), np.asarray(B, float)
    p = a + b
    r = 1.0
    for k in range(3):
        r *= hermite_E(la[k], lb[k], a, b, A[k] - B[k])[0] * math.sqrt(math.pi / p)
    return r
