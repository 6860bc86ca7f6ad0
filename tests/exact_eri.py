"""Exact contracted ERIs for pinning the oracle and the GPU kernels at high
angular momentum (independent of both: McMurchie-Davidson in mpmath).

For primitives x_A^i exp(-a r_A^2) and x_B^j exp(-b r_B^2) the product is
expanded in Hermite Gaussians about P = (aA + bB)/p (Helgaker, Jorgensen &
Olsen sec. 9.5):

  E^{00}_0 = exp(-ab/p X_AB^2),
  E^{i+1,j}_t = E^{ij}_{t-1} / (2p) + X_PA E^{ij}_t + (t+1) E^{ij}_{t+1},
  E^{i,j+1}_t = E^{ij}_{t-1} / (2p) + X_PB E^{ij}_t + (t+1) E^{ij}_{t+1},

and (ab|cd) = 2 pi^{5/2} / (p q sqrt(p+q)) sum E^{ab}_{tuv} E^{cd}_{tau nu phi}
(-1)^{tau+nu+phi} R_{t+tau,u+nu,v+phi} with the Hermite Coulomb integrals R of
boys_pin.py, all at 40 digits.  A contracted value sums the primitive values
with the basis functions' final weights (the flattened basis)."""
from functools import lru_cache

import mpmath as mp

from boys_pin import boys_mp

mp.mp.dps = 40


def _hermite_pair(i, j, a, b, A, B):
    p = a + b
    mu = a * b / p
    P = (a * A + b * B) / p
    XPA, XPB = P - A, P - B

    @lru_cache(maxsize=None)
    def E(i, j, t):
        if t < 0 or t > i + j:
            return mp.mpf(0)
        if i == 0 and j == 0:
            return mp.e ** (-mu * (A - B) ** 2)
        if i > 0:
            return E(i - 1, j, t - 1) / (2 * p) + XPA * E(i - 1, j, t) + (t + 1) * E(i - 1, j, t + 1)
        return E(i, j - 1, t - 1) / (2 * p) + XPB * E(i, j - 1, t) + (t + 1) * E(i, j - 1, t + 1)

    return [E(i, j, t) for t in range(i + j + 1)]


def primitive_eri(pa, pb, pc, pd):
    """(ab|cd) of unweighted primitives given as (centre, exponent, (i, j, k))."""
    (A, a, la), (B, b, lb), (C, c, lc), (D, d, ld) = pa, pb, pc, pd
    a, b, c, d = (mp.mpf(x) for x in (a, b, c, d))
    A, B, C, D = ([mp.mpf(x) for x in v] for v in (A, B, C, D))
    p, q = a + b, c + d
    rho = p * q / (p + q)
    PQ = [(a * A[k] + b * B[k]) / p - (c * C[k] + d * D[k]) / q for k in range(3)]
    T = rho * (PQ[0] ** 2 + PQ[1] ** 2 + PQ[2] ** 2)
    F = [boys_mp(n, T) for n in range(sum(la) + sum(lb) + sum(lc) + sum(ld) + 1)]

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

    Eab = [_hermite_pair(la[k], lb[k], a, b, A[k], B[k]) for k in range(3)]
    Ecd = [_hermite_pair(lc[k], ld[k], c, d, C[k], D[k]) for k in range(3)]
    s = mp.mpf(0)
    for t, et in enumerate(Eab[0]):
        for u, eu in enumerate(Eab[1]):
            for v, ev in enumerate(Eab[2]):
                e1 = et * eu * ev
                for tau, ft in enumerate(Ecd[0]):
                    for nu, fu in enumerate(Ecd[1]):
                        for ph, fv in enumerate(Ecd[2]):
                            sgn = -1 if (tau + nu + ph) % 2 else 1
                            s += sgn * e1 * ft * fu * fv * R(0, t + tau, u + nu, v + ph)
    return 2 * mp.pi ** mp.mpf(2.5) / (p * q * mp.sqrt(p + q)) * s


def contracted_eri(flat, i, j, k, l):
    """(ij|kl) of basis functions of a flattened basis (fields prim_cen,
    prim_xpn, prim_ang, bf_ptr, bf_prim, bf_w), as a float."""
    def terms(f):
        out = []
        for z in range(int(flat.bf_ptr[f]), int(flat.bf_ptr[f + 1])):
            pi = int(flat.bf_prim[z])
            out.append(((tuple(float(x) for x in flat.prim_cen[pi]), float(flat.prim_xpn[pi]),
                         tuple(int(x) for x in flat.prim_ang[pi])), mp.mpf(float(flat.bf_w[z]))))
        return out

    s = mp.mpf(0)
    for pa, wa in terms(i):
        for pb, wb in terms(j):
            for pc, wc in terms(k):
                for pd, wd in terms(l):
                    s += wa * wb * wc * wd * primitive_eri(pa, pb, pc, pd)
    return float(s)
