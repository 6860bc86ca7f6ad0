"""Pinning the CPU oracle (the checker) against the reference's own golden
values, a 50-digit Boys table, closed forms and an independent
McMurchie-Davidson evaluation."""
import math

import numpy as np
import pytest

import md_reference as md
import oracle
import quiqbox_amd as qb
from helpers import OracleEngine, boys_table, contracted, lih_basis, literals, molecule_basis


def test_boys_reference_literals():
    L = literals()["boys_points"]
    for x, n, v in L["points"]:
        assert oracle.boys(x, n) == pytest.approx(v, rel=L["rtol"])
    s = oracle.boys_seq(50.0, 100)
    assert s[16] == pytest.approx(L["points"][9][2], rel=L["rtol"])
    assert s[64] == pytest.approx(L["points"][10][2], rel=L["rtol"])


def test_boys_against_mpmath_table():
    tab = boys_table()
    worst = 0.0
    for T, row in tab["rows"]:
        seq = oracle.boys_seq(T, 16)
        for m in range(17):
            ref = row[m]
            worst = max(worst, abs(seq[m] - ref) / ref)
            assert abs(oracle.boys(T, m) - ref) <= 2e-13 * ref
    assert worst < 2e-13


def _prim(k):
    c, x, a = literals()["prim_gtos"][str(k)]
    return (c, x, a)


def test_primitive_nuclear_literals():
    L = literals()["coulomb1"]
    for i, j, v in L["values"]:
        assert oracle.prim_coulomb1(_prim(i), _prim(j), L["point"]) == pytest.approx(v, rel=1e-12)


def test_primitive_eri_literals():
    for q, v in literals()["eri_prim"]["values"]:
        assert oracle.prim_eri(*[_prim(k) for k in q]) == pytest.approx(v, rel=1e-11)


def test_primitive_kinetic_literal():
    (p, v) = literals()["kinetic"]["prim"]
    assert oracle.prim_kinetic(tuple(p), tuple(p)) == pytest.approx(v, rel=1e-13)


def test_contracted_kinetic_literal():
    K = literals()["kinetic"]["contracted"]
    e = OracleEngine([contracted(K)])
    assert e.kinetic()[0, 0] == pytest.approx(K["value"], rel=1e-12)


def test_lih_overlap_and_nuclear_literals():
    # the reference asserts these with isapprox (rtol = sqrt(eps)); some of its
    # literals differ from the exact values by ~1e-10 (see the closed form below)
    L = literals()["lih"]
    e = OracleEngine(lih_basis())
    np.testing.assert_allclose(e.overlap().ravel(), L["S"], rtol=1.5e-8, atol=1e-12)
    Z = [qb.getCharge(s) for s in L["nucs"]]
    np.testing.assert_allclose(e.nuclear(Z, L["coords"]).ravel(), L["V"], rtol=1.5e-8, atol=1e-12)


def test_lih_self_overlap_closed_form():
    # <phi|phi> of a contraction of normalized s primitives:
    # sum_ij c_i c_j (2 sqrt(a_i a_j) / (a_i + a_j))^{3/2}, evaluated at 40 digits
    import mpmath as mp
    mp.mp.dps = 40
    e = OracleEngine(lih_basis())
    S = e.overlap()
    for k, b in enumerate(literals()["lih"]["basis"][:3]):
        x = [mp.mpf(repr(v)) for v in b["xpns"]]
        c = [mp.mpf(repr(v)) for v in b["cons"]]
        ref = sum(c[i] * c[j] * (2 * mp.sqrt(x[i] * x[j]) / (x[i] + x[j])) ** mp.mpf(1.5)
                  for i in range(3) for j in range(3))
        assert S[k, k] == pytest.approx(float(ref), rel=5e-15)


def test_interface_overlaps_and_core_hamiltonian():
    L = literals()["interface"]
    bs = [contracted(L["cgf1"]), contracted(L["cgf2"])]
    e = OracleEngine(bs)
    np.testing.assert_allclose(e.overlap(), L["S"], rtol=1e-14, atol=0)
    Z = [qb.getCharge(s) for s in L["nucs"]]
    H = e.core_hamiltonian(Z, L["coords"])
    assert np.array_equal(H, e.kinetic() + e.nuclear(Z, L["coords"]))


def test_h_sto3g_primitive_and_weighted_overlaps():
    L = literals()["h_sto3g_prim_overlap"]
    prims = [qb.genGaussTypeOrb((0.0, 0.0, 0.0), x) for x in L["xpns"]]
    np.testing.assert_allclose(OracleEngine(prims).overlap(), L["S"], rtol=1e-14)
    pn = [qb.genGaussTypeOrb((0.0, 0.0, 0.0), x, renormalize=True) for x in L["xpns"]]
    c = np.array(L["cons"])
    np.testing.assert_allclose(OracleEngine(pn).overlap() * np.outer(c, c), L["weighted_renormalized_S"], rtol=1e-13)


def test_li_321g_overlaps():
    L = literals()["li_321g_overlap"]
    bs = qb.genGaussTypeOrbSeq(tuple(L["center"]), "Li", "3-21G")
    assert len(bs) == 9
    np.testing.assert_allclose(OracleEngine(bs).overlap().ravel(), L["S"], rtol=0, atol=1e-14)


def test_overlap_recurrence_identity():
    # +2a <l+1|r> - l <l-1|r> = -2b <l|r+1> + r <l|r-1>  (Overlap-test.jl:8-18)
    a, b, A, B, l, r = 1.0, 1.5, (2.0, 0, 0), (3.1, 0, 0), 2, 4
    S = lambda li, ri: oracle.prim_overlap((A, a, (li, 0, 0)), (B, b, (ri, 0, 0)))
    lhs = 2 * a * S(l + 1, r) - l * S(l - 1, r)
    rhs = -2 * b * S(l, r + 1) + r * S(l, r - 1)
    assert lhs == pytest.approx(rhs, rel=1e-12)


def test_closed_forms_s_functions():
    rng = np.random.default_rng(3)
    for _ in range(20):
        A, B, C, D = (rng.uniform(-2, 2, 3) for _ in range(4))
        a, b, c, d = rng.uniform(0.1, 5, 4)
        z = (0, 0, 0)
        p, q = a + b, c + d
        P, Q = (a * A + b * B) / p, (c * C + d * D) / q
        KAB = math.exp(-a * b / p * np.sum((A - B) ** 2))
        KCD = math.exp(-c * d / q * np.sum((C - D) ** 2))
        rho = p * q / (p + q)
        ss = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * KAB * KCD * md.boys(0, rho * np.sum((P - Q) ** 2))
        assert oracle.prim_eri((A, a, z), (B, b, z), (C, c, z), (D, d, z)) == pytest.approx(ss, rel=1e-13)
        # (s|1/r_C|s) and overlap / kinetic of s functions
        v = 2 * math.pi / p * KAB * md.boys(0, p * np.sum((P - C) ** 2))
        assert oracle.prim_coulomb1((A, a, z), (B, b, z), C) == pytest.approx(v, rel=1e-13)
        s = (math.pi / p) ** 1.5 * KAB
        assert oracle.prim_overlap((A, a, z), (B, b, z)) == pytest.approx(s, rel=1e-13)
        mu = a * b / p
        t = mu * (3 - 2 * mu * np.sum((A - B) ** 2)) * s
        assert oracle.prim_kinetic((A, a, z), (B, b, z)) == pytest.approx(t, rel=1e-12, abs=1e-15)


def _rand_ang(rng, lmax):
    L = int(rng.integers(0, lmax + 1))
    i = int(rng.integers(0, L + 1))
    j = int(rng.integers(0, L - i + 1))
    return (i, j, L - i - j)


@pytest.mark.parametrize("seed", range(6))
def test_oracle_against_mcmurchie_davidson(seed):
    rng = np.random.default_rng(100 + seed)
    for _ in range(6):
        P = [(rng.uniform(-1.5, 1.5, 3), float(rng.uniform(0.2, 4.0)), _rand_ang(rng, 2)) for _ in range(4)]
        v = oracle.prim_eri(*[(tuple(c), x, a) for c, x, a in P])
        ref = md.eri(*[v_ for p in P for v_ in p])
        assert v == pytest.approx(ref, rel=1e-10, abs=1e-13)
        C = rng.uniform(-1.5, 1.5, 3)
        v1 = oracle.prim_coulomb1((tuple(P[0][0]), P[0][1], P[0][2]), (tuple(P[1][0]), P[1][1], P[1][2]), C)
        assert v1 == pytest.approx(md.nuclear(*P[0], *P[1], C), rel=1e-10, abs=1e-14)
        s1 = oracle.prim_overlap((tuple(P[0][0]), P[0][1], P[0][2]), (tuple(P[1][0]), P[1][1], P[1][2]))
        assert s1 == pytest.approx(md.overlap(*P[0], *P[1]), rel=1e-11, abs=1e-15)


def test_kinetic_against_overlap_derivative():
    # T = -1/2 sum_axis <a| d^2/dx^2 |b>, d^2/dx^2 of x_B^j e^{-b x_B^2} expanded into
    # overlaps of shifted Cartesian exponents (independent of the oracle's recursion order)
    rng = np.random.default_rng(7)
    for _ in range(15):
        A, B = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        a, b = rng.uniform(0.3, 3, 2)
        la, lb = _rand_ang(rng, 2), _rand_ang(rng, 2)
        t = 0.0
        for k in range(3):
            j = lb[k]
            def S(dj):
                l = list(lb)
                l[k] += dj
                return md.overlap(A, a, la, B, b, tuple(l)) if l[k] >= 0 else 0.0
            d2 = j * (j - 1) * S(-2) - 2 * b * (2 * j + 1) * S(0) + 4 * b * b * S(2)
            t += -0.5 * d2
        assert oracle.prim_kinetic((tuple(A), a, la), (tuple(B), b, lb)) == pytest.approx(t, rel=1e-10, abs=1e-14)


def test_dense_tensor_symmetry_and_gcore():
    e = OracleEngine(lih_basis())
    E = e.eri()
    n = E.shape[0]
    for perm in ["jikl", "ijlk", "jilk", "klij", "lkij", "klji", "lkji"]:
        assert np.array_equal(E, np.einsum("ijkl->" + perm, E)) or np.allclose(E, np.einsum("ijkl->" + perm, E), atol=0)
    rng = np.random.default_rng(1)
    A = rng.standard_normal((n, n))
    D = A + A.T
    G = oracle.gcore(E, 2 * D, D)
    G2 = np.einsum("ijkl,lk->ij", E, 2 * D) - np.einsum("iklj,kl->ij", E, D)
    np.testing.assert_allclose(G, G2, rtol=1e-13, atol=1e-13)


def test_multipole_reduces_to_overlap():
    # multipoleMoment(cen1, (2,0,0), cgf2, cgf1) == overlap(cgf2, cgf3) (Interface-test.jl:38-42)
    L = literals()["interface"]
    c1, c2 = L["cgf1"], L["cgf2"]
    cgf3 = dict(c1)
    cgf3["ang"] = [3, 0, 0]
    bs = [contracted(c1), contracted(c2), contracted(cgf3)]
    e = OracleEngine(bs)
    M = e.ob.multipole(c1["center"], (2, 0, 0))
    assert M[1, 0] == pytest.approx(e.overlap()[1, 2], rel=1e-13)
    M0 = e.ob.multipole(c1["center"], (0, 0, 0))
    np.testing.assert_array_equal(M0, e.overlap())


def test_reference_orientation_loses_digits_stable_one_does_not():
    """The reference evaluates each primitive quartet in its canonical index
    order; for mixed tight/diffuse d quartets the electron-transfer step then
    loses ~1e-9 absolute.  The oracle's default orientation (larger exponent
    first in each pair, smaller exponent sum on electron 1) agrees with a
    McMurchie-Davidson evaluation (40-digit Boys values) to ~1e-13."""
    import itertools
    from helpers import random_spd_basis
    coords, bs = random_spd_basis(0, n_atoms=3)
    f = qb.flatten_basis(bs)
    ob = oracle.Basis(f.as_dict())

    def terms(k):
        return [(tuple(f.prim_cen[f.bf_prim[s]]), f.prim_xpn[f.bf_prim[s]], tuple(f.prim_ang[f.bf_prim[s]]), f.bf_w[s])
                for s in range(f.bf_ptr[k], f.bf_ptr[k + 1])]
    q = (9, 9, 9, 26)
    ref = sum(np.prod([x[3] for x in t]) * md.eri(*[v for x in t for v in x[:3]])
              for t in itertools.product(*[terms(k) for k in q]))
    try:
        oracle.set_reference_orientation(True)
        v_ref = ob.eri(*q)
    finally:
        oracle.set_reference_orientation(False)
    v = ob.eri(*q)
    assert abs(v - ref) < 1e-13
    assert abs(v_ref - ref) > 1e-10


def test_multipole_off_centre_closed_forms():
    """Moments about a point M away from both centres, at one of them (the
    swap branch) and of a primitive with itself, against closed forms of the
    Gaussian product exp(-a r_A^2) exp(-b r_B^2) = K exp(-p r_P^2):
    <s|(x-M)^n|s> = S (u^n + C(n,2) u^{n-2}/(2p) + 3 [n=4]/(4 p^2)), u = P_x - M_x,
    <x_A|(x-M)|s> = S ((P_x - A_x) u + 1/(2p))."""
    import math
    a, b = 0.9, 0.35
    A = (0.3, -0.2, 0.5)
    B = (-0.4, 0.6, 1.1)
    p = a + b
    P = [(a * A[i] + b * B[i]) / p for i in range(3)]
    K = math.exp(-a * b / p * sum((A[i] - B[i]) ** 2 for i in range(3)))
    S = K * (math.pi / p) ** 1.5

    def s_moment(u, n):
        return {1: u, 2: u * u + 1 / (2 * p), 3: u ** 3 + 3 * u / (2 * p),
                4: u ** 4 + 6 * u * u / (2 * p) + 3 / (4 * p * p)}[n]

    s0 = (0, 0, 0)
    for M in ((0.7, 0.1, -0.3), A, B):
        for t in range(3):
            for n in (1, 2, 3, 4):
                deg = [0, 0, 0]
                deg[t] = n
                u = P[t] - M[t]
                ref = S * s_moment(u, n)
                v = oracle.prim_multipole((A, a, s0), (B, b, s0), M, deg)
                assert v == pytest.approx(ref, rel=1e-13, abs=1e-15), (M, t, n)
                v2 = oracle.prim_multipole((B, b, s0), (A, a, s0), M, deg)
                assert v2 == pytest.approx(ref, rel=1e-13, abs=1e-15), (M, t, n)
        # p function on A: <x_A|(x - M)|s_B>
        u = P[0] - M[0]
        ref = S * ((P[0] - A[0]) * u + 1 / (2 * p))
        v = oracle.prim_multipole((A, a, (1, 0, 0)), (B, b, s0), M, (1, 0, 0))
        assert v == pytest.approx(ref, rel=1e-13, abs=1e-15)
        v2 = oracle.prim_multipole((B, b, s0), (A, a, (1, 0, 0)), M, (1, 0, 0))
        assert v2 == pytest.approx(ref, rel=1e-13, abs=1e-15)
    # a primitive with itself about an outside point: P = A, p = 2a
    M = (0.7, 0.1, -0.3)
    Sa = (math.pi / (2 * a)) ** 1.5
    v = oracle.prim_multipole((A, a, s0), (A, a, s0), M, (2, 0, 0))
    assert v == pytest.approx(Sa * ((A[0] - M[0]) ** 2 + 1 / (4 * a)), rel=1e-13)
