"""Primitive-level integrals on the GPU (qbx_prim_one_e, qbx_prim_eri): the
reference's per-primitive engines at angular momentum up to 10 per primitive.

- The reference's own literals (tests/golden/reference_literals.json): the
  one-body repulsion values of Coulomb-test.jl:27-40, including the l = 10
  primitive (1, 7, 2), and the primitive kinetic value of Kinetic-test.jl.
- The reference's primitive ERI literals (Coulomb-test.jl:51-81, total
  angular momentum up to 24, the l = 10 primitive in four positions).
- Random primitive pairs with every total angular momentum 0..10 against the
  oracle's primitive routines (overlap, kinetic, one-body repulsion,
  multipoles), and random primitive quartets up to l = 10 each and 24 in
  total against the oracle's primitive ERI, relative to the magnitude of the
  value."""
import numpy as np
import pytest

import oracle
import quiqbox_amd as qb
from exact_eri import primitive_eri
from helpers import literals
from quiqbox_amd import _lib
from quiqbox_amd.engine import default_context

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[gpu-test] " + msg, flush=True)


def _prims(keys):
    g = literals()["prim_gtos"]
    return [(tuple(g[str(k)][0]), g[str(k)][1], tuple(g[str(k)][2])) for k in keys]


def _run(kind, pairs, C=None, degree=None):
    A = [p[0] for p, _ in pairs]
    a = [p[1] for p, _ in pairs]
    la = [p[2] for p, _ in pairs]
    B = [q[0] for _, q in pairs]
    b = [q[1] for _, q in pairs]
    lb = [q[2] for _, q in pairs]
    return default_context().prim_one_e(kind, A, a, la, B, b, lb, C, degree)


def test_reference_coulomb_literals_up_to_l10():
    lit = literals()["coulomb1"]
    pairs = [tuple(_prims([i, j])) for i, j, _ in lit["values"]]
    ref = np.array([v for _, _, v in lit["values"]])
    got = _run(_lib.QBX_PRIM_COULOMB, pairs, C=lit["point"])
    for (i, j, v), g in zip(lit["values"], got):
        _say("(%d|1/r|%d): gpu %.16e reference %.16e rel %.1e" % (i, j, g, v, abs(g - v) / abs(v)))
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=0)


def test_reference_kinetic_literal():
    lit = literals()["kinetic"]
    c, x, ang = lit["prim"][0]
    p = (tuple(c), x, tuple(ang))
    got = _run(_lib.QBX_PRIM_KINETIC, [(p, p)])[0]
    _say("primitive kinetic: gpu %.16e reference %.16e" % (got, lit["prim"][1]))
    assert got == pytest.approx(lit["prim"][1], rel=1e-13)


def _random_pairs(rng, n):
    pairs = []
    for k in range(n):
        L1, L2 = k % 11, (7 * k + 3) % 11
        def ang(L):
            i = int(rng.integers(0, L + 1))
            j = int(rng.integers(0, L - i + 1))
            return (i, j, L - i - j)
        pairs.append(((tuple(rng.uniform(-1.0, 1.0, 3)), float(np.exp(rng.uniform(np.log(0.3), np.log(4.0)))), ang(L1)),
                      (tuple(rng.uniform(-1.0, 1.0, 3)), float(np.exp(rng.uniform(np.log(0.3), np.log(4.0)))), ang(L2))))
    return pairs


@pytest.mark.parametrize("kind", ["overlap", "kinetic", "coulomb", "multipole"])
def test_random_high_l_pairs_against_oracle(kind):
    rng = np.random.default_rng({"overlap": 1, "kinetic": 2, "coulomb": 3, "multipole": 4}[kind])
    pairs = _random_pairs(rng, 121)
    C = (0.25, -0.4, 0.6)
    deg = (2, 1, 3)
    code = {"overlap": _lib.QBX_PRIM_OVERLAP, "kinetic": _lib.QBX_PRIM_KINETIC,
            "coulomb": _lib.QBX_PRIM_COULOMB, "multipole": _lib.QBX_PRIM_MULTIPOLE}[kind]
    got = _run(code, pairs, C=C if kind in ("coulomb", "multipole") else None,
               degree=deg if kind == "multipole" else None)
    ref = []
    for p, q in pairs:
        if kind == "overlap":
            ref.append(oracle.prim_overlap(p, q))
        elif kind == "kinetic":
            ref.append(oracle.prim_kinetic(p, q))
        elif kind == "coulomb":
            ref.append(oracle.prim_coulomb1(p, q, C))
        else:
            ref.append(oracle.prim_multipole(p, q, C, deg))
    ref = np.array(ref)
    scale = np.max(np.abs(ref))
    rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-6 * scale)
    k = int(np.argmax(rel))
    _say("%s: %d pairs (l up to 10), worst rel %.2e at l %d/%d (gpu %.6e oracle %.6e)" % (
        kind, len(pairs), rel[k], sum(pairs[k][0][2]), sum(pairs[k][1][2]), got[k], ref[k]))
    assert rel[k] < 1e-10


def test_rejects_angular_momentum_above_limit():
    p = ((0.0, 0.0, 0.0), 1.0, (11, 0, 0))
    with pytest.raises(Exception):
        _run(_lib.QBX_PRIM_OVERLAP, [(p, p)])


def test_reference_primitive_eri_literals():
    lit = literals()["eri_prim"]
    quartets = [_prims(q) for q, _ in lit["values"]]
    ref = np.array([v for _, v in lit["values"]])
    got = default_context().prim_eri(quartets)
    # exact values (40-digit McMurchie-Davidson): the reference's own literals
    # are up to 4.7e-13 off them at l = 10
    exact = np.array([float(primitive_eri(*q)) for q in quartets])
    for (q, v), g, x in zip(lit["values"], got, exact):
        _say("(%d %d|%d %d): gpu %.16e reference %.16e exact %.16e" % (tuple(q) + (g, v, x)))
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=0)
    np.testing.assert_allclose(got, exact, rtol=2e-13, atol=0)


def test_random_primitive_quartets_against_exact_values():
    rng = np.random.default_rng(9)
    quartets = []
    for k in range(96):
        Ls = [int(x) for x in rng.integers(0, 11, 4)]
        while sum(Ls) > 24:
            Ls[int(np.argmax(Ls))] -= 1
        q = []
        for L in Ls:
            i = int(rng.integers(0, L + 1))
            j = int(rng.integers(0, L - i + 1))
            q.append((tuple(rng.uniform(-1.0, 1.0, 3)), float(np.exp(rng.uniform(np.log(0.3), np.log(4.0)))),
                      (i, j, L - i - j)))
        quartets.append(q)
    got = default_context().prim_eri(quartets)
    # against exact values: at these angular momenta the oracle's own
    # recurrence loses up to 1.3e-8 relative (quartet 4, l = 6, 7, 5, 6)
    exact = np.array([float(primitive_eri(*q)) for q in quartets])
    ref = np.array([oracle.prim_eri(*q) for q in quartets])
    scale = np.max(np.abs(exact))
    rel = np.abs(got - exact) / np.maximum(np.abs(exact), 1e-6 * scale)
    rel_o = np.abs(ref - exact) / np.maximum(np.abs(exact), 1e-6 * scale)
    k = int(np.argmax(rel))
    _say("primitive ERIs: %d quartets (l up to 10 each, 24 total), worst rel to exact %.2e at l %s (gpu %.6e exact "
         "%.6e); the oracle's worst %.2e" % (len(quartets), rel[k], [sum(p[2]) for p in quartets[k]], got[k],
                                             exact[k], float(np.max(rel_o))))
    assert rel[k] < 1e-10
