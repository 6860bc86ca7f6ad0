"""Primitive-pair one-electron integrals on the GPU (qbx_prim_one_e): the
reference's per-primitive engines at angular momentum up to 10 per primitive.

- The reference's own literals (tests/golden/reference_literals.json): the
  one-body repulsion values of Coulomb-test.jl:27-40, including the l = 10
  primitive (1, 7, 2), and the primitive kinetic value of Kinetic-test.jl.
- Random primitive pairs with every total angular momentum 0..10 against the
  oracle's primitive routines (overlap, kinetic, one-body repulsion,
  multipoles), relative to the magnitude of the value."""
import numpy as np
import pytest

import oracle
import quiqbox_amd as qb
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
