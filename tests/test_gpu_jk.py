"""Direct J/K (qbx_jk) against the oracle's dense getGcore, at small sizes,
and size-independent properties at the measured workload sizes."""
import numpy as np
import pytest

import oracle
import quiqbox_amd as qb
from helpers import OracleEngine, molecule_basis, literals, random_spd_basis
from quiqbox_amd.systems import h10_sto3g, random_cluster

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[gpu-test] " + msg, flush=True)


def _sym(rng, n, scale=0.1):
    A = rng.standard_normal((n, n))
    return scale * (A + A.T)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_jk_against_dense_oracle(seed):
    coords, bs = random_spd_basis(10 + seed, n_atoms=3 if seed < 2 else 4)
    gb = qb.GPUBasis(bs)
    eng = OracleEngine(gb.flat)
    E = eng.eri()
    rng = np.random.default_rng(seed)
    n = gb.n
    _say("jk seed %d n_bf=%d" % (seed, n))
    DJ = np.stack([_sym(rng, n), _sym(rng, n)])
    DK = np.stack([_sym(rng, n), _sym(rng, n)])
    G = gb.jk(DJ, DK, 0.0)
    for s in range(2):
        Gref = oracle.gcore(E, DJ[s], DK[s])
        err = np.max(np.abs(G[s] - Gref))
        _say("  set %d max |G - G_oracle| = %.3e" % (s, err))
        assert err < 1e-11
    # n_set = 1 path, screening at 1e-12 changes nothing measurable here
    G1 = gb.jk(DJ[0], DK[0], 1e-12)
    assert np.max(np.abs(G1 - G[0])) < 1e-10


def test_jk_h10_and_hho():
    for label, (nuc, bs) in [("H10", h10_sto3g()),
                             ("HHO", (None, molecule_basis(["H", "H", "O"], [(-0.7, 0, 0), (0.7, 0, 0), (0, 0, 0)],
                                                           "STO-3G")))]:
        gb = qb.GPUBasis(bs)
        E = OracleEngine(gb.flat).eri()
        rng = np.random.default_rng(4)
        D = _sym(rng, gb.n)
        _say("jk %s" % label)
        G = gb.jk(2 * D, D, 1e-12)
        assert np.max(np.abs(G - oracle.gcore(E, 2 * D, D))) < 1e-11


def test_jk_mixed_precision_error_bound():
    coords, bs = random_spd_basis(21, n_atoms=4)
    gb = qb.GPUBasis(bs)
    E = OracleEngine(gb.flat).eri()
    rng = np.random.default_rng(5)
    D = _sym(rng, gb.n)
    Gref = oracle.gcore(E, 2 * D, D)
    _say("jk mixed precision n_bf=%d" % gb.n)
    G = gb.jk(2 * D, D, 1e-12, qb.QBX_PREC_MIXED)
    err = np.max(np.abs(G - Gref))
    _say("  mixed max |G - G_oracle| = %.3e (|G|max %.3e)" % (err, np.max(np.abs(Gref))))
    assert err < 1e-5 * max(1.0, np.max(np.abs(Gref)))
    gb.ctx.set_mixed_cut(0.0)  # every quartet above the cut: FP64 only
    G64 = gb.jk(2 * D, D, 1e-12, qb.QBX_PREC_MIXED)
    gb.ctx.set_mixed_cut(1e-5)
    assert np.max(np.abs(G64 - Gref)) < 1e-11


def test_jk_config2_properties():
    """Full config-2 size (64 atoms, 256 functions): linearity, symmetry,
    determinism to rounding, screening monotonicity."""
    nuc, bs = random_cluster(64, 0)
    gb = qb.GPUBasis(bs)
    n = gb.n
    rng = np.random.default_rng(6)
    D1, D2 = _sym(rng, n, 0.01), _sym(rng, n, 0.01)
    _say("jk config-2 properties n_bf=%d" % n)
    G1 = gb.jk(2 * D1, D1, 1e-12)
    G2 = gb.jk(2 * D2, D2, 1e-12)
    G12 = gb.jk(2 * (D1 + 0.5 * D2), D1 + 0.5 * D2, 1e-12)
    scale = np.max(np.abs(G12))
    assert np.max(np.abs(G12 - (G1 + 0.5 * G2))) < 1e-12 * max(scale, 1.0) * 100
    assert np.max(np.abs(G1 - G1.T)) == 0.0
    G1b = gb.jk(2 * D1, D1, 1e-12)
    assert np.max(np.abs(G1b - G1)) < 1e-12 * max(1.0, np.max(np.abs(G1)))
    G0 = gb.jk(2 * D1, D1, 0.0)
    assert np.max(np.abs(G0 - G1)) < 1e-9
    st = gb.jk_stats(1e-12)
    st0 = gb.jk_stats(0.0)
    assert st["shell_quartets"].sum() <= st0["shell_quartets"].sum()
    npair = 128 * 129 // 2
    assert st0["shell_quartets"].sum() == npair * (npair + 1) // 2


def test_jk_config2_sampled_against_oracle():
    """Config-2 size: J and K columns of G checked against the oracle's
    contracted quartets for a random subset of rows."""
    nuc, bs = random_cluster(64, 1)
    gb = qb.GPUBasis(bs)
    n = gb.n
    rng = np.random.default_rng(7)
    D = _sym(rng, n, 0.01)
    _say("jk config-2 sampled n_bf=%d" % n)
    G = gb.jk(2 * D, D, 0.0)
    ob = oracle.Basis(gb.flat.as_dict())
    for mu, nu in [(0, 0), (5, 17), (100, 200), (255, 3)]:
        q = []
        for l in range(n):
            for s in range(n):
                q.append((mu, nu, l, s))
        J = ob.eri_list(np.array(q), threads=8).reshape(n, n)
        q = []
        for l in range(n):
            for s in range(n):
                q.append((mu, l, s, nu))
        K = ob.eri_list(np.array(q), threads=8).reshape(n, n)
        g = np.sum(2 * D.T * J) - np.sum(D * K)
        _say("  G[%d,%d] gpu %.12e oracle %.12e" % (mu, nu, G[mu, nu], g))
        assert abs(G[mu, nu] - g) < 1e-11
