"""f and g shells (cc-pVTZ, cc-pVQZ; the reference's N2 cc-pVTZ benchmark
system, benchmark/benchmarks.jl:70-72) through the GPU path: one-electron
integrals, the dense ERI tensor, the direct J/K build and RHF energies
against the oracle.  Quartets with an f or g shell run the generic kernel;
the rest the generated class kernels, so every test mixes both."""
import os

import numpy as np
import pytest

import oracle
import quiqbox_amd as qb
from exact_eri import contracted_eri
from helpers import OracleEngine, molecule_basis, oracle_eri_dense

pytestmark = pytest.mark.gpu
THREADS = max(1, min(16, os.cpu_count() or 1))


def _say(msg):
    print("[gpu-test] " + msg, flush=True)


def _water(basis):
    coords = [(0.0, 0.0, 0.0), (1.8, 0.0, 0.0), (-0.45, 1.75, 0.0)]
    return qb.NuclearCluster(["O", "H", "H"], coords, pairwiseSort=False), \
        molecule_basis(["O", "H", "H"], coords, basis)


def _n2():
    coords = [(0.0, 0.0, 0.0), (2.074, 0.0, 0.0)]
    return qb.NuclearCluster(["N", "N"], coords, pairwiseSort=False), molecule_basis(["N", "N"], coords, "cc-pVTZ")


def _small_fg():
    """Two centres with s, d, f (2 primitives) and p, g shells."""
    rng = np.random.default_rng(3)
    bs = []
    for cen, shells in (((0.1, -0.2, 0.3), ((0, 2), (2, 1), (3, 2))), ((1.2, 0.7, -0.4), ((1, 2), (4, 1)))):
        for L, k in shells:
            xp = list(np.exp(rng.uniform(np.log(0.3), np.log(4.0), k)))
            co = list(rng.uniform(0.2, 1.0, k))
            for ijk in qb.subshellXYZs(L):
                cons = [c * qb.get3DimPGTOrbNormFactor(x, ijk) for x, c in zip(xp, co)]
                bs.append(qb.genGaussTypeOrb(cen, xp, cons, ijk, outerRenormalize=True))
    return bs


@pytest.mark.parametrize("system", ["water_ccpvtz", "n2_ccpvtz", "water_ccpvqz"])
def test_one_electron_integrals_against_oracle(system):
    nuc, bs = _n2() if system == "n2_ccpvtz" else _water("cc-pVTZ" if system == "water_ccpvtz" else "cc-pVQZ")
    gb = qb.GPUBasis(bs)
    eng = OracleEngine(gb.flat)
    info = gb.info()
    _say("%s: %d functions, highest L %d" % (system, gb.n, info["max_l"]))
    assert info["max_l"] == (4 if system == "water_ccpvqz" else 3)
    Z, R = nuc.charges, nuc.coord_array
    np.testing.assert_allclose(gb.overlap(), eng.overlap(), rtol=0, atol=1e-12)
    np.testing.assert_allclose(gb.kinetic(), eng.kinetic(), rtol=0, atol=1e-12)
    np.testing.assert_allclose(gb.nuclear(Z, R), eng.nuclear(Z, R), rtol=0, atol=1e-11)
    for deg in [(1, 0, 0), (0, 2, 1), (2, 2, 0)]:
        np.testing.assert_allclose(gb.multipole((0.3, -0.2, 0.5), deg), eng.ob.multipole((0.3, -0.2, 0.5), deg),
                                   rtol=0, atol=1e-11)


@pytest.mark.parametrize("system", ["small_fg", "water_ccpvtz", "n2_ccpvtz"])
def test_dense_eri_and_jk_against_oracle(system):
    if system == "small_fg":
        bs = _small_fg()
    else:
        bs = (_n2() if system == "n2_ccpvtz" else _water("cc-pVTZ"))[1]
    gb = qb.GPUBasis(bs)
    n = gb.n
    ob = oracle.Basis(gb.flat.as_dict())
    _say("%s: %d functions, oracle tensor" % (system, n))
    E = oracle_eri_dense(ob, n, THREADS)
    Eg = gb.eri_dense()
    dif = np.abs(Eg - E)
    err = float(np.max(dif))
    _say("  dense ERI max|d| %.3e (max %.3e)" % (err, float(np.max(np.abs(E)))))
    Ls = [int(np.sum(gb.flat.prim_ang[gb.flat.bf_prim[gb.flat.bf_ptr[i]]])) for i in range(n)]
    for flat in np.argsort(dif.ravel())[::-1][:4]:
        idx = tuple(int(x) for x in np.unravel_index(flat, dif.shape))
        ex = contracted_eri(gb.flat, *idx)
        _say("    (%d %d|%d %d) L %s: gpu %.16e oracle %.16e exact %.16e" %
             (idx + (tuple(Ls[i] for i in idx), Eg[idx], E[idx], ex)))
        # where the two engines differ most, the GPU is as close to the exact
        # value as the oracle's conditioning allows (N2: x^3 functions on both
        # atoms along the bond, both engines about 2e-12 off 0.36)
        assert abs(Eg[idx] - ex) <= 4e-12 * max(1.0, abs(ex))
    assert err < 5e-12
    rng = np.random.default_rng(8)
    A = [rng.standard_normal((n, n)) for _ in range(4)]
    DJ = np.stack([0.05 * (A[0] + A[0].T), 0.05 * (A[1] + A[1].T)])
    DK = np.stack([0.05 * (A[2] + A[2].T), 0.05 * (A[3] + A[3].T)])
    G = gb.jk(DJ, DK, 0.0)
    Gm = gb.jk(DJ, DK, 0.0, qb.QBX_PREC_MIXED)
    for s in range(2):
        Gref = oracle.gcore(E, DJ[s], DK[s])
        e64 = float(np.max(np.abs(G[s] - Gref)))
        emx = float(np.max(np.abs(Gm[s] - Gref)))
        _say("  set %d: max|G - G_oracle| %.3e, mixed %.3e" % (s, e64, emx))
        assert e64 < 1e-11
        assert emx < 1e-5 * max(1.0, float(np.max(np.abs(Gref))))
    # Schwarz screening at 1e-12 changes nothing measurable
    G12 = gb.jk(DJ, DK, 1e-12)
    assert float(np.max(np.abs(G12 - G))) < 1e-10


@pytest.mark.parametrize("system", ["water_ccpvtz", "n2_ccpvtz"])
def test_rhf_with_f_shells_matches_oracle(system):
    nuc, bs = _n2() if system == "n2_ccpvtz" else _water("cc-pVTZ")
    cfg = qb.HFconfig(initial="CoreH", strategy=qb.SCFconfig(("ADIIS", "DIIS"), (1e-4, 1e-10)), maxStep=200)
    gb = qb.GPUBasis(bs)
    eng = OracleEngine(gb.flat)
    eng._eri = oracle_eri_dense(eng.ob, gb.n, THREADS)
    _say("%s RHF: GPU direct vs oracle dense" % system)
    res = qb.runHartreeFock(nuc, gb, cfg, printInfo=False)
    ref = qb.runHartreeFock(nuc, eng, cfg, printInfo=False)
    _say("  E_gpu %.12f E_oracle %.12f (converged %s %s)" % (res.energy[0], ref.energy[0], res.converged,
                                                             ref.converged))
    assert res.converged and ref.converged
    assert abs(res.energy[0] - ref.energy[0]) < 1e-10
