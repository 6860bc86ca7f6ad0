"""Textbook pin: H2 in STO-3G (zeta = 1.24) at R = 1.4 bohr, the worked
example of Szabo and Ostlund, Modern Quantum Chemistry, section 3.5.2.  The
book lists the integrals and the SCF result to four decimals: S12 = 0.6593,
T11 = 0.7600, T12 = 0.2365, V11(A) = -1.2266, V12(A) = -0.5974, V22(A) =
-0.6538, (11|11) = 0.7746, (11|22) = 0.5697, (21|21) = 0.2970, (21|11) = 0.4441,
orbital energies -0.5782 and 0.6703, total energy -1.1167 Eh.  Checked for the
oracle (CPU) and for the GPU path."""
import numpy as np
import pytest

import quiqbox_amd as qb
from helpers import OracleEngine, molecule_basis

NUCS = ["H", "H"]
COORDS = [(0.0, 0.0, 0.0), (0.0, 0.0, 1.4)]
TOL = 1e-4      # the book's four decimals (rounded)


def _check_integrals(S, T, VA, E):
    assert S[0, 1] == pytest.approx(0.6593, abs=TOL)
    assert T[0, 0] == pytest.approx(0.7600, abs=TOL)
    assert T[0, 1] == pytest.approx(0.2365, abs=TOL)
    assert VA[0, 0] == pytest.approx(-1.2266, abs=TOL)
    assert VA[0, 1] == pytest.approx(-0.5974, abs=TOL)
    assert VA[1, 1] == pytest.approx(-0.6538, abs=TOL)
    assert E[0, 0, 0, 0] == pytest.approx(0.7746, abs=TOL)
    assert E[0, 0, 1, 1] == pytest.approx(0.5697, abs=TOL)
    assert E[1, 0, 1, 0] == pytest.approx(0.2970, abs=TOL)
    assert E[1, 0, 0, 0] == pytest.approx(0.4441, abs=TOL)


def _check_scf(res):
    assert res.converged
    assert sum(res.energy) == pytest.approx(-1.1167, abs=TOL)
    np.testing.assert_allclose(res.occu[0][0], [-0.5782, 0.6703], atol=TOL)


def _run(two_body_source):
    nuc = qb.NuclearCluster(NUCS, COORDS)
    cfg = qb.HFconfig(initial="CoreH", strategy=qb.SCFconfig(("DIIS",), (1e-12,)))
    return qb.runHartreeFock(nuc, two_body_source, cfg, printInfo=False)


def test_oracle_h2_sto3g_textbook():
    bs = molecule_basis(NUCS, COORDS, "STO-3G")
    eng = OracleEngine(bs)
    R = np.array(COORDS)
    _check_integrals(eng.overlap(), eng.kinetic(), eng.nuclear([1.0], R[:1]), eng.eri())
    _check_scf(_run(eng))


@pytest.mark.gpu
def test_gpu_h2_sto3g_textbook():
    bs = molecule_basis(NUCS, COORDS, "STO-3G")
    gb = qb.GPUBasis(bs)
    R = np.array(COORDS)
    _check_integrals(gb.overlap(), gb.kinetic(), gb.nuclear([1.0], R[:1]), gb.eri_dense())
    _check_scf(_run(gb))
