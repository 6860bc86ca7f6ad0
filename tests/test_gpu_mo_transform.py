"""changeOrbitalBasis (Interface.jl:376-407) on the GPU: one FP64 matrix-core
GEMM per tensor index (qbx_mo_transform) and the opposite-spin Coulomb matrix
(qbx_mo_coulomb_ab).  Against the oracle on random tensors (square and
rectangular C), through the reference's own check that the RHF/UHF energy from
MO-basis integrals equals the SCF energy (HartreeFock-test.jl:88-91,148-151,
H2O2 STO-3G), and at n_ao = 130, where the first quarter's GEMM has 130^3 rows
(more row tiles than one grid dimension of 65535 holds), by an orthogonal
round trip and single elements contracted on the host."""
import ctypes

import numpy as np
import pytest

import oracle
import quiqbox_amd as qb
from helpers import literals, molecule_basis
from quiqbox_amd import _lib

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[gpu-test] " + msg, flush=True)


@pytest.mark.parametrize("nao,nmo", [(11, 11), (11, 7), (9, 13)])
def test_random_tensors_against_oracle(nao, nmo):
    rng = np.random.default_rng(nao * 100 + nmo)
    A4 = rng.standard_normal((nao,) * 4)
    A2 = rng.standard_normal((nao, nao))
    C = rng.standard_normal((nao, nmo))
    C2 = rng.standard_normal((nao, nmo + 2))
    ctx = qb.default_context()
    _say("mo transform n_ao %d n_mo %d" % (nao, nmo))
    for A in (A2, A4):
        got = ctx.mo_transform(A, C)
        ref = oracle.change_orbital_basis(A, C)
        err = float(np.abs(got - ref).max())
        _say("  rank %d: max err %.2e (max %.2e)" % (A.ndim, err, float(np.abs(ref).max())))
        assert err < 1e-12 * max(1.0, float(np.abs(ref).max()))
    got = ctx.mo_coulomb_ab(A4, C, C2)
    ref = oracle.coulomb_ab(A4, C, C2)
    err = float(np.abs(got - ref).max())
    _say("  Jab: max err %.2e (max %.2e)" % (err, float(np.abs(ref).max())))
    assert err < 1e-12 * max(1.0, float(np.abs(ref).max()))


def test_hho_mo_energies_match_scf():
    L = literals()["hho_sto3g"]
    nuc = qb.NuclearCluster(L["nucs"], L["coords"])
    bs = molecule_basis(L["nucs"], L["coords"], "STO-3G")
    gb = qb.GPUBasis(bs)
    H = gb.core_hamiltonian(nuc.charges, nuc.coord_array)
    E = gb.eri_dense()
    scf = qb.SCFconfig(("ADIIS", "DIIS", "EDIIS", "DD"), (1e-4, 1e-8, 1e-10, 1e-15))
    res = qb.runHartreeFock(nuc, gb, qb.HFconfig(initial="CoreH", strategy=scf), printInfo=False)
    C = res.coeff[0]
    e_mo = qb.getEhfMO((qb.changeOrbitalBasis(H, C),), (qb.changeOrbitalBasis(E, C),), (5,))
    _say("HHO RHF: E_scf %.12f E_mo %.12f" % (res.energy[0], e_mo))
    assert e_mo == pytest.approx(res.energy[0], abs=1e-12)
    assert e_mo == pytest.approx(L["E_rhf"], abs=L["tol_E"])
    ures = qb.runHartreeFock(nuc, gb, qb.HFconfig(qb.UOHartreeFock(), initial="GWH", strategy=scf), printInfo=False)
    a, b, c = qb.changeOrbitalBasis(E, *ures.coeff)
    e_mo = qb.getEhfMO(tuple(qb.changeOrbitalBasis(H, Cs) for Cs in ures.coeff), (a, b), (5, 5), c)
    _say("HHO UHF: E_scf %.12f E_mo %.12f" % (ures.energy[0], e_mo))
    assert e_mo == pytest.approx(ures.energy[0], abs=1e-12)


def test_n130_round_trip_and_elements():
    n = 130
    rng = np.random.default_rng(130)
    _say("n_ao 130: tensor of %.2f GB" % (8.0 * n ** 4 / 1e9))
    A = rng.standard_normal((n,) * 4)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ctx = qb.default_context()
    B = ctx.mo_transform(A, Q)
    for idx in ((0, 0, 0, 0), (129, 3, 77, 128), (5, 129, 0, 64)):
        ref = A
        for i in idx:
            ref = np.tensordot(Q[:, i], ref, axes=([0], [0]))
        _say("  element %s: %.15f host %.15f" % (idx, B[idx], float(ref)))
        assert abs(B[idx] - float(ref)) < 1e-11
    back = ctx.mo_transform(B, Q.T)
    err = float(np.abs(back - A).max())
    _say("  round trip max err %.2e" % err)
    assert err < 1e-11


def test_argument_errors():
    """QBX_EINVAL (-1) for a bad rank, an empty basis or a null array; ValueError
    from the host checks."""
    ctx = qb.default_context()
    L = _lib.load()
    x = np.zeros(16)
    pd = ctypes.POINTER(ctypes.c_double)
    p = x.ctypes.data_as(pd)
    assert L.qbx_mo_transform(ctx.h, 3, 2, 2, p, p, p) == -1      # rank must be 2 or 4
    assert L.qbx_mo_transform(ctx.h, 4, 0, 2, p, p, p) == -1      # empty basis
    assert L.qbx_mo_transform(ctx.h, 2, 2, 2, None, p, p) == -1   # null tensor
    assert L.qbx_mo_coulomb_ab(ctx.h, 2, 0, 2, p, p, p, p) == -1
    with pytest.raises(ValueError):
        ctx.mo_transform(np.zeros((3, 3, 3)), np.eye(3))                       # neither one- nor two-body
    with pytest.raises(ValueError):
        ctx.mo_transform(np.zeros((4, 4)), np.eye(3))                          # C rows != n_ao
    with pytest.raises(ValueError):
        qb.changeOrbitalBasis(np.zeros((3, 3)), np.eye(3), np.eye(3))          # two C only for 4-index
