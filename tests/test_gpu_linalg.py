"""FP64 matrix-core GEMM (qbx_gemm_f64) against numpy (ragged sizes with 32 x 32
blocks; 2113 x 2049 and 70000 x 100 with 64 x 64 blocks), the purified density
(qbx_density_purify) against the eigenvector density of the reference's SCF
step (eigen of X' F X, D = C_occ C_occ^T), and SCF runs with
densitySolver="purify" against the eigen steps (same energies to 1e-10 Eh)."""
import ctypes

import numpy as np
import pytest

import quiqbox_amd as qb
from helpers import literals, molecule_basis
from quiqbox_amd import _lib
from quiqbox_amd.engine import default_context
from quiqbox_amd.hartree_fock import getC, getD, getOrthonormalization

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _say(msg):
    print("[gpu-test] " + msg, flush=True)


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(37, 53, 29), (64, 64, 64), (1, 70, 5), (131, 17, 200),
                                   (2113, 2049, 45), (70000, 100, 24)])
def test_gemm_f64_all_transposes_ragged_sizes(ta, tb, M, N, K):
    rng = np.random.default_rng(M * 7 + N + K + 2 * ta + tb)
    A = rng.normal(size=(K, M) if ta else (M, K))
    B = rng.normal(size=(N, K) if tb else (K, N))
    C0 = rng.normal(size=(M, N))
    dA, dB, dC = (torch.tensor(x, dtype=torch.float64).cuda() for x in (A, B, C0))
    _say("gemm_f64 ta=%d tb=%d %dx%dx%d" % (ta, tb, M, N, K))
    L = _lib.load()
    _lib.check(L.qbx_gemm_f64(ta, tb, M, N, K, 1.5, _p(dA), A.shape[1], _p(dB), B.shape[1], -0.5, _p(dC), N, None))
    torch.cuda.synchronize()
    ref = 1.5 * (A.T if ta else A) @ (B.T if tb else B) - 0.5 * C0
    got = dC.cpu().numpy()
    # FP64 products, FP64 sums in another order: a few ulps of the row-column dot products
    bound = 1e-13 * (np.abs(A.T if ta else A) @ np.abs(B.T if tb else B) + np.abs(C0))
    assert np.all(np.abs(got - ref) <= bound + 1e-300)


def test_gemm_f64_rejects_bad_leading_dimension():
    L = _lib.load()
    d = torch.zeros(16, dtype=torch.float64).cuda()
    rc = L.qbx_gemm_f64(0, 0, 4, 4, 4, 1.0, _p(d), 3, _p(d), 4, 0.0, _p(d), 4, None)
    assert rc == -1      # QBX_EINVAL, before any launch


def _projector_case(n, nocc, seed, gap=0.5):
    """Symmetric F = Q diag(lam) Q' with a gap above eigenvalue nocc, and an
    SPD orthonormalization X; the reference density C_occ C_occ' from eigh."""
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    lam = np.sort(np.concatenate([rng.uniform(-20.0, -gap / 2, nocc), rng.uniform(gap / 2, 15.0, n - nocc)]))
    Fo = (Q * lam) @ Q.T
    B = rng.normal(size=(n, n)) / np.sqrt(n)
    S = np.eye(n) + 0.3 * (B @ B.T)
    X = getOrthonormalization(S)
    Xi = np.linalg.inv(X)
    F = Xi.T @ Fo @ Xi          # X' F X = Fo
    return X, F


@pytest.mark.parametrize("n,nocc", [(40, 7), (257, 100), (600, 120)])
def test_purified_density_equals_eigenvector_density(n, nocc):
    X, F = _projector_case(n, nocc, seed=n)
    ctx = default_context()
    _say("purify n=%d nocc=%d" % (n, nocc))
    D, it, idem = ctx.purified_density(X, F, nocc)
    Dref = getD(getC(X, F), nocc)
    _say("  %d products, idempotency %.2e, max |D - D_eigen| %.2e" % (it, idem, np.abs(D - Dref).max()))
    assert np.abs(D - Dref).max() < 1e-10 * max(1.0, np.abs(Dref).max())
    assert it < 100


def test_purified_density_without_gap_raises():
    n, nocc = 30, 10
    rng = np.random.default_rng(3)
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    lam = np.sort(rng.uniform(-5.0, 5.0, n))
    lam[nocc] = lam[nocc - 1]            # degenerate at the Fermi level
    F = (Q * lam) @ Q.T
    with pytest.raises(_lib.QbxError):
        default_context().purified_density(np.eye(n), F, nocc, max_iter=60)


def test_scf_with_purified_density_matches_eigen_steps():
    L = literals()["h2o2_631g"]
    coords = [tuple(L["angstrom_to_bohr"] * x for x in c) for c in L["coords_angstrom"]]
    nuc = qb.NuclearCluster(L["nucs"], coords)
    bs = molecule_basis(nuc.symbols, nuc.coords, "6-31G")
    gb = qb.GPUBasis(bs)
    cfg = qb.HFconfig(initial="CoreH", strategy=qb.SCFconfig(("ADIIS", "DIIS"), (1e-4, 1e-10)))
    _say("H2O2 6-31G RHF eigen vs purify")
    ref = qb.runHartreeFock(nuc, gb, cfg, printInfo=False)
    res = qb.runHartreeFock(nuc, gb, cfg, printInfo=False, densitySolver="purify")
    _say("  E eigen %.12f purify %.12f, %d purified steps, %d fallbacks" % (
        ref.energy[0], res.energy[0], res.timing["purified_steps"], res.timing["purify_fallbacks"]))
    assert ref.converged and res.converged
    assert abs(res.energy[0] - ref.energy[0]) < 1e-10
    assert res.timing["purified_steps"] > 0 and res.timing["purify_fallbacks"] == 0
    # the result's coefficients are the eigenvectors of a Fock matrix: C' S C = I
    S = gb.overlap()
    C = res.coeff[0]
    np.testing.assert_allclose(C.T @ S @ C, np.eye(len(C)), atol=1e-9)
    np.testing.assert_allclose(res.density[0], getD(C, 9), atol=1e-8)


def test_uhf_with_purified_density_matches_eigen_steps():
    L = literals()["hho_sto3g"]
    nuc = qb.NuclearCluster(L["nucs"], L["coords"])
    bs = molecule_basis(L["nucs"], L["coords"], "STO-3G")
    scf = qb.SCFconfig(("ADIIS", "DIIS"), (1e-4, 1e-10))
    cfg = qb.HFconfig(qb.UOHartreeFock(), initial="GWH", strategy=scf)
    _say("HHO UHF eigen vs purify")
    ref = qb.runHartreeFock(nuc, bs, cfg, printInfo=False)
    res = qb.runHartreeFock(nuc, bs, cfg, printInfo=False, densitySolver="purify")
    _say("  E eigen %.12f purify %.12f" % (ref.energy[0], res.energy[0]))
    assert ref.converged and res.converged
    assert abs(res.energy[0] - ref.energy[0]) < 1e-10


@pytest.mark.parametrize("system", ["water24_ccpvdz", "h2o2_631g"])
def test_fock_solve_matches_host_eigen_step(system):
    """qbx_fock_solve (X' F X and X C' on the matrix cores, rocSOLVER dsyevd)
    against the host step solveFockMatrix/getD (numpy eigh): orbital
    energies, column signs (C[0, j] >= 0), coefficients of nondegenerate
    orbitals and the occupied density."""
    from quiqbox_amd.hartree_fock import solveFockMatrix
    from quiqbox_amd.systems import water_box
    if system == "water24_ccpvdz":
        nuc, bs = water_box(24, 0, basis="cc-pVDZ")
    else:
        L = literals()["h2o2_631g"]
        coords = [tuple(L["angstrom_to_bohr"] * x for x in c) for c in L["coords_angstrom"]]
        nuc = qb.NuclearCluster(L["nucs"], coords)
        bs = molecule_basis(nuc.symbols, nuc.coords, "6-31G")
    gb = qb.GPUBasis(bs)
    X = gb.orthonormalization()
    H = gb.core_hamiltonian(nuc.charges, nuc.coord_array)
    nocc = int(round(sum(nuc.charges))) // 2
    # a Fock-like matrix: the core Hamiltonian plus the two-electron part of its guess density
    D0 = getD(getC(X, H), nocc)
    F = H + gb.jk(2 * D0, D0, 1e-12)
    C_ref, e_ref = solveFockMatrix(X, F)
    D_ref = getD(C_ref, nocc)
    C, e, D = default_context().fock_solve(X, F, nocc)
    err_e = float(np.max(np.abs(e - e_ref)))
    err_D = float(np.max(np.abs(D - D_ref)))
    gaps = np.minimum(np.diff(e_ref, prepend=-np.inf), np.diff(e_ref, append=np.inf))
    nondeg = gaps > 1e-6
    # the sign rule (C[0, j] >= 0, as the reference's) leaves the sign of a
    # column whose first coefficient vanishes by symmetry to rounding: such
    # columns (H2O2's C2 axis gives some) are compared up to sign
    dC = np.abs(C - C_ref)
    free = np.abs(C_ref[0, :]) < 1e-8
    dC[:, free] = np.minimum(dC[:, free], np.abs(C + C_ref)[:, free])
    err_C = float(np.max(dC[:, nondeg]))
    _say("%s (n %d): max|de| %.2e, max|dD| %.2e, max|dC| (nondegenerate) %.2e" % (system, gb.n, err_e, err_D, err_C))
    assert np.all(C[0, :] >= 0)
    assert err_e < 1e-11 * max(1.0, float(np.max(np.abs(e_ref))))
    assert err_D < 1e-10
    assert err_C < 1e-8


def test_scf_default_device_eigen_matches_host_eigen():
    """runHartreeFock's step on the GPU (densitySolver="eigen-device", the default from
    DEVICE_EIGEN_MIN_N functions on) against
    the same SCF with the host eigen step: same energy to 1e-10 Eh, same
    number of Fock builds (the trajectory only differs by rounding)."""
    L = literals()["h2o2_631g"]
    coords = [tuple(L["angstrom_to_bohr"] * x for x in c) for c in L["coords_angstrom"]]
    nuc = qb.NuclearCluster(L["nucs"], coords)
    gb = qb.GPUBasis(molecule_basis(nuc.symbols, nuc.coords, "6-31G"))
    cfg = qb.HFconfig(initial="CoreH", strategy=qb.SCFconfig(("ADIIS", "DIIS"), (1e-4, 1e-10)), maxStep=200)
    dev = qb.runHartreeFock(nuc, gb, cfg, printInfo=False, densitySolver="eigen-device")
    host = qb.runHartreeFock(nuc, gb, cfg, printInfo=False, densitySolver="eigen-host")
    _say("H2O2 SCF: device eigen E %.12f (%d device steps), host eigen E %.12f" % (
        dev.energy[0], dev.timing.get("device_eigen_steps", 0), host.energy[0]))
    assert dev.converged and host.converged
    assert dev.timing.get("device_eigen_steps", 0) > 0
    assert abs(dev.energy[0] - host.energy[0]) < 1e-10
