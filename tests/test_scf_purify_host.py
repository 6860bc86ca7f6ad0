"""Host side of densitySolver="purify": coefficients formed on first use, and
the diagonalization taking over where the device solver reports no gap."""
import numpy as np

from quiqbox_amd import _lib
from quiqbox_amd.hartree_fock import LazyCoeff, PurifiedDensity, getC, getD, getOrthonormalization


def _case(n=12, seed=0):
    rng = np.random.default_rng(seed)
    B = rng.normal(size=(n, n))
    S = np.eye(n) + 0.05 * (B @ B.T)
    F = rng.normal(size=(n, n))
    return getOrthonormalization(S), 0.5 * (F + F.T)


def test_lazy_coefficients_match_getC_and_act_as_arrays():
    X, F = _case()
    lc = LazyCoeff(X, F)
    assert lc._C is None
    C = getC(X, F)
    np.testing.assert_array_equal(np.asarray(lc), C)
    assert lc.value() is lc.value()


class _NoGap:
    def purified_density(self, X, F, nocc, max_iter):
        raise _lib.QbxError(-1, "purification did not converge")


class _Echo:
    """Stands in for the device: returns the eigenvector density, slightly asymmetric."""

    def purified_density(self, X, F, nocc, max_iter):
        D = getD(getC(X, F), nocc)
        D[0, 1] += 1e-14
        return D, 31, 1e-15


def test_purified_density_falls_back_to_diagonalization():
    X, F = _case(seed=1)
    solver = PurifiedDensity(_NoGap())
    D = solver(X, F, 4)
    np.testing.assert_allclose(D, getD(getC(X, F), 4), atol=1e-14)
    assert solver.calls == 1 and solver.fallbacks == 1 and solver.products == 0


def test_purified_density_is_symmetrized_and_counted():
    X, F = _case(seed=2)
    solver = PurifiedDensity(_Echo())
    D = solver(X, F, 5)
    np.testing.assert_array_equal(D, D.T)
    assert solver.calls == 1 and solver.fallbacks == 0 and solver.products == 31
