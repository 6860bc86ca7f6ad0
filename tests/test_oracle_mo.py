"""The oracle's orbital-basis change (changeOrbitalBasis, Interface.jl:376-380)
and opposite-spin Coulomb matrix (getJab, Interface.jl:382-393), pinned by
what the mathematics fixes: the identity, column selection (fixes which index
of C is contracted and the output index order), a diagonal scaling, an
orthogonal round trip, and brute-force loops on a tiny tensor.  The
reference's own check (MO-basis energy equals the SCF energy,
HartreeFock-test.jl:88-91,148-151) runs in test_scf_cpu.py."""
import itertools

import numpy as np

import oracle


def _tensor(n, seed):
    return np.random.default_rng(seed).standard_normal((n,) * 4)


def test_identity_and_column_selection():
    A = _tensor(6, 0)
    assert np.array_equal(oracle.change_orbital_basis(A, np.eye(6)), A)
    cols = [4, 1, 3]
    sel = np.eye(6)[:, cols]
    np.testing.assert_array_equal(oracle.change_orbital_basis(A, sel), A[np.ix_(cols, cols, cols, cols)])
    np.testing.assert_array_equal(oracle.change_orbital_basis(A[0, 0], sel), A[0, 0][np.ix_(cols, cols)])


def test_diagonal_scaling():
    A = _tensor(5, 1)
    s = np.array([1.0, -2.0, 0.5, 3.0, 0.25])
    out = oracle.change_orbital_basis(A, np.diag(s))
    np.testing.assert_allclose(out, A * np.einsum("i,j,k,l->ijkl", s, s, s, s), rtol=1e-14)


def test_orthogonal_round_trip():
    A = _tensor(7, 2)
    Q, _ = np.linalg.qr(np.random.default_rng(3).standard_normal((7, 7)))
    back = oracle.change_orbital_basis(oracle.change_orbital_basis(A, Q), Q.T)
    np.testing.assert_allclose(back, A, atol=1e-12)


def test_brute_force_tiny():
    n = 3
    rng = np.random.default_rng(4)
    A = rng.standard_normal((n,) * 4)
    C = rng.standard_normal((n, 2))
    C2 = rng.standard_normal((n, 4))
    out = oracle.change_orbital_basis(A, C)
    jab = oracle.coulomb_ab(A, C, C2)
    for i, j, k, l in itertools.product(range(2), repeat=4):
        ref = sum(A[a, b, c, d] * C[a, i] * C[b, j] * C[c, k] * C[d, l]
                  for a, b, c, d in itertools.product(range(n), repeat=4))
        assert abs(out[i, j, k, l] - ref) < 1e-12
    for m in range(2):
        for q in range(4):
            ref = sum(A[a, b, c, d] * C[a, m] * C[b, m] * C2[c, q] * C2[d, q]
                      for a, b, c, d in itertools.product(range(n), repeat=4))
            assert abs(jab[m, q] - ref) < 1e-12
