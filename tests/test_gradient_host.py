"""Host side of the basis-parameter gradient: the parametrization and the
normalization derivatives used by the chain rule."""
import math

import numpy as np
import pytest

from quiqbox_amd.basis import PrimitiveOrb, flatten_basis, get3DimPGTOrbNormFactor, prim_overlap
from quiqbox_amd.gradient import ParamBasis, _lnS_derivatives
from quiqbox_amd.systems import h_cluster_sp


def test_param_basis_reproduces_the_config_generator():
    nuc, bs = h_cluster_sp(32, 0)
    pb = ParamBasis.random_h_cluster(32, 0)
    a, b = flatten_basis(pb.orbitals()), flatten_basis(bs)
    assert a.n_bf == b.n_bf == 128
    np.testing.assert_array_equal(a.prim_xpn, b.prim_xpn)
    np.testing.assert_array_equal(a.bf_w, b.bf_w)
    np.testing.assert_array_equal(a.prim_cen, b.prim_cen)
    assert pb.n_theta == 608
    th = pb.theta()
    pb2 = pb.with_theta(th)
    np.testing.assert_array_equal(pb2.theta(), th)
    with pytest.raises(ValueError):
        pb.with_theta(th[:-1])


def test_lnS_derivatives_closed_form_single_s_primitive_pair():
    # two s primitives at one centre: S = sum_kl c_k c_l N_k N_l (pi/(a_k+a_l))^{3/2}
    # with N(a) = (2a/pi)^{3/4}; d ln S/dc_k = 2 sum_l c_l M_kl / S
    x = np.array([0.7, 2.3])
    c = np.array([0.4, 0.9])
    N = (2 * x / math.pi) ** 0.75
    M = np.array([[N[k] * N[l] * (math.pi / (x[k] + x[l])) ** 1.5 for l in range(2)] for k in range(2)])
    S = c @ M @ c
    dc, dx = _lnS_derivatives(0, x, c)
    np.testing.assert_allclose(dc, 2 * (M @ c) / S, rtol=1e-8)
    # d M_kl / d a_k = M_kl (3/(4 a_k) - 3/(2 (a_k + a_l)))
    dM = np.array([[M[k, l] * (0.75 / x[k] - 1.5 / (x[k] + x[l])) for l in range(2)] for k in range(2)])
    np.testing.assert_allclose(dx, 2 * c * (dM @ c) / S, rtol=1e-7)


def _lnS_brute(L_ijk, x, c):
    """ln S_mumu from the package's primitive norms and overlaps, summed term by term."""
    S = 0.0
    for a, ca in zip(x, c):
        for b, cb in zip(x, c):
            S += ca * cb * get3DimPGTOrbNormFactor(a, L_ijk) * get3DimPGTOrbNormFactor(b, L_ijk) * \
                prim_overlap(PrimitiveOrb((0.0, 0.0, 0.0), a, L_ijk), PrimitiveOrb((0.0, 0.0, 0.0), b, L_ijk))
    return math.log(S)


@pytest.mark.parametrize("ijk", [(0, 0, 0), (1, 0, 0), (0, 0, 1), (2, 0, 0), (1, 1, 0), (0, 1, 1), (2, 1, 0),
                                 (1, 1, 1)])
def test_lnS_derivatives_against_central_differences_for_every_component(ijk):
    """The closed form (one result per shell) against central differences of
    ln S built term by term for each Cartesian component: the ijk factors of
    the norms and overlaps cancel in ln S."""
    L = sum(ijk)
    rng = np.random.default_rng(L * 7 + ijk[0])
    x = np.exp(rng.uniform(np.log(0.1), np.log(20.0), size=4))
    c = rng.uniform(0.1, 1.0, size=4)
    dc, dx = _lnS_derivatives(L, x, c)
    h = 1e-5
    for k in range(4):
        cp, cm = c.copy(), c.copy()
        cp[k] *= 1 + h
        cm[k] *= 1 - h
        fd_c = (_lnS_brute(ijk, x, cp) - _lnS_brute(ijk, x, cm)) / (2 * h * c[k])
        xp, xm = x.copy(), x.copy()
        xp[k] *= 1 + h
        xm[k] *= 1 - h
        fd_x = (_lnS_brute(ijk, xp, c) - _lnS_brute(ijk, xm, c)) / (2 * h * x[k])
        assert dc[k] == pytest.approx(fd_c, rel=1e-7, abs=1e-9)
        assert dx[k] == pytest.approx(fd_x, rel=1e-7, abs=1e-9)
