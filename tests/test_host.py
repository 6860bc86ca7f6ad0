"""Host-side logic without a GPU: basis construction, flattening and the
C-ABI library's exported symbols."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import quiqbox_amd as qb
from quiqbox_amd import _lib
from quiqbox_amd.basis import prim_overlap, primitive_expansion
from helpers import OracleEngine, contracted, literals

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_header_symbol():
    header = open(os.path.join(ROOT, "include", "qbx.h")).read()
    names = set(re.findall(r"\b(qbx_[a-z0-9_]+)\s*\(", header))
    assert len(names) >= 18
    assert os.path.exists(_lib.LIB_PATH), "libqbx_hip.so not built"
    L = ctypes.CDLL(_lib.LIB_PATH)
    for n in sorted(names):
        assert hasattr(L, n), n
    assert names == set(_lib.SIGNATURES), names ^ set(_lib.SIGNATURES)


def test_norm_factor_normalizes():
    for L in range(4):
        for ijk in qb.subshellXYZs(L):
            a = 0.7
            p = qb.PrimitiveOrb((0.1, 0.2, 0.3), a, ijk)
            n = qb.get3DimPGTOrbNormFactor(a, ijk)
            assert n * n * prim_overlap(p, p) == pytest.approx(1.0, rel=1e-14)


def test_renormalization_semantics():
    # Overlap-test.jl:74-94: inner/outer renormalization combinations
    cen, xp, co = (1.0, 1.5, 1.1), [1.5, 0.6], [1.0, 0.8]
    g = lambda i, o: OracleEngine([qb.genGaussTypeOrb(cen, xp, co, (1, 0, 0), innerRenormalize=i,
                                                      outerRenormalize=o)]).overlap()[0, 0]
    assert g(False, True) == pytest.approx(1.0, rel=1e-14)
    assert g(True, True) == pytest.approx(1.0, rel=1e-14)
    assert not math.isclose(g(True, False), 1.0, rel_tol=1e-6)
    e = OracleEngine([qb.genGaussTypeOrb(cen, xp, co, (1, 0, 0)),
                      qb.genGaussTypeOrb(cen, xp, co, (1, 0, 0), outerRenormalize=True)])
    S = e.overlap()
    assert S[0, 0] == pytest.approx(S[0, 1] ** 2, rel=1e-13)


def test_primitive_renormalize_weight():
    p = qb.genGaussTypeOrb((0.1, -0.2, -0.3), 1.5, (1, 0, 0), renormalize=True)
    ((pr, w),) = primitive_expansion(p)
    assert w * w * prim_overlap(pr, pr) == pytest.approx(1.0, rel=1e-14)


def test_basis_library_parsing():
    assert len(qb.genGaussTypeOrbSeq((0.0, 0.0, 0.0), "H", "STO-3G")) == 1
    assert len(qb.genGaussTypeOrbSeq((0.0, 0.0, 0.0), "O", "STO-3G")) == 5
    assert len(qb.genGaussTypeOrbSeq((0.0, 0.0, 0.0), "Li", "3-21G")) == 9
    assert len(qb.genGaussTypeOrbSeq((0.0, 0.0, 0.0), "O", "6-31G")) == 9
    assert len(qb.genGaussTypeOrbSeq((0.0, 0.0, 0.0), "C", "cc-pVDZ")) == 15  # Cartesian d: 3s 2p 1d
    with pytest.raises(ValueError):
        qb.genGaussTypeOrbSeq((0.0, 0.0, 0.0), "K", "cc-pVDZ")


def test_flattening_dedups_and_marks_standard():
    bs = qb.genGaussTypeOrbSeq((0.0, 0.0, 0.0), "O", "6-31G")
    f = qb.flatten_basis(bs)
    assert f.n_bf == 9 and f.is_standard()
    # the p components of one SP shell share exponents but not Cartesian exponents
    assert f.n_prim == 6 + 3 + 3 * 3 + 1 + 3
    mixed = qb.CompositeOrb([qb.PrimitiveOrb((0, 0, 0), 1.0), qb.PrimitiveOrb((0, 0, 1), 1.0)], [1.0, 1.0])
    assert not qb.flatten_basis([mixed]).is_standard()


def test_expand_to_standard_roundtrip():
    from quiqbox_amd.basis import expand_to_standard
    mixed = qb.CompositeOrb([qb.PrimitiveOrb((0, 0, 0), 1.0), qb.PrimitiveOrb((0, 0, 1), 1.3, (1, 0, 0))],
                            [1.0, -0.4])
    bs = [mixed] + qb.genGaussTypeOrbSeq((0.3, 0.0, 0.0), "H", "STO-3G")
    f = qb.flatten_basis(bs)
    aux, T = expand_to_standard(f)
    S_direct = OracleEngine(f).overlap()
    S_aux = OracleEngine(aux).overlap()
    np.testing.assert_allclose(T.T @ S_aux @ T, S_direct, rtol=1e-13, atol=1e-15)


def test_spin_configuration_and_nuclear_repulsion():
    nuc = qb.NuclearCluster(["O", "H", "H"], [(0, 0, 0), (1.4, 0, 0), (0, 1.4, 0)])
    assert nuc.symbols == ["H", "H", "O"]
    s = qb.prepareSpinConfiguration(nuc)
    assert list(s) == [5, 5]
    with pytest.raises(ValueError):
        qb.prepareSpinConfiguration(nuc, 1)
    assert qb.nucRepulsion(qb.NuclearCluster(["H", "H"], [(0, 0, 0), (2.0, 0, 0)])) == pytest.approx(0.5)


def test_workload_generators():
    from quiqbox_amd.systems import random_cluster, h10_sto3g
    nuc, bs = random_cluster(64, 0)
    assert len(bs) == 256 and int(sum(nuc.charges)) % 2 == 0
    nuc3, bs3 = random_cluster(160, 0, with_d=True)
    assert len(bs3) == 592 and int(sum(nuc3.charges)) == 624
    d = np.array(nuc3.coords)
    r = np.sqrt(((d[:, None] - d[None]) ** 2).sum(-1)) + np.eye(len(d)) * 99
    assert r.min() >= 1.8
    nuc1, bs1 = h10_sto3g()
    assert len(bs1) == 10


def test_library_built_from_these_sources():
    """libqbx_hip.so reports the digest of the sources it was built from
    (qbx_source_digest); it must be the digest of the sources in this tree,
    i.e. the library was rebuilt after the last source change."""
    from quiqbox_amd import _lib
    prov = _lib.provenance()
    assert prov["sources"] is not None and len(prov["library"]) == 16
    assert prov["match"], "libqbx_hip.so was built from other sources (%s, tree %s): run build()" % (
        prov["library"], prov["sources"])
