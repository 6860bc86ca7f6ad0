"""The Julia glue's flattening (the `Basis(data::OrbitalCoreData)` method of
the integration guide), restated in Python over a structure shaped like the
reference's OrbitalCoreData: `source` a MemoryPair of deduplicated primitive
fields and their categories, `format` one OrbCorePointer per basis function
whose `inner` MemoryPair holds (OneToIndex, weight) pairs (Framework.jl:134-173,
OrbitalBases.jl:439-451).  The loop below is the glue's, statement for
statement; it must reproduce FlatBasis (the package's own flattening) and,
through qbx_basis_create, the same integrals."""
from dataclasses import dataclass

import numpy as np
import pytest

import quiqbox_amd as qb
from helpers import molecule_basis
from quiqbox_amd.basis import primitive_expansion


@dataclass(frozen=True)
class OneToIndex:          # Query.jl:4-16: a 1-based index
    idx: int


@dataclass
class MemoryPair:          # Query.jl:692-705: two equal-length memories
    left: list
    right: list


@dataclass
class OrbCorePointer:      # Framework.jl:134-153
    inner: MemoryPair


@dataclass
class PrimGaussTypeOrbInfo:   # GaussianOrbitals.jl:6-16
    cen: tuple
    xpn: float
    ang: tuple


@dataclass
class OrbitalCoreData:     # Framework.jl:165-169
    source: MemoryPair     # fields => category
    format: list           # OrbCorePointer per basis function


def orbital_core_data(orbs):
    """What initializeHartreeFock hands to computeOrbDataIntegral: the fields of
    MultiOrbitalData (deduplicated, first-appearance order) and the weights
    getOrbCorePointers built (normalization of buildOrbCoreWeight!)."""
    fields, index, fmt = [], {}, []
    for orb in orbs:
        left, right = [], []
        for p, w in primitive_expansion(orb):
            k = p.key()
            if k not in index:
                fields.append(PrimGaussTypeOrbInfo(tuple(k[0]), k[1], tuple(k[2])))
                index[k] = len(fields)          # 1-based
            left.append(OneToIndex(index[k]))
            right.append(w)
        fmt.append(OrbCorePointer(MemoryPair(left, right)))
    return OrbitalCoreData(MemoryPair(fields, ["PrimGaussTypeOrb"] * len(fields)), fmt)


def glue_flatten(data):
    """The glue's loop: centres, exponents and Cartesian exponents of the
    fields, then per function its (idx - 1, weight) pairs in CSR form."""
    prims = data.source.left
    cen = [x for p in prims for x in p.cen]
    xpn = [p.xpn for p in prims]
    ang = [a for p in prims for a in p.ang]
    bfptr, bfprim, bfw = [0], [], []
    for ptr in data.format:
        idxs, ws = ptr.inner.left, ptr.inner.right
        for k in range(len(idxs)):
            bfprim.append(idxs[k].idx - 1)
            bfw.append(ws[k])
        bfptr.append(len(bfprim))
    return qb.FlatBasis(np.array(cen, dtype=np.float64).reshape(-1, 3), np.array(xpn, dtype=np.float64),
                        np.array(ang, dtype=np.int32).reshape(-1, 3), np.array(bfptr, dtype=np.int32),
                        np.array(bfprim, dtype=np.int32), np.array(bfw, dtype=np.float64))


def _bases():
    L = 0.7
    yield "H2O2 6-31G", molecule_basis(["O", "O", "H", "H"], [(0, 0, 0), (2.7, 0, 0), (-0.4, 1.7, 0),
                                                              (3.1, 1.7, 0.9)], "6-31G")
    yield "water cc-pVDZ", molecule_basis(["O", "H", "H"], [(0, 0, 0), (1.8, 0, 0), (-0.45, 1.75, 0)], "cc-pVDZ")
    # renormalized contractions (outerRenormalize) and a renormalized primitive
    yield "renormalized", [qb.genGaussTypeOrb((0, 0, L), [1.3, 0.4], [0.6, 0.5], (1, 0, 0), outerRenormalize=True),
                           qb.genGaussTypeOrb((0, 0, -L), [0.9], [1.0], (0, 0, 0), innerRenormalize=True)]


@pytest.mark.parametrize("case", range(3))
def test_glue_flattening_matches_flatbasis(case):
    name, orbs = list(_bases())[case]
    flat = qb.flatten_basis(orbs)
    glue = glue_flatten(orbital_core_data(orbs))
    assert glue.n_bf == flat.n_bf
    np.testing.assert_array_equal(glue.prim_cen, flat.prim_cen)
    np.testing.assert_array_equal(glue.prim_xpn, flat.prim_xpn)
    np.testing.assert_array_equal(glue.prim_ang, flat.prim_ang)
    np.testing.assert_array_equal(glue.bf_ptr, flat.bf_ptr)
    np.testing.assert_array_equal(glue.bf_prim, flat.bf_prim)
    np.testing.assert_array_equal(glue.bf_w, flat.bf_w)
    # 0-based indices inside the primitive table
    assert glue.bf_prim.min() == 0 and glue.bf_prim.max() == glue.n_prim - 1


@pytest.mark.gpu
def test_glue_flattening_through_the_abi():
    for name, orbs in _bases():
        glue = glue_flatten(orbital_core_data(orbs))
        gb = qb.GPUBasis(glue)
        ref = qb.GPUBasis(orbs)
        print("[gpu-test] glue flattening %s: %d functions" % (name, gb.n), flush=True)
        assert np.max(np.abs(gb.overlap() - ref.overlap())) == 0.0
        n = gb.n
        rng = np.random.default_rng(1)
        A = rng.standard_normal((n, n))
        D = 0.05 * (A + A.T)
        assert np.max(np.abs(gb.jk(2 * D, D, 0.0) - ref.jk(2 * D, D, 0.0))) < 1e-13
