"""Test helpers: oracle-backed integral engine and golden-data loaders.

The oracle (oracle/liboracle.so) is the checker; it is only used from tests.
"""
import json
import os

import numpy as np

import oracle
import quiqbox_amd as qb

HERE = os.path.dirname(os.path.abspath(__file__))
_LIT = None
_BOYS = None


def literals():
    global _LIT
    if _LIT is None:
        with open(os.path.join(HERE, "golden", "reference_literals.json")) as f:
            _LIT = json.load(f)
    return _LIT


def boys_table():
    global _BOYS
    if _BOYS is None:
        with open(os.path.join(HERE, "golden", "boys_mpmath.json")) as f:
            _BOYS = json.load(f)
    return _BOYS


class OracleEngine:
    """CPU engine with the same duck type as GPUBasis (for runHartreeFock)."""

    def __init__(self, orbs):
        self.flat = orbs if isinstance(orbs, qb.FlatBasis) else qb.flatten_basis(orbs)
        self.n = self.flat.n_bf
        self.ob = oracle.Basis(self.flat.as_dict())
        self._eri = None

    def overlap(self):
        return self.ob.overlap()

    def kinetic(self):
        return self.ob.kinetic()

    def nuclear(self, Z, R):
        return self.ob.nuclear(Z, R)

    def core_hamiltonian(self, Z, R):
        return self.ob.kinetic() + self.ob.nuclear(Z, R)

    def eri(self):
        if self._eri is None:
            self._eri = self.ob.eri_dense()
        return self._eri

    def two_body(self, threshold=0.0, precision=0):
        return qb.DenseERI(self.eri())


def contracted(spec, innerRenormalize=False, outerRenormalize=False):
    return qb.genGaussTypeOrb(tuple(spec["center"]), spec["xpns"], spec["cons"], tuple(spec["ang"]),
                              innerRenormalize=innerRenormalize, outerRenormalize=outerRenormalize)


def lih_basis():
    L = literals()["lih"]
    return [contracted(b, innerRenormalize=True) for b in L["basis"]]


def molecule_basis(nucs, coords, key):
    bs = []
    for s, c in zip(nucs, coords):
        bs += qb.genGaussTypeOrbSeq(tuple(c), s, key)
    return bs


def random_spd_basis(seed, n_atoms=4, box=4.0, lmax=2, nprim=(3, 3, 2)):
    """Small seeded cluster with s, p (and d) shells on every atom."""
    rng = np.random.default_rng(seed)
    bs = []
    coords = []
    for a in range(n_atoms):
        c = tuple(rng.uniform(0, box, 3))
        coords.append(c)
        for L in range(lmax + 1):
            k = nprim[L]
            xp = np.exp(rng.uniform(np.log(0.15), np.log(20.0), k))
            co = rng.uniform(0.1, 1.0, k)
            for ijk in qb.subshellXYZs(L):
                cons = [cc * qb.get3DimPGTOrbNormFactor(x, ijk) for x, cc in zip(xp, co)]
                bs.append(qb.genGaussTypeOrb(c, list(xp), cons, ijk, outerRenormalize=True))
    return coords, bs


def oracle_eri_dense(ob, n, threads=16):
    """The oracle's dense ERI tensor from its unique quartets (i >= j, k >= l,
    ij >= kl) evaluated on several threads, scattered 8-fold."""
    ii, jj = np.tril_indices(n)
    m = len(ii)
    P, Q = np.tril_indices(m)
    q = np.stack([ii[P], jj[P], ii[Q], jj[Q]], axis=1).astype(np.int32)
    v = ob.eri_list(q, threads=threads)
    E = np.zeros((n, n, n, n))
    i, j, k, l = q.T
    for a, b, c, d in ((i, j, k, l), (j, i, k, l), (i, j, l, k), (j, i, l, k),
                       (k, l, i, j), (l, k, i, j), (k, l, j, i), (l, k, j, i)):
        E[a, b, c, d] = v
    return E
