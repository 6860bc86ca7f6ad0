"""Public integral functions with Quiqbox's names and argument meaning
(src/Integration/Interface.jl).  Every value is computed by libqbx_hip.so.

The `lazyCompute` keyword is accepted for drop-in compatibility; the GPU path
evaluates every requested integral directly (Quiqbox's lazy and eager paths
are bit-identical by contract, Coulomb-test.jl:102,149).
"""
import collections
import hashlib
import threading

import numpy as np

from .basis import flatten_basis
from .engine import GPUBasis, default_context
from .particles import NuclearCluster, getCharge

# Small bases (the single-element calls: two or four functions) are kept on
# the GPU while their flattened content is unchanged, so repeated calls on the
# same functions build them once; the oldest is dropped beyond _CACHE_MAX.
_CACHE_MAX_BF = 64
_CACHE_MAX = 32
_cache = collections.OrderedDict()
_cache_lock = threading.Lock()


def _basis(orbs):
    flat = flatten_basis(list(orbs))
    if flat.n_bf > _CACHE_MAX_BF:
        return GPUBasis(flat)
    h = hashlib.sha1()
    for a in (flat.prim_cen, flat.prim_xpn, flat.prim_ang, flat.bf_ptr, flat.bf_prim, flat.bf_w):
        a = np.ascontiguousarray(a)
        h.update(str(a.dtype).encode() + str(a.shape).encode())
        h.update(a.tobytes())
    key = h.digest()
    with _cache_lock:
        gb = _cache.get(key)
        if gb is not None:
            _cache.move_to_end(key)
            return gb
    gb = GPUBasis(flat)
    with _cache_lock:
        _cache[key] = gb
        while len(_cache) > _CACHE_MAX:
            _cache.popitem(last=False)
    return gb


def _nuclei(nucs, coords):
    if isinstance(nucs, NuclearCluster):
        return nucs.charges, nucs.coord_array
    Z = np.array([getCharge(s) for s in nucs], dtype=np.float64)
    R = np.array([tuple(c) for c in coords], dtype=np.float64).reshape(-1, 3)
    if len(Z) != len(R):
        raise ValueError("nuclei and coordinates differ in length")
    return Z, R


def overlaps(basisSet, lazyCompute=True):
    """Interface.jl:46-54"""
    return _basis(basisSet).overlap()


def overlap(orbL, orbR, lazyCompute=True):
    """Interface.jl:21-33"""
    return _basis([orbL, orbR]).overlap()[0, 1]


def multipoleMoments(center, degrees, basisSet, lazyCompute=True):
    """Interface.jl:94-105"""
    return _basis(basisSet).multipole(center, degrees)


def multipoleMoment(center, degrees, orbL, orbR, lazyCompute=True):
    """Interface.jl:71-80"""
    return _basis([orbL, orbR]).multipole(center, degrees)[0, 1]


def elecKinetics(basisSet, lazyCompute=True):
    """Interface.jl:143-153"""
    return _basis(basisSet).kinetic()


def elecKinetic(orbL, orbR, lazyCompute=True):
    """Interface.jl:120-130"""
    return _basis([orbL, orbR]).kinetic()[0, 1]


def nucAttractions(nucs, coords_or_basis, basisSet=None, lazyCompute=True):
    """nucAttractions(nucs, coords, bs) or nucAttractions(nucInfo, bs) (Interface.jl:236-256)"""
    if basisSet is None:
        Z, R = _nuclei(nucs, None)
        basisSet = coords_or_basis
    else:
        Z, R = _nuclei(nucs, coords_or_basis)
    return _basis(basisSet).nuclear(Z, R)


def nucAttraction(nucs, *args, lazyCompute=True):
    """nucAttraction(nucs, coords, orbL, orbR) or nucAttraction(nucInfo, orbL, orbR) (Interface.jl:186-205)"""
    if isinstance(nucs, NuclearCluster):
        orbL, orbR = args
        Z, R = _nuclei(nucs, None)
    else:
        coords, orbL, orbR = args
        Z, R = _nuclei(nucs, coords)
    return _basis([orbL, orbR]).nuclear(Z, R)[0, 1]


def coreHamiltonian(nucs, coords_or_basis, basisSet=None, lazyCompute=True):
    """kinetic + nuclear attraction (Interface.jl:288-312, Samplers.jl:110-131)"""
    if basisSet is None:
        Z, R = _nuclei(nucs, None)
        basisSet = coords_or_basis
    else:
        Z, R = _nuclei(nucs, coords_or_basis)
    return _basis(basisSet).core_hamiltonian(Z, R)


def elecRepulsions(basisSet, lazyCompute=True):
    """Dense (ij|kl) tensor, chemists' notation (Interface.jl:345-365)."""
    return _basis(basisSet).eri_dense()


def elecRepulsion(orb1, orb2, orb3, orb4, lazyCompute=True):
    """(12|34) (Interface.jl:320-342)"""
    return _basis([orb1, orb2, orb3, orb4]).eri_dense()[0, 1, 2, 3]


def changeOrbitalBasis(DbodyInt, *C):
    """Transform one- or two-body AO integrals with orbital coefficients
    (Interface.jl:376-406) on the GPU, one matrix-core GEMM per index.  With
    two coefficient matrices and a 4-index tensor: (ERI_alpha, ERI_beta,
    J_alphabeta)."""
    ctx = default_context()
    A = np.asarray(DbodyInt)
    if len(C) == 1:
        if A.ndim not in (2, 4):
            raise ValueError("one- or two-body integrals expected")
        return ctx.mo_transform(A, C[0])
    if len(C) == 2 and A.ndim == 4:
        return ctx.mo_transform(A, C[0]), ctx.mo_transform(A, C[1]), ctx.mo_coulomb_ab(A, C[0], C[1])
    raise ValueError("unsupported changeOrbitalBasis arguments")
