"""Synthetic molecular systems and basis sets for the measured workloads.

  h10_sto3g()                 linear H10 chain, 1.4 bohr spacing, STO-3G s (10 functions)
  random_cluster(n_atoms, seed, with_d)
                              seeded C/H/O cluster in a 20-bohr cube, minimum separation
                              1.8 bohr, closed shell; per atom 3-primitive s and p shells
                              with exponents log-uniform in [0.1, 50] and coefficients
                              uniform in [0.1, 1] (x primitive normalization), each
                              contraction normalized (w <- w / sqrt(S_mumu)).
                              with_d: element-dependent shells (H: s; C: s, p; O: s, p, d
                              with 2-primitive d, exponents log-uniform in [0.2, 5]) and
                              composition 50% H / 30% C / 20% O in seeded random order:
                              592 functions and 624 electrons at 160 atoms.
  h_cluster_sp(n_atoms, seed) random H cluster with floating 4-primitive s and p shells.
"""
import math

import numpy as np

from .basis import CompositeOrb, PrimitiveOrb, genGaussTypeOrb, get3DimPGTOrbNormFactor, subshellXYZs
from .particles import NuclearCluster, getCharge

H10_XPN = (3.42525091, 0.62391373, 0.16885540)
H10_CON = (0.15432897, 0.53532814, 0.44463454)


def h10_sto3g(n=10, spacing=1.4):
    """Linear H_n chain along x with STO-3G s functions (coefficient x primitive norm)."""
    coords = [(spacing * k, 0.0, 0.0) for k in range(n)]
    nuc = NuclearCluster(["H"] * n, coords, pairwiseSort=False)
    bs = []
    for c in coords:
        cons = [d * get3DimPGTOrbNormFactor(a, (0, 0, 0)) for a, d in zip(H10_XPN, H10_CON)]
        bs.append(genGaussTypeOrb(c, list(H10_XPN), cons, (0, 0, 0)))
    return nuc, bs


def _place_atoms(rng, n, box, dmin):
    pts = []
    tries = 0
    while len(pts) < n:
        p = rng.uniform(0.0, box, size=3)
        tries += 1
        if tries > 1000000:
            raise RuntimeError("could not place %d atoms" % n)
        if all(np.sum((p - q) ** 2) >= dmin * dmin for q in pts):
            pts.append(p)
    return [tuple(float(x) for x in p) for p in pts]


def _shell(center, L, xpns, coefs):
    """Contracted functions of one shell, each normalized (outer renormalization)."""
    out = []
    for ijk in subshellXYZs(L):
        cons = [c * get3DimPGTOrbNormFactor(a, ijk) for a, c in zip(xpns, coefs)]
        out.append(genGaussTypeOrb(center, list(xpns), cons, ijk, outerRenormalize=True))
    return out


def random_cluster(n_atoms=64, seed=0, with_d=False, box=20.0, dmin=1.8):
    rng = np.random.Generator(np.random.PCG64(seed))
    if with_d:
        nH, nC = int(round(0.5 * n_atoms)), int(round(0.3 * n_atoms))
        elems = ["H"] * nH + ["C"] * nC + ["O"] * (n_atoms - nH - nC)
        elems = [elems[k] for k in rng.permutation(n_atoms)]
    else:
        elems = list(rng.choice(["C", "H", "O"], size=n_atoms))
    # closed shell: an even electron count needs an even number of H atoms
    if sum(getCharge(e) for e in elems) % 2 == 1:
        elems[-1] = "C" if elems[-1] == "H" else "H"
    coords = _place_atoms(rng, n_atoms, box, dmin)
    nuc = NuclearCluster(elems, coords, pairwiseSort=False)
    bs = []
    for e, c in zip(elems, coords):
        shells = ["s", "p"] if not with_d else {"H": ["s"], "C": ["s", "p"], "O": ["s", "p", "d"]}[e]
        for sh in shells:
            if sh == "d":
                xp = np.exp(rng.uniform(math.log(0.2), math.log(5.0), size=2))
                co = rng.uniform(0.1, 1.0, size=2)
                bs += _shell(c, 2, xp, co)
            else:
                xp = np.exp(rng.uniform(math.log(0.1), math.log(50.0), size=3))
                co = rng.uniform(0.1, 1.0, size=3)
                bs += _shell(c, 0 if sh == "s" else 1, xp, co)
    return nuc, bs


def h_cluster_sp(n_atoms=32, seed=0, box=10.0, dmin=1.2, nprim=4):
    """Random H cluster with floating nprim-primitive s and p shells."""
    rng = np.random.Generator(np.random.PCG64(seed))
    coords = _place_atoms(rng, n_atoms, box, dmin)
    nuc = NuclearCluster(["H"] * n_atoms, coords, pairwiseSort=False)
    bs = []
    for c in coords:
        for L in (0, 1):
            xp = np.exp(rng.uniform(math.log(0.1), math.log(10.0), size=nprim))
            co = rng.uniform(0.1, 1.0, size=nprim)
            bs += _shell(c, L, xp, co)
    return nuc, bs
