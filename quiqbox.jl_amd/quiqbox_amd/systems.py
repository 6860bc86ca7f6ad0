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
  water_box(n_mol, seed, basis)
                              closed-shell water cluster at liquid density (random
                              orientations), random shells or a library basis set.
  h_cluster_sp(n_atoms, seed) random H cluster with floating 4-primitive s and p shells.
"""
import math

import numpy as np

from .basis import (CompositeOrb, PrimitiveOrb, genGaussTypeOrb, genGaussTypeOrbSeq, get3DimPGTOrbNormFactor,
                    subshellXYZs)
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


def _random_shells(rng, elem, center):
    """Element-dependent random shells of random_cluster(with_d=True)."""
    bs = []
    for sh in {"H": ["s"], "C": ["s", "p"], "O": ["s", "p", "d"]}[elem]:
        if sh == "d":
            xp = np.exp(rng.uniform(math.log(0.2), math.log(5.0), size=2))
            co = rng.uniform(0.1, 1.0, size=2)
            bs += _shell(center, 2, xp, co)
        else:
            xp = np.exp(rng.uniform(math.log(0.1), math.log(50.0), size=3))
            co = rng.uniform(0.1, 1.0, size=3)
            bs += _shell(center, 0 if sh == "s" else 1, xp, co)
    return bs


def water_box(n_mol=50, seed=0, density=1.0, doo=5.0, basis=None):
    """Closed-shell water cluster at liquid density: rigid H2O (O-H 1.809 bohr,
    H-O-H 104.5 deg) in random orientations, oxygens at least `doo` bohr apart
    in a cube sized for `density` g/cm^3.  With basis=None every atom carries
    the element-dependent random shells of random_cluster(with_d=True) (O: s,
    p, d; H: s), so 50 molecules give 150 atoms and 600 functions; with a
    library name (e.g. "cc-pVDZ": 25 Cartesian functions per molecule, 24
    molecules give 600) the standard basis set."""
    rng = np.random.Generator(np.random.PCG64(seed))
    # 18.015 g/mol, 1 bohr = 0.529177 A
    vol_a3 = n_mol * 18.015 / (density * 0.6022140857)
    box = vol_a3 ** (1.0 / 3.0) / 0.529177210903
    oxy = _place_atoms(rng, n_mol, box, doo)
    roh, half = 1.809, math.radians(104.5) / 2.0
    elems, coords = [], []
    for o in oxy:
        # random rotation from a normalized quaternion
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        Rm = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                       [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                       [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        h1 = Rm @ np.array([roh * math.sin(half), 0.0, roh * math.cos(half)])
        h2 = Rm @ np.array([-roh * math.sin(half), 0.0, roh * math.cos(half)])
        o = np.asarray(o)
        for e, c in (("O", o), ("H", o + h1), ("H", o + h2)):
            elems.append(e)
            coords.append(tuple(float(v) for v in c))
    nuc = NuclearCluster(elems, coords, pairwiseSort=False)
    bs = []
    for e, c in zip(elems, coords):
        bs += _random_shells(rng, e, c) if basis is None else genGaussTypeOrbSeq(c, e, basis)
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
