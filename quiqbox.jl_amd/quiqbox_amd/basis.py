"""Gaussian-type orbital construction and the flattening handed to the C-ABI.

Mirrors Quiqbox.jl's orbital constructors:
  genGaussTypeOrb       OrbitalBases.jl:199-247
  genGaussTypeOrbSeq    OrbitalBases.jl:502-593 (Gaussian-format basis text)
  get3DimPGTOrbNormFactor  OrbitalBases.jl:476-485
  PrimitiveOrb / CompositeOrb (renormalize flags)
and the pointer/weight flattening of MultiOrbitalData (OrbitalBases.jl:439-468)
with the normalization of buildOrbCoreWeight! (Framework.jl:748-773).

The flattened basis (`FlatBasis`) is exactly the argument list of
qbx_basis_create: deduplicated primitives (centre, exponent, Cartesian
exponents) and, per basis function, the primitive indices with final weights.
"""
import json
import math
import os
import re
from dataclasses import dataclass

import numpy as np

from .particles import ELEMENTS, getCharge

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "basis_sets.json")
_BASIS_TEXTS = None

# Cartesian component order of a shell (Lexicons.jl:17-36)


def subshellXYZs(L):
    return [(i, j, L - i - j) for i in range(L, -1, -1) for j in range(L - i, -1, -1)]


AngularSubShellDict = {"S": 0, "P": 1, "D": 2, "F": 3, "G": 4, "H": 5, "I": 6}


def _odd_fact_scaled(a, coeff):
    r = 1.0
    for i in range(1, a + 1, 2):
        r *= i * coeff
    return r


def _axial_factor(p, j):
    """(2j-1)!!/(2p)^j sqrt(pi/p)  (GaussianOrbitals.jl:94-97)"""
    f = _odd_fact_scaled(2 * j - 1, 1.0 / (2.0 * p)) if j > 0 else 1.0
    return math.sqrt(math.pi) / math.sqrt(p) * f


def get3DimPGTOrbNormFactor(xpn, carteAng):
    """Normalization constant of x^i y^j z^k exp(-xpn r^2) (OrbitalBases.jl:476-485)."""
    i, j, k = (int(v) for v in carteAng)
    if min(i, j, k) < 0:
        raise ValueError("angular momentum components must be non-negative")
    angSum = i + j + k
    xpnPart = xpn ** ((2 * angSum + 3) * 0.25)
    f = math.factorial
    angPart = math.pi ** -0.75 * 2.0 ** (1.5 * angSum + 0.75) * math.sqrt(
        f(i) * f(j) * f(k) / (f(2 * i) * f(2 * j) * f(2 * k)))
    return xpnPart * angPart


class OrbitalBasis:
    """Common base of primitive and composite orbitals."""


class PrimitiveOrb(OrbitalBasis):
    """Primitive Cartesian GTO x^i y^j z^k exp(-xpn |r - center|^2)."""

    def __init__(self, center, xpn, ang=(0, 0, 0), renormalize=False):
        self.center = tuple(float(c) for c in center)
        if len(self.center) != 3:
            raise ValueError("only 3-dimensional orbitals are supported by the GPU engine")
        self.xpn = float(xpn)
        if not self.xpn > 0:
            raise ValueError("the exponent must be positive")
        self.ang = tuple(int(a) for a in ang)
        if len(self.ang) != 3 or min(self.ang) < 0:
            raise ValueError("ang must be three non-negative integers")
        self.renormalize = bool(renormalize)

    def key(self):
        return (self.center, self.xpn, self.ang)

    def __repr__(self):
        return "PrimitiveOrb(center=%s, xpn=%g, ang=%s%s)" % (
            self.center, self.xpn, self.ang, ", renormalize=True" if self.renormalize else "")

    def __call__(self, r):
        d = np.asarray(r, dtype=float) - np.asarray(self.center)
        v = math.exp(-self.xpn * float(np.dot(d, d)))
        for t in range(3):
            v *= d[t] ** self.ang[t]
        return v * (1.0 / math.sqrt(prim_overlap(self, self)) if self.renormalize else 1.0)


class CompositeOrb(OrbitalBasis):
    """Linear combination of orbitals (OrbitalBases.jl:57-72)."""

    def __init__(self, basis, weights, renormalize=False):
        basis = list(basis)
        weights = [float(w) for w in weights]
        if len(basis) == 0:
            raise ValueError("a composite orbital needs at least one component")
        if len(basis) != len(weights):
            raise ValueError("basis and weights must have the same length")
        self.basis = basis
        self.weights = weights
        self.renormalize = bool(renormalize)

    def __repr__(self):
        return "CompositeOrb(%d terms%s)" % (len(self.basis), ", renormalize=True" if self.renormalize else "")

    def __call__(self, r):
        w = primitive_expansion(self)
        return sum(c * PrimitiveOrb(p.center, p.xpn, p.ang)(r) for p, c in w)


def genGaussTypeOrb(center, xpn, cons=None, ijk=None, *, renormalize=False,
                    innerRenormalize=False, outerRenormalize=False):
    """genGaussTypeOrb(center, xpn, ijk=(0,0,0); renormalize) -> PrimitiveOrb
    genGaussTypeOrb(center, xpns, cons, ijk=(0,0,0); innerRenormalize, outerRenormalize)
    -> CompositeOrb (OrbitalBases.jl:199-247)."""
    if np.ndim(xpn) == 0:
        ang = (0, 0, 0) if cons is None else cons
        if ijk is not None:
            raise TypeError("a primitive orbital takes (center, xpn, ijk)")
        return PrimitiveOrb(center, xpn, ang, renormalize=renormalize)
    xpns = [float(x) for x in xpn]
    if cons is None:
        raise TypeError("contracted orbitals need contraction coefficients")
    cons = [float(c) for c in cons]
    if len(xpns) == 0:
        raise ValueError("`xpns` must not be empty")
    if len(xpns) != len(cons):
        raise ValueError("the length of `xpns` must equal the output length of `cons`")
    ijk = (0, 0, 0) if ijk is None else ijk
    prims = [PrimitiveOrb(center, x, ijk, renormalize=innerRenormalize) for x in xpns]
    return CompositeOrb(prims, cons, renormalize=outerRenormalize)


def _basis_texts():
    global _BASIS_TEXTS
    if _BASIS_TEXTS is None:
        with open(_DATA) as f:
            _BASIS_TEXTS = json.load(f)
    return _BASIS_TEXTS


AtomicGTOrbSetNames = ["STO-2G", "STO-3G", "STO-6G", "3-21G", "6-31G", "cc-pVDZ", "cc-pVTZ", "cc-pVQZ"]


def _parse_number(tok):
    return float(tok.replace("D", "E").replace("d", "e"))


def genGaussTypeOrbSeq(center, content, basisKey=None, *, innerRenormalize=False,
                       outerRenormalize=False, unlinkCenter=False):
    """Basis functions of one atom from Gaussian-format text, or from an element
    symbol and basis-set name (OrbitalBases.jl:502-593).  SP shells split into
    an s and a p shell; each Cartesian component's coefficients carry the
    primitive normalization factor get3DimPGTOrbNormFactor(xpn, ijk)."""
    if basisKey is not None:
        atm = str(content)
        texts = _basis_texts().get(basisKey)
        Z = ELEMENTS.get(atm)
        if texts is None or Z is None or Z > len(texts) or texts[Z - 1] is None:
            raise ValueError("Quiqbox does not have this basis-set configuration pre-stored: (%s, %s)"
                             % (atm, basisKey))
        content = texts[Z - 1]
    lines = [l.split() for l in content.strip("\n").splitlines()]
    bfs = []
    for n, toks in enumerate(lines):
        if len(toks) <= 2 or toks[0] == "X":
            continue
        try:
            _parse_number(toks[0])
            continue  # a numeric line
        except ValueError:
            pass
        shell = toks[0]
        nprim = int(toks[1])
        rows = [[_parse_number(t) for t in lines[n + 1 + k]] for k in range(nprim)]
        xpns = [r[0] for r in rows]
        angNums = (0, 1) if shell == "SP" else (AngularSubShellDict[shell],)
        for col, L in enumerate(angNums):
            for ijk in subshellXYZs(L):
                cons = [r[1 + col] * get3DimPGTOrbNormFactor(x, ijk) for r, x in zip(rows, xpns)]
                bfs.append(genGaussTypeOrb(center, xpns, cons, ijk, innerRenormalize=innerRenormalize,
                                           outerRenormalize=outerRenormalize))
    return bfs


# ---------------------------------------------------------------------------
# primitive overlap (used for the normalization weights only)
# ---------------------------------------------------------------------------

def _ov1d(PA, PB, inv2p, ni, nj):
    S = np.zeros((ni + 1, nj + 1))
    S[0, 0] = 1.0
    for i in range(ni + 1):
        if i > 0:
            S[i, 0] = PA * S[i - 1, 0] + (inv2p * (i - 1) * S[i - 2, 0] if i > 1 else 0.0)
        for j in range(1, nj + 1):
            v = PB * S[i, j - 1]
            if i > 0:
                v += inv2p * i * S[i - 1, j - 1]
            if j > 1:
                v += inv2p * (j - 1) * S[i, j - 2]
            S[i, j] = v
    return S


def prim_overlap(p1, p2):
    """<p1|p2> of two unnormalized primitives (Obara-Saika 1D recurrence)."""
    if p1.key() == p2.key():
        r = 1.0
        for l in p1.ang:
            r *= _axial_factor(2.0 * p1.xpn, l)
        return r
    a, b = p1.xpn, p2.xpn
    p = a + b
    mu = a * b / p
    r = 1.0
    for t in range(3):
        A, B = p1.center[t], p2.center[t]
        P = (a * A + b * B) / p
        S = _ov1d(P - A, P - B, 0.5 / p, p1.ang[t], p2.ang[t])
        r *= S[p1.ang[t], p2.ang[t]] * math.sqrt(math.pi / p) * math.exp(-mu * (A - B) ** 2)
    return r


def primitive_expansion(orb):
    """Final (primitive, weight) list of an orbital with the normalization of
    buildOrbCoreWeight! (Framework.jl:748-773): a renormalized primitive gets
    1/sqrt(S_aa), a renormalized composite divides its weights by
    sqrt(w^T S w).  Nested composites are expanded recursively."""
    if isinstance(orb, PrimitiveOrb):
        w = 1.0 / math.sqrt(prim_overlap(orb, orb)) if orb.renormalize else 1.0
        return [(PrimitiveOrb(orb.center, orb.xpn, orb.ang), w)]
    if not isinstance(orb, CompositeOrb):
        raise TypeError("unsupported orbital type %r" % type(orb))
    terms = []
    for sub, c in zip(orb.basis, orb.weights):
        for p, w in primitive_expansion(sub):
            terms.append((p, c * w))
    if orb.renormalize:
        s = 0.0
        for i, (pi, wi) in enumerate(terms):
            for j, (pj, wj) in enumerate(terms):
                s += wi * wj * prim_overlap(pi, pj)
        f = 1.0 / math.sqrt(abs(s))
        terms = [(p, w * f) for p, w in terms]
    return terms


@dataclass
class FlatBasis:
    """Arguments of qbx_basis_create (0-based, C order)."""
    prim_cen: np.ndarray   # (n_prim, 3) float64
    prim_xpn: np.ndarray   # (n_prim,)   float64
    prim_ang: np.ndarray   # (n_prim, 3) int32
    bf_ptr: np.ndarray     # (n_bf+1,)   int32
    bf_prim: np.ndarray    # (nnz,)      int32
    bf_w: np.ndarray       # (nnz,)      float64

    @property
    def n_bf(self):
        return len(self.bf_ptr) - 1

    @property
    def n_prim(self):
        return len(self.prim_xpn)

    def as_dict(self):
        return dict(prim_cen=self.prim_cen, prim_xpn=self.prim_xpn, prim_ang=self.prim_ang,
                    bf_ptr=self.bf_ptr, bf_prim=self.bf_prim, bf_w=self.bf_w)

    def is_standard(self):
        """Every basis function is a single-centre, single-(i,j,k) contraction."""
        for f in range(self.n_bf):
            idx = self.bf_prim[self.bf_ptr[f]:self.bf_ptr[f + 1]]
            if len(idx) == 0:
                return False
            c = self.prim_cen[idx]
            a = self.prim_ang[idx]
            if not (np.all(c == c[0]) and np.all(a == a[0])):
                return False
        return True


def flatten_basis(orbs):
    """MultiOrbitalData-equivalent flattening (OrbitalBases.jl:439-468):
    deduplicated primitive fields plus per-function weights."""
    orbs = list(orbs)
    if len(orbs) == 0:
        raise ValueError("`orbitals` must not be empty")
    prims = {}
    cen, xpn, ang = [], [], []
    ptr = [0]
    pidx, wts = [], []
    for orb in orbs:
        terms = primitive_expansion(orb)
        # merge repeated primitives inside one function
        acc = {}
        order = []
        for p, w in terms:
            k = p.key()
            if k not in acc:
                acc[k] = 0.0
                order.append(k)
            acc[k] += w
        for k in order:
            if k not in prims:
                prims[k] = len(xpn)
                cen.append(k[0])
                xpn.append(k[1])
                ang.append(k[2])
            pidx.append(prims[k])
            wts.append(acc[k])
        ptr.append(len(pidx))
    return FlatBasis(np.array(cen, dtype=np.float64).reshape(-1, 3), np.array(xpn, dtype=np.float64),
                     np.array(ang, dtype=np.int32).reshape(-1, 3), np.array(ptr, dtype=np.int32),
                     np.array(pidx, dtype=np.int32), np.array(wts, dtype=np.float64))


def expand_to_standard(flat):
    """For bases whose functions mix centres or Cartesian exponents: an
    auxiliary basis with one function per primitive and the matrix T
    (n_aux x n_bf) such that phi_f = sum_a T[a, f] chi_a."""
    n = flat.n_bf
    T = np.zeros((flat.n_prim, n))
    for f in range(n):
        for s in range(flat.bf_ptr[f], flat.bf_ptr[f + 1]):
            T[flat.bf_prim[s], f] += flat.bf_w[s]
    aux = FlatBasis(flat.prim_cen.copy(), flat.prim_xpn.copy(), flat.prim_ang.copy(),
                    np.arange(flat.n_prim + 1, dtype=np.int32), np.arange(flat.n_prim, dtype=np.int32),
                    np.ones(flat.n_prim))
    return aux, T


def splitOrb(orb):
    """Components of a composite orbital (or the primitive itself)."""
    return list(orb.basis) if isinstance(orb, CompositeOrb) else [orb]
