"""Hartree-Fock SCF driver with Quiqbox's call surface (src/HartreeFock.jl).

The control flow -- initial guesses, the DD/DIIS/ADIIS/EDIIS method sequence,
the convex-combination solvers, the three convergence criteria, oscillation
detection, early termination with rollback -- follows HartreeFock.jl step by
step (file:line at each function).  The two-electron part of every Fock
matrix comes from a `TwoBodyOperator`: by default the GPU direct J/K build of
libqbx_hip.so (`DirectJK`); a dense tensor (`DenseERI`) can be given instead.

Linear algebra (S^{-1/2}, Fock diagonalization, DIIS algebra) runs in FP64.
"""
import math
import sys
import time
from dataclasses import dataclass, field

import numpy as np
import scipy.optimize

from . import _lib
from .basis import FlatBasis
from .engine import GPUBasis
from .particles import NuclearCluster, OccupationState, nucRepulsion, prepareSpinConfiguration

# module constants (HartreeFock.jl:13-29)
defaultDS = 0.75
defaultDIISsize = 15
defaultDIISsolver = "LBFGS"
SADHFmaxStep = 50
defaultHFinfoLevel = 2
defaultHFmaxStep = 200
defaultHFsaveTrace = (False, False, False, True)
HFminItr = 10
HFinterEsOscConvSize = 15
HFinterMatStoreSizes = (2, 3, 2)
defaultSCFconfigArgs = (("DD", "ADIIS", "DIIS"), (5e-3, 1e-4, 1e-9))
defaultSecConvRatio = (1000.0, 1000.0)
defaultOscThreshold = 5e-6
ATOL = 4e-16            # getAtolVal(Float64)
ATOL_DIGITS = 15        # getAtolDigits(Float64)


class RCHartreeFock:
    name = "Restricted closed-shell (RHF)"

    def __eq__(self, o):
        return isinstance(o, RCHartreeFock)

    def __hash__(self):
        return 1


class UOHartreeFock:
    name = "Unrestricted open-shell (UHF)"

    def __eq__(self, o):
        return isinstance(o, UOHartreeFock)

    def __hash__(self):
        return 2


# ---------------------------------------------------------------------------
# two-body operators (the Fock-build hook, HartreeFock.jl:143-151, 305-335)
# ---------------------------------------------------------------------------

class TwoBodyOperator:
    def getGcore(self, DJ, DK):
        raise NotImplementedError

    def getGcores(self, pairs):
        return [self.getGcore(dj, dk) for dj, dk in pairs]


class DenseERI(TwoBodyOperator):
    """Dense (ij|kl) tensor; getGcore as HartreeFock.jl:305-319."""

    def __init__(self, eri):
        self.eri = np.asarray(eri, dtype=np.float64)

    def getGcore(self, DJ, DK):
        J = np.einsum("ijkl,lk->ij", self.eri, DJ, optimize=True)
        K = np.einsum("ikjl,kj->il", self.eri, DK, optimize=True)
        return J - K


class DirectJK(TwoBodyOperator):
    """Direct, Schwarz-screened J/K on the GPU (qbx_jk).

    precision: QBX_PREC_FP64, or QBX_PREC_MIXED (FP32 arithmetic for quartets
        whose Schwarz bound is below the context's FP64 cut).  A mixed operator
        switches itself to FP64 for good once the SCF reports |dE| below
        `switch_dE` (the final FP64 Fock rebuilds near convergence).
    incremental: build G(D) = G(D_ref) + G(D - D_ref) from the last computed
        pair, with density screening at `density_threshold` (quartets with
        Q_ab Q_cd max|D - D_ref| below it are skipped; default: the Schwarz
        `threshold`); a full build every `rebuild_every` builds and at the
        precision switch.  A looser density threshold trades accuracy for
        time: on 24 H2O in cc-pVDZ (learned guess, FP64 incremental builds,
        purified densities, |dE| < 1e-8 Eh) 1e-12 took 0.740 s, 1e-11 0.695 s
        and 3e-11 0.670 s, all 4.6-4.7e-9 Eh from the tightly converged energy
        (1e-10 needed 14 builds and missed by 3e-8 Eh), but on a 24-atom C/H/O
        cluster converged to 1e-10 Eh, 1e-11 left the energy 2.8e-9 Eh from the
        plain FP64 one.
    """

    def __init__(self, gpu_basis, threshold=1e-12, precision=_lib.QBX_PREC_FP64, incremental=False,
                 rebuild_every=8, switch_dE=1e-5, density_threshold=None):
        self.basis = gpu_basis
        self.threshold = threshold
        self.density_threshold = threshold if density_threshold is None else density_threshold
        self.precision = precision
        self.incremental = incremental
        self.rebuild_every = rebuild_every
        self.switch_dE = switch_dE
        self.n_builds = 0
        self.n_mixed = 0
        self.builds = []      # (full build?, precision, seconds) of every build
        self._ref = None      # (DJ, DK, G) of the last build
        self._since_full = 0

    def observe(self, dE, dF):
        """Progress report from the SCF loop (|dE| of the last iteration)."""
        if self.precision == _lib.QBX_PREC_MIXED and abs(dE) < self.switch_dE:
            self.precision = _lib.QBX_PREC_FP64
            self._ref = None   # next build is a full FP64 rebuild

    def _build(self, DJ, DK, density_thr, precision):
        ctx = self.basis.ctx
        t0 = time.perf_counter()
        ctx.set_density_screening(density_thr)
        try:
            G = self.basis.jk(DJ, DK, self.threshold, precision)
        finally:
            ctx.set_density_screening(0.0)
        self.builds.append((density_thr == 0.0, precision, time.perf_counter() - t0))
        self.n_builds += 1
        if precision == _lib.QBX_PREC_MIXED:
            self.n_mixed += 1
        return G

    def getGcores(self, pairs):
        DJ = np.stack([p[0] for p in pairs])
        DK = np.stack([p[1] for p in pairs])
        ref = self._ref
        if (self.incremental and ref is not None and ref[0].shape == DJ.shape
                and self._since_full < self.rebuild_every):
            G = ref[2] + self._build(DJ - ref[0], DK - ref[1], self.density_threshold, self.precision)
            self._since_full += 1
        else:
            G = self._build(DJ, DK, 0.0, self.precision)
            self._since_full = 0
        if self.incremental:
            self._ref = (DJ, DK, G)
        return [G[s] for s in range(len(pairs))]

    def getGcore(self, DJ, DK):
        return self.getGcores([(DJ, DK)])[0]


def getG(twoBody, D):
    """RHF: G = J[2D] - K[D]; UHF: G_s = J[Da+Db] - K[D_s] (HartreeFock.jl:322-327)."""
    if len(D) == 1:
        return tuple(twoBody.getGcores([(2.0 * D[0], D[0])]))
    Dt = D[0] + D[1]
    return tuple(twoBody.getGcores([(Dt, D[0]), (Dt, D[1])]))


def getF(Hcore, twoBody, D):
    """F = Hcore + G (HartreeFock.jl:330-335)"""
    return tuple(Hcore + g for g in getG(twoBody, D))


# ---------------------------------------------------------------------------
# elementary steps
# ---------------------------------------------------------------------------

def getOrthonormalization(S):
    """X = inv(sqrt(Hermitian(S))), symmetric orthonormalization (HartreeFock.jl:39-42)."""
    w, V = np.linalg.eigh(S)
    if w.min() <= 0:
        raise ValueError("the overlap matrix is not positive definite (smallest eigenvalue %g)" % w.min())
    return (V / np.sqrt(w)) @ V.T


def solveFockMatrix(X, F, stabilizeSign=True):
    """eigen(X' F X) ascending; C = X C'; column signs fixed so that C[0, j] >= 0
    (HartreeFock.jl:46-56)."""
    Fp = X.T @ F @ X
    e, Cp = np.linalg.eigh(0.5 * (Fp + Fp.T))
    C = X @ Cp
    if stabilizeSign:
        C = C * np.where(C[0, :] < 0, -1.0, 1.0)[None, :]
    return C, e


def getC(X, F, stabilizeSign=True):
    return solveFockMatrix(X, F, stabilizeSign)[0]


def getD(C, N):
    """D = C[:, :N] C[:, :N]' (HartreeFock.jl:296-299)"""
    c = C[:, :N]
    return c @ c.T


def getE(Hcore, F, D):
    """E = <D, Hcore + F> / 2 (HartreeFock.jl:339-342)"""
    return float(np.vdot(D, Hcore + F)) / 2.0


def get2SpinQuantity(O):
    """RHF (one sector): 2 O; UHF (two sectors): O_a + O_b (HartreeFock.jl:344)."""
    return abs(2 - (len(O) - 1)) * sum(O)


def getEhfCore(Hcore, F, D):
    return get2SpinQuantity([getE(Hcore, f, d) for f, d in zip(F, D)])


def getEhf(Hcore, twoBody, C, Ns):
    """Electronic energy of coefficient matrices (HartreeFock.jl:353-359)."""
    if isinstance(twoBody, np.ndarray):
        twoBody = DenseERI(twoBody)
    D = tuple(getD(c, n) for c, n in zip(C, Ns))
    F = getF(Hcore, twoBody, D)
    return getEhfCore(Hcore, F, D)


def getEhfMO(HcoreMO, HeeIMO, Ns, Jab=None):
    """Electronic energy from MO-basis integrals (HartreeFock.jl:362-389)."""
    if len(Ns) == 1:
        (h,), (g,), (n,) = HcoreMO, HeeIMO, Ns
        t1 = 2.0 * np.trace(h[:n, :n])
        t2 = 0.0
        for i in range(n):
            for j in range(n):
                t2 += 2.0 * g[i, i, j, j] - g[i, j, j, i]
        return t1 + t2
    res = 0.0
    for h, g, n in zip(HcoreMO, HeeIMO, Ns):
        res += np.trace(h[:n, :n])
        for j in range(n):
            for i in range(j):
                res += g[i, i, j, j] - g[i, j, j, i]
    return res + Jab[:Ns[0], :Ns[1]].sum()


class LazyCoeff:
    """Orbital coefficients getC(X, F), computed on first use: the SCF steps
    whose density comes from purification keep the Fock matrix instead."""
    __slots__ = ("X", "F", "_C")

    def __init__(self, X, F):
        self.X, self.F, self._C = X, F, None

    def value(self):
        if self._C is None:
            self._C = getC(self.X, self.F)
        return self._C

    def __array__(self, dtype=None, copy=None):
        C = self.value()
        return C if dtype is None else C.astype(dtype)


def _coeff(c):
    return c.value() if isinstance(c, LazyCoeff) else c


class PurifiedDensity:
    """Density solver for the SCF steps: D = C[:, :N] C[:, :N]^T of eigen(X' F X)
    by trace-correcting purification on the GPU (Context.purified_density),
    with the diagonalization of getC/getD where purification does not settle
    (no gap at the Fermi level)."""

    def __init__(self, ctx, max_iter=100):
        self.ctx = ctx
        self.max_iter = max_iter
        self.calls = 0
        self.fallbacks = 0
        self.products = 0

    def __call__(self, X, F, N):
        self.calls += 1
        try:
            D, it, _ = self.ctx.purified_density(X, F, N, self.max_iter)
        except _lib.QbxError:
            self.fallbacks += 1
            return getD(getC(X, F), N)
        self.products += it
        return 0.5 * (D + D.T)


# Basis size from which the default SCF step runs on the GPU.  One step
# (X' F X, eigen, C = X C', D) measured on the box, device against numpy on
# the host: n = 128 2.44 / 1.18 ms, 256 6.18 / 4.38 ms, 384 8.86 / 8.79 ms,
# 512 12.0 / 23.0 ms, 600 13.6 / 18.0 ms (rocSOLVER's fixed costs dominate
# small problems).
DEVICE_EIGEN_MIN_N = 384


class DeviceEigen:
    """getC + getD of the SCF steps on the GPU (Context.fock_solve): X' F X
    and X C' on the FP64 matrix cores, the symmetric eigenproblem by rocSOLVER;
    the same matrices as solveFockMatrix/getD up to rounding."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.calls = 0

    def __call__(self, X, F, N):
        self.calls += 1
        C, _, D = self.ctx.fock_solve(X, F, N)
        return C, D


def getCDFE(Hcore, twoBody, X, Ns, F):
    """One SCF step: F -> C -> D -> F -> E (HartreeFock.jl:392-403).  With a
    density solver on the two-body operator (runHartreeFock's
    densitySolver="purify") D comes from it and C is formed only if used;
    with an eigen solver ("eigen", the default with the GPU engine) C and D come
    from the device."""
    solver = getattr(twoBody, "density_solver", None)
    eig = getattr(twoBody, "eigen_solver", None)
    if solver is None and eig is not None:
        CD = [eig(X, f, n) for f, n in zip(F, Ns)]
        Cn = tuple(c for c, _ in CD)
        Dn = tuple(d for _, d in CD)
    elif solver is None:
        Cn = tuple(getC(X, f) for f in F)
        Dn = tuple(getD(c, n) for c, n in zip(Cn, Ns))
    else:
        Cn = tuple(LazyCoeff(X, f) for f in F)
        Dn = tuple(solver(X, f, n) for f, n in zip(F, Ns))
    Fn = getF(Hcore, twoBody, Dn)
    En = tuple(getE(Hcore, f, d) for f, d in zip(Fn, Dn))
    Dt = get2SpinQuantity(Dn)
    Et = get2SpinQuantity(En)
    return [(c, d, f, e, Dt, Et) for c, d, f, e in zip(Cn, Dn, Fn, En)]


def rmsOf(a):
    a = np.asarray(a)
    return float(np.linalg.norm(a) / math.sqrt(a.size))


def getEresidual(F, D, S):
    """FDS - SDF (HartreeFock.jl:1298-1299).  F, D and S are symmetric, so
    SDF = (FDS)^T: two products instead of four."""
    M = F @ D @ S
    return M - M.T


# ---------------------------------------------------------------------------
# SCF state (HartreeFock.jl:406-508)
# ---------------------------------------------------------------------------

class HFinterrelatedVars:
    def __init__(self, Dtots=None, Etots=None):
        self.Dtots = [] if Dtots is None else Dtots
        self.Etots = [] if Etots is None else Etots


class HFtempInfo:
    def __init__(self, N, C, D, F, E, shared):
        self.N = N
        self.Cs = [C]
        self.Ds = [D]
        self.Fs = [F]
        self.Es = [E]
        self.shared = shared


def _push(lst, val, maxLen, truncate):
    while truncate and len(lst) >= maxLen:
        lst.pop(0)
    lst.append(val)


def updateHFtempInfo(maxLen, vars_, ress):
    for tv, res in zip(vars_, ress):
        for lst, r, tr in zip((tv.Cs, tv.Ds, tv.Fs, tv.Es), res[:4], (True, True, True, False)):
            _push(lst, r, maxLen, tr)
    sh = vars_[0].shared
    _push(sh.Dtots, ress[0][4], maxLen, True)
    _push(sh.Etots, ress[0][5], maxLen, False)


def popHFtempInfo(vars_, counts=1):
    for tv in vars_:
        for lst in (tv.Cs, tv.Ds, tv.Fs, tv.Es):
            for _ in range(counts):
                lst.pop()
    for lst in (vars_[0].shared.Dtots, vars_[0].shared.Etots):
        for _ in range(counts):
            lst.pop()


def clearHFtempInfo(saveTrace, vars_):
    for tv in vars_:
        for bl, name in zip(saveTrace, ("Cs", "Ds", "Fs", "Es")):
            if not bl:
                setattr(tv, name, getattr(tv, name)[-1:])
    sh = vars_[0].shared
    if not saveTrace[1]:
        sh.Dtots = sh.Dtots[-1:]
    if not saveTrace[3]:
        sh.Etots = sh.Etots[-1:]


def formatSpinConfiguration(HFT, spin):
    a, b = spin[0], spin[1]
    if isinstance(HFT, RCHartreeFock):
        if a != b:
            raise AssertionError("Close-shell system must have the same number of spins for both spin orientations.")
        return (a,)
    return (a, b)


@dataclass
class ElecHamiltonianConfig:
    spin: OccupationState
    orthonormalization: np.ndarray
    overlap: np.ndarray
    oneBody: np.ndarray
    twoBody: TwoBodyOperator


def initializeSCF(HFT, config, C):
    """HartreeFock.jl:526-542"""
    Ns = formatSpinConfiguration(HFT, config.spin)
    D = tuple(getD(c, n) for c, n in zip(C, Ns))
    F = getF(config.oneBody, config.twoBody, D)
    E = tuple(getE(config.oneBody, f, d) for f, d in zip(F, D))
    shared = HFinterrelatedVars([get2SpinQuantity(D)], [get2SpinQuantity(E)])
    return tuple(HFtempInfo(n, c, d, f, e, shared) for n, c, d, f, e in zip(Ns, C, D, F, E))


# ---------------------------------------------------------------------------
# configuration objects (HartreeFock.jl:545-705, 822-904)
# ---------------------------------------------------------------------------

class SCFconfig:
    """SCFconfig(methods, intervals, config={}; secondaryConvRatio, oscillateThreshold)
    or SCFconfig(threshold=...) (HartreeFock.jl:659-688)."""

    def __init__(self, methods=None, intervals=None, config=None, *, threshold=None,
                 secondaryConvRatio=defaultSecConvRatio, oscillateThreshold=defaultOscThreshold):
        if methods is None:
            methods = defaultSCFconfigArgs[0]
            iv = list(defaultSCFconfigArgs[1])
            if threshold is not None:
                iv[-1] = float(threshold)
            intervals = tuple(iv)
        methods = tuple(str(m).lstrip(":") for m in methods)
        intervals = tuple(float(x) for x in intervals)
        if len(methods) == 0 or len(methods) != len(intervals):
            raise ValueError("methods and intervals must be non-empty and of equal length")
        for m in methods:
            if m not in ("DD", "DIIS", "ADIIS", "EDIIS"):
                raise ValueError("unknown SCF method %r" % m)
        for x in intervals:
            if not (x >= 0 or math.isnan(x)):
                raise ValueError("thresholds must be non-negative")
        kw = [dict() for _ in methods]
        for i, pairs in (config or {}).items():
            kw[i - 1] = dict((str(k).lstrip(":"), v) for k, v in (pairs.items() if isinstance(pairs, dict) else pairs))
        if np.ndim(secondaryConvRatio) == 0:
            secondaryConvRatio = (secondaryConvRatio, secondaryConvRatio)
        self.method = methods
        self.interval = intervals
        self.methodConfig = tuple(kw)
        self.secondaryConvRatio = (float(secondaryConvRatio[0]), float(secondaryConvRatio[-1]))
        self.oscillateThreshold = float(oscillateThreshold)

    def cache_sizes(self):
        """getSCFcacheSizes (HartreeFock.jl:690-705)"""
        res = []
        for m, kw in zip(self.method, self.methodConfig):
            if m == "DD":
                res.append(1)
            else:
                s = int(kw.get("DIISsize", defaultDIISsize))
                if s <= 0:
                    raise ValueError("DIISsize must be positive")
                res.append(s)
        return res


class OrbCoeffInitialConfig:
    """Initial-guess method (:Direct, :CoreH, :GWH, :SAD) and data (HartreeFock.jl:157-173)."""

    def __init__(self, HFT, method_or_data="SAD"):
        self.HFT = HFT
        if isinstance(method_or_data, str):
            m = method_or_data.lstrip(":")
            if m not in ("CoreH", "GWH", "SAD", "Direct"):
                raise ValueError("unknown initial-guess method %r" % method_or_data)
            self.method = m
            self.data = None
        else:
            data = tuple(np.asarray(c, dtype=np.float64) for c in method_or_data)
            need = 1 if isinstance(HFT, RCHartreeFock) else 2
            if len(data) != need:
                raise ValueError("%d coefficient matrices expected" % need)
            self.method = "Direct"
            self.data = data


class HFconfig:
    """HFconfig(type=RCHartreeFock(); initial=:SAD, strategy=SCFconfig(), maxStep=200,
    earlyStop=true, saveTrace) (HartreeFock.jl:874-904)."""

    def __init__(self, type=None, *, initial="SAD", strategy=None, orthonormalization="Symmetric",
                 maxStep=defaultHFmaxStep, earlyStop=True, saveTrace=defaultHFsaveTrace):
        self.type = type if type is not None else RCHartreeFock()
        if not isinstance(initial, OrbCoeffInitialConfig):
            initial = OrbCoeffInitialConfig(self.type, initial)
        self.initial = initial
        self.strategy = strategy if strategy is not None else SCFconfig()
        if orthonormalization != "Symmetric":
            raise ValueError("only symmetric orthonormalization is available")
        self.orthonormalization = orthonormalization
        if maxStep <= 0:
            raise ValueError("maxStep must be positive")
        self.maxStep = int(maxStep)
        self.earlyStop = bool(earlyStop)
        self.saveTrace = tuple(bool(x) for x in saveTrace)


# ---------------------------------------------------------------------------
# initial guesses (HartreeFock.jl:71-140, 178-293)
# ---------------------------------------------------------------------------

def breakCoeffSymmetryCore(X, C1, C2):
    """HartreeFock.jl:71-100"""
    if C1.shape != C2.shape:
        raise ValueError("DimensionMismatch: coefficient matrices differ in size")
    inv = np.linalg.inv(0.5 * (X + X.T))
    c1 = inv @ C1
    c2 = inv @ C2
    for k in range(c1.shape[1]):
        col = c1[:, k] if (k + 1) % 2 == 1 else c2[:, k]
        idx = int(np.argmax(np.abs(col)))
        mag = abs(col[idx])
        val = col[idx]
        if abs(mag - 1.0) < math.sqrt(np.finfo(float).eps):
            col[idx] *= -val
        else:
            col[idx] = 0.0
            nrm = np.linalg.norm(col)
            if nrm > 0:
                col /= nrm
    return (X @ c1, X @ c2)


def breakCoeffSymmetry(HFT, X, C):
    """HartreeFock.jl:102-124"""
    if isinstance(HFT, RCHartreeFock):
        return tuple(C)
    if len(C) == 1:
        return breakCoeffSymmetryCore(X, C[0], C[0])
    inv = np.linalg.inv(0.5 * (X + X.T))
    diff = rmsOf(inv @ (C[0] - C[1]))
    if diff < 0.1:
        return breakCoeffSymmetryCore(X, C[0], C[1])
    return tuple(C)


def breakCoeffSymmetryFromDensity(HFT, X, Hcore, twoBody, Dpair):
    """HartreeFock.jl:126-140"""
    if isinstance(HFT, RCHartreeFock):
        D1 = (Dpair[0] + Dpair[1]) / 2.0
        return tuple(getC(X, f) for f in getF(Hcore, twoBody, (D1,)))
    C = tuple(getC(X, f) for f in getF(Hcore, twoBody, tuple(Dpair)))
    return breakCoeffSymmetry(UOHartreeFock(), X, C)


def guessCoreH(HFT, X, Hcore):
    """HartreeFock.jl:221-226"""
    return breakCoeffSymmetry(HFT, X, (getC(X, Hcore),))


def guessGWH(HFT, X, S, Hcore):
    """Generalized Wolfsberg-Helmholz: H_ij = 3 S_ij (H_ii + H_jj) / 8 (HartreeFock.jl:235-254)."""
    d = np.diag(Hcore)
    H = 3.0 * S * (d[:, None] + d[None, :]) / 8.0
    return breakCoeffSymmetry(HFT, X, (getC(X, H),))


def guessSAD(HFT, nuc, config, atom_core_h):
    """Quiqbox's superposition of atomic densities (HartreeFock.jl:256-293): for
    each nucleus, a UHF ADIIS run in the full basis with that nucleus's
    attraction only and the whole system's electron count; densities averaged."""
    atmCfg = SCFconfig(("ADIIS",), (max(1e-2, 10 * ATOL),))
    Da = np.zeros_like(config.overlap)
    Db = np.zeros_like(config.overlap)
    nAtm = len(nuc)
    for k in range(nAtm):
        Hk = atom_core_h(k)
        sysk = ElecHamiltonianConfig(config.spin, config.orthonormalization, config.overlap, Hk, config.twoBody)
        # the reference passes the overlap matrix in the orthonormalization slot here (HartreeFock.jl:284)
        C0 = guessCoreH(UOHartreeFock(), config.overlap, Hk)
        vars_, _ = runHartreeFockCore(UOHartreeFock(), atmCfg, sysk, C0, SADHFmaxStep, True,
                                      (False, False, False, False))
        Da += vars_[0].Ds[-1]
        Db += vars_[1].Ds[-1]
    return breakCoeffSymmetryFromDensity(HFT, config.orthonormalization, config.oneBody, config.twoBody,
                                         (Da / nAtm, Db / nAtm))


# ---------------------------------------------------------------------------
# DIIS family (HartreeFock.jl:1245-1485)
# ---------------------------------------------------------------------------

def shiftLastEle(c, shiftVal):
    """Tools.jl:561-573 (on a copy)"""
    s = float(np.sum(c))
    shiftVal = abs(shiftVal)
    if s < shiftVal:
        sg = -1.0 if s < 0 else 1.0
        c = c.copy()
        c[-1] += sg * shiftVal
        s += sg * shiftVal
    return s, c


def _xdiis_f(v, B, shift):
    def f(c):
        s, c = shiftLastEle(np.asarray(c, dtype=float), shift)
        return float(v @ c / s + c @ B @ c / (2.0 * s * s))
    return f


def _xdiis_g(v, B, shift):
    def g(c):
        s, c = shiftLastEle(np.asarray(c, dtype=float), shift)
        return v / s + (B + B.T) @ c / (2.0 * s * s) - (v @ c / s ** 2 + c @ B @ c / s ** 3)
    return g


_blas_ctl = None


def _one_blas_thread():
    """Context that runs BLAS single-threaded: the DIIS subproblems are m <= 15
    variables, where waking a BLAS thread pool per call costs more than the
    arithmetic (0.8 ms per L-BFGS-B iteration with 16 OpenBLAS threads)."""
    global _blas_ctl
    if _blas_ctl is None:
        from threadpoolctl import ThreadpoolController
        _blas_ctl = ThreadpoolController()
    return _blas_ctl.limit(limits=1, user_api="blas")


def LBFGSBsolver(convex, c, v, B):
    """L-BFGS-B on the normalized quadratic (HartreeFock.jl:1416-1430); scipy's
    L-BFGS-B is the same Fortran algorithm (v3.0) the reference wraps."""
    shift = ATOL
    lb = 0.0 if convex else None
    with _one_blas_thread():
        res = scipy.optimize.minimize(_xdiis_f(v, B, shift), c, jac=_xdiis_g(v, B, shift), method="L-BFGS-B",
                                      bounds=[(lb, None)] * len(c),
                                      options=dict(maxcor=min(ATOL_DIGITS, 50), ftol=1e5 * np.finfo(float).eps,
                                                   gtol=10.0 ** -ATOL_DIGITS, maxfun=10000, maxiter=10000))
    s, x = shiftLastEle(res.x, shift)
    return x / s


def SPGBsolver(convex, c, v, B):
    """Spectral projected gradient with box constraints (Birgin, Martinez & Raydan),
    the method of the reference's SPGBox option (HartreeFock.jl:1432-1443)."""
    shift = ATOL
    f = _xdiis_f(v, B, shift)
    g = _xdiis_g(v, B, shift)
    lo = 0.0 if convex else -np.inf
    proj = lambda x: np.maximum(x, lo)
    x = proj(np.asarray(c, dtype=float))
    gx = g(x)
    lam = 1.0
    fs = [f(x)]
    for _ in range(10000):
        d = proj(x - lam * gx) - x
        if np.max(np.abs(proj(x - gx) - x)) <= 10.0 ** -ATOL_DIGITS:
            break
        fmax = max(fs[-10:])
        t = 1.0
        fx = fs[-1]
        while True:
            xn = x + t * d
            fn = f(xn)
            if fn <= fmax + 1e-4 * t * (gx @ d) or t < 1e-20:
                break
            t *= 0.5
        s = xn - x
        gn = g(xn)
        y = gn - gx
        sy = s @ y
        lam = min(1e30, max(1e-30, (s @ s) / sy)) if sy > 0 else 1e30
        x, gx = xn, gn
        fs.append(fn)
    s, x = shiftLastEle(x, shift)
    return x / s


def CMsolver(convex, c, v, B, perturbation=1000 * ATOL):
    """Lagrange-multiplier solve with non-negativity by subset elimination
    (HartreeFock.jl:1445-1477)."""
    n = len(v)
    B = np.array(B, dtype=float)
    b = np.concatenate([-v, [1.0]])

    def getA(M):
        A = np.zeros((n + 1, n + 1))
        A[:n, :n] = M
        A[:n, n] = 1.0
        A[n, :n] = 1.0
        return A

    import itertools
    while True:
        A = getA(B)
        while np.linalg.det(A) == 0:
            B = B + perturbation * np.eye(n)
            A = getA(B)
        x = np.linalg.solve(A, b)[:n]
        if not (convex and np.any(x < 0)):
            return x
        order = list(np.argsort(np.abs(x), kind="stable"))
        for r in range(1, n + 1):
            for is_ in itertools.combinations(order, r):
                keep = [i for i in range(n + 1) if i not in is_]
                At = A[np.ix_(keep, keep)]
                if np.linalg.det(At) == 0:
                    continue
                cl = list(np.linalg.solve(At, b[keep])[:-1])
                for i in sorted(is_):
                    cl.insert(i, 0.0)
                x = np.array(cl)
                if not np.any(x < 0):
                    return x
        B = B + perturbation * np.eye(n)


ConstraintSolvers = dict(LCM=CMsolver, LBFGS=LBFGSBsolver, SPGB=SPGBsolver)


class _ResidualCache:
    """X'(FDS-SDF)X per stored (D, F), so the DIIS B matrix needs no recomputation."""

    def __init__(self, S, X):
        self.S = S
        self.X = X
        self.store = {}

    def get(self, D, F):
        k = (id(D), id(F))
        r = self.store.get(k)
        if r is None:
            r = self.X.T @ getEresidual(F, D, self.S) @ self.X
            self.store[k] = (r, D, F)
            return r
        return r[0]

    def prune(self, live):
        keep = {(id(d), id(f)) for d, f in live}
        self.store = {k: v for k, v in self.store.items() if k in keep}


def DIIScore(Ds, Fs, Es, cache):
    R = [cache.get(d, f) for d, f in zip(Ds, Fs)]
    n = len(Ds)
    B = np.empty((n, n))
    for i in range(n):
        for j in range(i + 1):
            B[i, j] = B[j, i] = float(np.vdot(R[i], R[j]))
    return np.zeros(n), B


def EDIIScore(Ds, Fs, Es, cache):
    """B_ij = -<D_i - D_j, F_i - F_j> (HartreeFock.jl:1273-1285), from the Gram
    matrix G_ij = <D_i, F_j>: B_ij = -(G_ii + G_jj - G_ij - G_ji)."""
    Dm = np.stack([np.ravel(d) for d in Ds])
    Fm = np.stack([np.ravel(f) for f in Fs])
    G = Dm @ Fm.T
    g = np.diag(G)
    B = -(g[:, None] + g[None, :] - G - G.T)
    return np.array(Es, dtype=float), B


def ADIIScore(Ds, Fs, Es, cache):
    """v_i = <D_i - D_L, F_L>, B_ij = <D_i - D_L, F_j - F_L> (HartreeFock.jl:1287-1296)."""
    DL, FL = np.ravel(Ds[-1]), np.ravel(Fs[-1])
    dD = np.stack([np.ravel(d) - DL for d in Ds])
    dF = np.stack([np.ravel(f) - FL for f in Fs])
    return dD @ FL, dD @ dF.T


DIISconfigs = dict(DIIS=(False, DIIScore), EDIIS=(True, EDIIScore), ADIIS=(True, ADIIScore))


def initializeHFcore(method, vars_, config, kws):
    Ns = formatSpinConfiguration(_hft_of(vars_), config.spin)
    X, S, Hcore, twoBody = config.orthonormalization, config.overlap, config.oneBody, config.twoBody
    if method == "DD":
        ds = float(kws.get("dampStrength", defaultDS))
        if not (0 <= ds <= 1):
            raise ValueError("The value of `dampStrength` should be between 0 and 1.")

        def DDcore():
            Fs = [tv.Fs[-1] for tv in vars_]
            Ds = [tv.Ds[-1] for tv in vars_]
            Dnew = tuple((1 - ds) * getD(getC(X, F), n) + ds * D for n, F, D in zip(Ns, Fs, Ds))
            return getCDFE(Hcore, twoBody, X, Ns, getF(Hcore, twoBody, Dnew))
        return DDcore, dict(dampStrength=ds)

    resetThreshold = float(kws.get("resetThreshold", 1000 * ATOL))
    DIISsize = int(kws.get("DIISsize", defaultDIISsize))
    solver = str(kws.get("solver", defaultDIISsolver)).lstrip(":")
    if DIISsize < 2:
        raise ValueError("%s space need to be at least 2." % method)
    if solver not in ConstraintSolvers:
        raise ValueError("unknown DIIS solver %r" % solver)
    lens = [len(getattr(vars_[0], f)) for f in ("Ds", "Fs", "Es")]
    init = min(DIISsize, *lens)
    cs = [list(np.arange(1, init + 1, dtype=float)) for _ in vars_]
    Dss = [list(tv.Ds[-init:]) for tv in vars_]
    Fss = [list(tv.Fs[-init:]) for tv in vars_]
    Ess = [list(tv.Es[-init:]) for tv in vars_]
    cvx, mDIIS = DIISconfigs[method]
    caches = [_ResidualCache(S, X) for _ in vars_]
    solve = ConstraintSolvers[solver]

    def xDIIScore():
        Fn = []
        for k in range(len(vars_)):
            v, B = mDIIS(Dss[k], Fss[k], Ess[k], caches[k])
            c = solve(cvx, np.array(cs[k]), v, B)
            cs[k][:] = list(c)
            Fn.append(sum(ci * fi for ci, fi in zip(c, Fss[k])))
        res = getCDFE(Hcore, twoBody, X, Ns, tuple(Fn))
        for k in range(len(vars_)):
            c, Ds, Fs, Es = cs[k], Dss[k], Fss[k], Ess[k]
            c.append(1.0)
            Ds.append(res[k][1])
            Fs.append(res[k][2])
            Es.append(res[k][3])
            if len(Es) > 2 and Es[-1] - Es[-2] > resetThreshold:
                keep = len(Es) - 2
                for lst in (c, Ds, Fs, Es):
                    e = lst[keep]
                    del lst[1:]
                    lst[0] = e
            elif len(c) > DIISsize:
                pop = int(np.argmax(Es))
                for lst in (c, Ds, Fs, Es):
                    lst.pop(pop)
            caches[k].prune(list(zip(Ds, Fs)))
        return res

    return xDIIScore, dict(resetThreshold=resetThreshold, DIISsize=DIISsize, solver=solver)


def _hft_of(vars_):
    return RCHartreeFock() if len(vars_) == 1 else UOHartreeFock()


def getErrorNrms(vars_, S):
    """HartreeFock.jl:1230-1237"""
    return sum(rmsOf(getEresidual(tv.Fs[-1], tv.Ds[-1], S)) for tv in vars_) / len(vars_)


def isOscillateConverged(seq, tarThreshold, target=float("nan"), stdThreshold=None, minLen=5, maxRemains=10):
    """Tools.jl:291-316"""
    if stdThreshold is None:
        stdThreshold = 0.8 * tarThreshold
    if len(seq) < minLen:
        tail = [0.0]
    else:
        sl = min(maxRemains, len(seq) // 3 + 1)
        tail = list(seq[-sl:])
    if len(tail) < 2:
        return False, tail
    valleys = sorted(tail)[: len(tail) // 2 + 1]
    if math.isnan(target):
        target = min(valleys)
    b = np.std(valleys, ddof=1) <= stdThreshold and abs(target - seq[-1]) <= tarThreshold
    return bool(b), float(np.std(tail, ddof=1))


def runHartreeFockCore(HFT, scf, config, C0, maxStep, earlyStop, saveTrace, printInfo=False,
                       infoLevel=defaultHFinfoLevel, log=None):
    """The SCF iteration (HartreeFock.jl:1049-1228)."""
    log = log if log is not None else (lambda *a: print(*a, flush=True))
    S = config.overlap
    vars_ = initializeSCF(HFT, config, C0)
    ratio = scf.secondaryConvRatio
    sh = vars_[0].shared
    Etots = sh.Etots
    dEs = [0.0]
    dDrms = [0.0]
    dFrms = [getErrorNrms(vars_, S)]
    L = len(scf.method)
    ivs = scf.interval
    dFbreak = [ratio[0] * x for x in ivs]
    dDbreak = [ratio[1] * x for x in ivs]
    dEosc = [max(x, scf.oscillateThreshold) for x in ivs]
    dFosc = [ratio[0] * x for x in dEosc]
    dDosc = [ratio[1] * x for x in dEosc]
    detect = not math.isnan(ivs[-1])
    converged = True
    rollback = range(0, HFminItr // 3 + 1)
    rbCount = len(rollback)
    memLens = [max(*HFinterMatStoreSizes, s) for s in scf.cache_sizes()]
    i = 0
    if printInfo and infoLevel > 0:
        log("•Initial HF energy E: %.10f Ha" % Etots[-1])
        log("•Initial RMS(FDS-SDF): %.10f" % dFrms[-1])
        log("•Convergence Threshold of E: %g Ha" % ivs[-1])
    for l, (method, kws, bp, mlen) in enumerate(zip(scf.method, scf.methodConfig, ivs, memLens)):
        core, keyArgs = initializeHFcore(method, vars_, config, kws)
        fluc = max(10 * bp, 1.5e-3)
        converged = False
        endM = l == L - 1
        n = 0
        if printInfo and infoLevel > 1:
            log("|–<%d>–[:%s]" % (l + 1, method))
        while True:
            if not i < maxStep:
                break
            i += 1
            n += 1
            updateHFtempInfo(mlen, vars_, core())
            dEs.append(Etots[-1] - Etots[-2])
            if endM or printInfo:
                dDrms.append(rmsOf(sh.Dtots[-1] - sh.Dtots[-2]))
                dFrms.append(getErrorNrms(vars_, S))
            dE, dD, dF = dEs[-1], dDrms[-1], dFrms[-1]
            observe = getattr(config.twoBody, "observe", None)
            if observe is not None:
                observe(dE, dF)
            if printInfo and infoLevel > 0:
                log("| %d | %.12f | %.3e | %.3e | %.3e" % (i, Etots[-1], dE, dF, dD))
            thr = (bp, dDbreak[l])
            if dF > dFbreak[l]:
                thr = (0.0, 0.0)
            if abs(dE) <= thr[0] and dD <= thr[1]:
                converged = True
                break
            if n > 1 and i > HFminItr and dE > fluc:
                isOsc = False
                if scf.oscillateThreshold > 0:
                    isOsc, _ = isOscillateConverged(Etots, 10 * dEosc[l], minLen=HFminItr,
                                                    maxRemains=HFinterEsOscConvSize)
                    if isOsc and abs(dE) <= dEosc[l] and ((dF <= dFosc[l] and dD <= dDosc[l]) if endM else True):
                        converged = 1
                        break
                if earlyStop and not isOsc:
                    if all(dEs[-1 - j] > 10 * fluc for j in rollback):
                        memLen = len(vars_[0].Cs)
                        rb = min(rbCount, max(memLen - 1, 0))
                        popHFtempInfo(vars_, rb)
                        if printInfo:
                            log("Early termination of %s due to its poor performance." % method)
                        i -= rb
                        break
    if printInfo:
        status = ("converged to an oscillation" if converged == 1 and converged is not True else
                  ("converged" if converged else "stopped but not converged")) if detect else "stopped"
        log("The SCF iteration of %s has %s at step %d: |ΔE| → %.3e Ha, RMS(FDS-SDF) → %.3e, RMS(ΔD) → %.3e."
            % (type(HFT).__name__, status, i, abs(dEs[-1]), dFrms[-1], dDrms[-1]))
    clearHFtempInfo(saveTrace, vars_)
    if not detect:
        converged = None
    elif converged is not True and converged is not False:
        converged = True
    return vars_, converged


def getOrbitalOccupations(HFT, X, Ns, Fs):
    """Orbital energies and spin occupations (HartreeFock.jl:708-750)."""
    out = []
    if isinstance(HFT, RCHartreeFock):
        e = solveFockMatrix(X, Fs[0])[1]
        nel = 2 * Ns[0]
        occ = [(True, True) if 2 * (k + 1) <= nel else (False, False) for k in range(len(e))]
        return ((e, occ),)
    for s, (N, F) in enumerate(zip(Ns, Fs)):
        e = solveFockMatrix(X, F)[1]
        mark = (True, False) if s == 0 else (False, True)
        out.append((e, [mark if k + 1 <= N else (False, False) for k in range(len(e))]))
    return tuple(out)


@dataclass
class HFfinalInfo:
    """Final SCF state (HartreeFock.jl:791-819): energy = (E_elec, E_nn)."""
    system: dict
    energy: tuple
    coeff: tuple
    density: tuple
    fock: tuple
    occu: tuple
    memory: tuple
    converged: object
    basis: object
    timing: dict = field(default_factory=dict)


def _is_engine(obj):
    return all(hasattr(obj, k) for k in ("n", "overlap", "core_hamiltonian", "two_body"))


def initializeHartreeFock(initCfg, basis, system, twoBody=None, schwarz=1e-12, precision=_lib.QBX_PREC_FP64,
                          incremental=False):
    """Integrals, orthonormalization and initial coefficients (HartreeFock.jl:178-210).
    `basis` is an integral engine (GPUBasis or any object with n, overlap(),
    core_hamiltonian(Z, R), two_body(threshold, precision)) or orbitals."""
    spin, nuc = system["spin"], system["geometry"]
    gb = basis if _is_engine(basis) else GPUBasis(basis)
    S = gb.overlap()
    Z, R = nuc.charges, nuc.coord_array
    Hcore = gb.core_hamiltonian(Z, R)
    if twoBody is None:
        twoBody = gb.two_body(schwarz, precision, incremental=incremental) if incremental else \
            gb.two_body(schwarz, precision)
    X = gb.orthonormalization() if hasattr(gb, "orthonormalization") else getOrthonormalization(S)
    config = ElecHamiltonianConfig(spin, X, S, Hcore, twoBody)
    HFT = initCfg.HFT
    if initCfg.method == "Direct":
        for c in initCfg.data:
            if c.shape != (gb.n, gb.n):
                raise ValueError("DimensionMismatch: the size of the input initial orbital coefficient matrix "
                                 "%s does not match the basis-set size (%d)." % (c.shape, gb.n))
        C0 = initCfg.data
    elif initCfg.method == "CoreH":
        C0 = guessCoreH(HFT, X, Hcore)
    elif initCfg.method == "GWH":
        C0 = guessGWH(HFT, X, S, Hcore)
    else:
        C0 = guessSAD(HFT, nuc, config, lambda k: gb.core_hamiltonian(Z[k:k + 1], R[k:k + 1]))
    return config, C0, gb


def runHartreeFock(nucInfo, basis, config=None, *, printInfo=True, infoLevel=defaultHFinfoLevel,
                   twoBody=None, schwarzThreshold=1e-12, precision=_lib.QBX_PREC_FP64, incremental=False,
                   densitySolver="eigen", log=None):
    """runHartreeFock(nucInfo, basis, config) or runHartreeFock((spin, nucInfo), basis, config)
    (HartreeFock.jl:953-1046).  `basis` is a list of orbitals, a FlatBasis or a
    GPUBasis.  Returns HFfinalInfo with energy = (E_HF electronic, E_nuclear).
    Beyond the reference: `precision` (QBX_PREC_FP64 / QBX_PREC_MIXED) and
    `incremental` (density-screened incremental Fock builds) select the GPU
    Fock-build mode; `densitySolver` selects each step's orbitals and density:
    "eigen" (default) diagonalizes X' F X as the reference does, on the GPU
    when the Fock build is and the basis has at least DEVICE_EIGEN_MIN_N
    functions (matrix-core GEMMs, rocSOLVER eigensolver), on the host (numpy)
    below that; "eigen-device" and "eigen-host" force either; "purify" forms
    the density by purification on the GPU instead (the same matrix; the
    coefficients of the result still come from one diagonalization).  The
    defaults reproduce the reference's full FP64 builds and eigen steps."""
    if densitySolver not in ("eigen", "eigen-device", "eigen-host", "purify"):
        raise ValueError("densitySolver must be 'eigen', 'eigen-device', 'eigen-host' or 'purify'")
    t0 = time.time()
    if isinstance(nucInfo, tuple) and len(nucInfo) == 2 and isinstance(nucInfo[1], NuclearCluster):
        spin, nuc = nucInfo
        if not isinstance(spin, OccupationState):
            spin = OccupationState(spin)
    else:
        nuc = nucInfo
        spin = prepareSpinConfiguration(nuc)
    system = dict(spin=spin, geometry=nuc)
    total = spin.total()
    if config is None:
        config = HFconfig(RCHartreeFock() if total % 2 == 0 else UOHartreeFock())
    HFT = config.type
    minElec = 1 if isinstance(HFT, RCHartreeFock) else 0
    if total <= minElec:
        raise ValueError("%s requires more than %d electrons." % (type(HFT).__name__, minElec))
    if _is_engine(basis):
        nb = basis.n
    elif isinstance(basis, FlatBasis):
        nb = basis.n_bf
    else:
        nb = len(basis)
    if nb < max(spin.layout):
        raise ValueError("The basis-set size should be no less than %d." % max(spin.layout))
    hamil, C0, gb = initializeHartreeFock(config.initial, basis, system, twoBody, schwarzThreshold, precision,
                                          incremental)
    solver = None
    eig = None
    if densitySolver == "purify":
        if not isinstance(hamil.twoBody, DirectJK):
            raise ValueError("densitySolver='purify' runs with the GPU Fock build")
        solver = hamil.twoBody.density_solver = PurifiedDensity(gb.ctx)
    elif (densitySolver in ("eigen", "eigen-device") and isinstance(hamil.twoBody, DirectJK) and
          hasattr(getattr(gb, "ctx", None), "fock_solve") and
          (densitySolver == "eigen-device" or nb >= DEVICE_EIGEN_MIN_N)):
        eig = hamil.twoBody.eigen_solver = DeviceEigen(gb.ctx)
    t1 = time.time()
    if printInfo and infoLevel > 1:
        (log or print)("Hartree–Fock (HF) Initialization:\n•HF Type: %s\n•Basis Set Size: %d\n"
                       "•Initial Guess Method: %s" % (HFT.name, nb, config.initial.method))
    vars_, converged = runHartreeFockCore(HFT, config.strategy, hamil, C0, config.maxStep, config.earlyStop,
                                          config.saveTrace, printInfo, infoLevel, log)
    t2 = time.time()
    Ns = tuple(tv.N for tv in vars_)
    Fs = tuple(tv.Fs[-1] for tv in vars_)
    occu = getOrbitalOccupations(HFT, hamil.orthonormalization, Ns, Fs)
    Enuc = nucRepulsion(nuc)
    Eel = vars_[0].shared.Etots[-1]
    Cfin = tuple(_coeff(tv.Cs[-1]) for tv in vars_)
    timing = dict(init_s=t1 - t0, scf_s=t2 - t1, fock_builds=getattr(hamil.twoBody, "n_builds", None),
                  mixed_builds=getattr(hamil.twoBody, "n_mixed", None),
                  builds=list(getattr(hamil.twoBody, "builds", [])))
    if solver is not None:
        hamil.twoBody.density_solver = None
        timing.update(purified_steps=solver.calls - solver.fallbacks, purify_fallbacks=solver.fallbacks,
                      purify_products=solver.products)
    if eig is not None:
        hamil.twoBody.eigen_solver = None
        timing.update(device_eigen_steps=eig.calls)
    res = HFfinalInfo(system, (Eel, Enuc), Cfin, tuple(tv.Ds[-1] for tv in vars_),
                      Fs, occu, vars_, converged, gb, timing)
    if printInfo:
        (log or print)("Hartree–Fock Energy ¦ Nuclear Repulsion ¦ Total Energy\n%.10f Ha   %.10f Ha   %.10f Ha"
                       % (Eel, Enuc, Eel + Enuc))
    return res
