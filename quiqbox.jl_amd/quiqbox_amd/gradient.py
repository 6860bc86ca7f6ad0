"""RHF and UHF energy gradients with respect to the basis parameters, and basis-set
optimization built on them.

Quiqbox v0.6.3 keeps the parameters of a basis in boxes (`Parameters.jl:14-43`,
`genCellParam` `:387-405`) that a caller can mark as variables, but it offers
no derivative of the Hartree-Fock energy; this module adds one for bases of
contracted Cartesian Gaussian shells whose parameters are

    theta = (exponents, contraction coefficients, centres),

each function built as genGaussTypeOrb(centre, exponents, c_k N(alpha_k, ijk),
ijk; outerRenormalize=true), i.e. weights w_k = c_k N(alpha_k, ijk) / sqrt(S_mumu)
(`OrbitalBases.jl:476-485, 502-534`, `Framework.jl:748-773`).  Centres belong
to sites (atoms) and may be shared by several shells; the nuclei stay fixed
(floating orbitals).

The gradient of a converged RHF energy is

    dE/dtheta = sum D dH/dtheta - sum W dS/dtheta + 1/2 sum Gamma d(mn|ls)/dtheta,

D the total density, W = 2 C_occ diag(eps) C_occ^T; the GPU (qbx_rhf_gradient)
returns it per radial primitive and shell with every function's weights held
fixed, and `rhf_gradient` chains it through the normalization above:

    dE/dc_k     = g_lin_k / c_k                          - 1/2 sum_mu t_mu dln S_mumu/dc_k
    dE/dalpha_k = g_xpn_k + (3/4 + L/2) g_lin_k / alpha_k - 1/2 sum_mu t_mu dln S_mumu/dalpha_k
    dE/dR_site  = sum over the site's shells of g_cen

with t_mu = 2 (D F - W S)_mumu, the derivative with respect to scaling
function mu (zero at self-consistency, kept for finite convergence).

For a UHF state (`uhf_gradient`, qbx_uhf_gradient) Gamma takes the exchange
of each spin density, W sums both spins' C_occ diag(eps) C_occ^T, and
t_mu = 2 sum_s (D_s F_s)_mumu - 2 (W S)_mumu.
"""
import math

import numpy as np

from . import systems
from .basis import subshellXYZs
from .engine import GPUBasis
from .hartree_fock import HFconfig, OrbCoeffInitialConfig, RCHartreeFock, SCFconfig, runHartreeFock


class ParamBasis:
    """Contracted Cartesian shells with exponents, coefficients and site
    centres as parameters.  shells: list of (site, L, exponents, coefficients)."""

    def __init__(self, centers, shells):
        self.centers = np.array(centers, dtype=np.float64).reshape(-1, 3)
        self.shells = [(int(s), int(L), np.array(x, dtype=np.float64), np.array(c, dtype=np.float64))
                       for s, L, x, c in shells]
        for s, L, x, c in self.shells:
            if not 0 <= s < len(self.centers):
                raise ValueError("shell site %d out of range" % s)
            if len(x) != len(c) or len(x) == 0:
                raise ValueError("every shell needs as many coefficients as exponents")
            if np.any(x <= 0):
                raise ValueError("exponents must be positive")

    @classmethod
    def random_h_cluster(cls, n_atoms=32, seed=0, box=10.0, dmin=1.2, nprim=4):
        """The parameters of systems.h_cluster_sp (same seed, same values)."""
        rng = np.random.Generator(np.random.PCG64(seed))
        coords = systems._place_atoms(rng, n_atoms, box, dmin)
        shells = []
        for i in range(n_atoms):
            for L in (0, 1):
                xp = np.exp(rng.uniform(math.log(0.1), math.log(10.0), size=nprim))
                co = rng.uniform(0.1, 1.0, size=nprim)
                shells.append((i, L, xp, co))
        return cls(coords, shells)

    # ---- parameter vector: all exponents, all coefficients, all centres ----
    @property
    def n_theta(self):
        return 2 * sum(len(x) for _, _, x, _ in self.shells) + self.centers.size

    def theta(self):
        xs = np.concatenate([x for _, _, x, _ in self.shells])
        cs = np.concatenate([c for _, _, _, c in self.shells])
        return np.concatenate([xs, cs, self.centers.ravel()])

    def with_theta(self, th):
        th = np.asarray(th, dtype=np.float64)
        if th.shape != (self.n_theta,):
            raise ValueError("theta has %s entries, expected %d" % (th.shape, self.n_theta))
        nx = sum(len(x) for _, _, x, _ in self.shells)
        xs, cs, R = th[:nx], th[nx:2 * nx], th[2 * nx:]
        out, k = [], 0
        for s, L, x, _ in self.shells:
            n = len(x)
            out.append((s, L, xs[k:k + n], cs[k:k + n]))
            k += n
        return ParamBasis(R.reshape(-1, 3), out)

    def orbitals(self):
        bs = []
        for s, L, x, c in self.shells:
            bs += systems._shell(tuple(self.centers[s]), L, x, c)
        return bs


def _lnS_derivatives(L, x, c):
    """d ln S_mumu / d c_k and d ln S_mumu / d alpha_k of one function of a
    shell of angular momentum L before renormalization,
    S = sum_kl c_k c_l N_k N_l <g_k|g_l> (same centre).  The primitive norm
    N(alpha, ijk) is alpha^((2L+3)/4) times a factor of ijk alone, and
    <g_k|g_l> is (alpha_k + alpha_l)^-(L+3/2) times a factor of ijk alone, so
    with f_k = c_k alpha_k^a (a = 3/4 + L/2) and O_kl = (alpha_k + alpha_l)^-b
    (b = L + 3/2), ln S = const + ln(f' O f) for every Cartesian component:
        d ln S/dc_k     = 2 alpha_k^a (O f)_k / S
        d ln S/dalpha_k = 2 f_k (a/alpha_k (O f)_k - b (O'' f)_k) / S,
    O''_kl = O_kl / (alpha_k + alpha_l)."""
    x = np.asarray(x, dtype=np.float64)
    c = np.asarray(c, dtype=np.float64)
    a, b = 0.75 + 0.5 * L, L + 1.5
    xa = x ** a
    f = c * xa
    P = x[:, None] + x[None, :]
    O = P ** (-b)
    Of = O @ f
    S = float(f @ Of)
    dc = 2.0 * xa * Of / S
    dx = 2.0 * f * (a / x * Of - b * ((O / P) @ f)) / S
    return dc, dx


def _chain(pb, gx, gl, gc, sh, t):
    """dE/dtheta of the ParamBasis from the kernels' per-primitive and
    per-shell sums (weights held fixed) and t_mu, the derivative with respect
    to scaling function mu, through the normalization (module docstring)."""
    # kernel shell of each user shell: match the first component's function
    first_bf = {}
    for k in range(len(sh["L"])):
        comps = [f for f in sh["bf"][k] if f >= 0]
        first_bf[min(comps)] = k
    g_x, g_c = [], []
    gR = np.zeros_like(pb.centers)
    f0 = 0
    for s, L, x, c in pb.shells:
        k = first_bf.get(f0)
        if k is None or sh["L"][k] != L or sh["nprim"][k] != len(x):
            raise RuntimeError("the kernels' shell table does not match the parametrized basis")
        o = sh["poff"][k]
        if not np.allclose(sh["xpn"][o:o + len(x)], x, rtol=0, atol=0):
            raise RuntimeError("radial primitives out of order in the kernels' shell table")
        gxs, gls = gx[o:o + len(x)], gl[o:o + len(x)]
        dEdc = gls / c
        dEdx = gxs + (0.75 + 0.5 * L) * gls / x
        # every component of the shell has the same d ln S (see _lnS_derivatives)
        tsum = float(np.sum(t[f0:f0 + len(subshellXYZs(L))]))
        if tsum != 0.0:
            dc, dx = _lnS_derivatives(L, x, c)
            dEdc = dEdc - 0.5 * tsum * dc
            dEdx = dEdx - 0.5 * tsum * dx
        g_x.append(dEdx)
        g_c.append(dEdc)
        gR[s] += gc[k]
        f0 += len(subshellXYZs(L))
    return np.concatenate([np.concatenate(g_x), np.concatenate(g_c), gR.ravel()])


def _occupied(res, sector):
    """Occupied coefficients and orbital energies of one spin sector."""
    C = res.coeff[sector]
    eps = np.asarray(res.occu[sector][0])
    occ = np.array([bool(o[0] or o[1]) for o in res.occu[sector][1]])
    return C[:, occ], eps[occ]


def rhf_gradient(nuc, pb, config=None, schwarz=1e-12, res=None, gpu_basis=None):
    """E_RHF (electronic + nuclear) and dE/dtheta of the ParamBasis `pb`.
    Runs the GPU SCF unless a converged result `res` of the same basis is given."""
    gb = gpu_basis or GPUBasis(pb.orbitals())
    if res is None:
        cfg = config or HFconfig(initial="CoreH", strategy=SCFconfig(("ADIIS", "DIIS"), (1e-4, 1e-11)), maxStep=200)
        res = runHartreeFock(nuc, gb, cfg, printInfo=False, schwarzThreshold=schwarz)
    if len(res.coeff) != 1:
        raise ValueError("rhf_gradient needs a restricted closed-shell result")
    C = res.coeff[0]
    eps = np.asarray(res.occu[0][0])
    nocc = int(sum(1 for o in res.occu[0][1] if o[0]))
    Co = C[:, :nocc]
    Dt = 2.0 * Co @ Co.T
    W = 2.0 * (Co * eps[:nocc]) @ Co.T
    gx, gl, gc, sh = gb.rhf_gradient(Dt, W, nuc.charges, nuc.coord_array, schwarz)
    F = res.fock[0]
    S = gb.overlap()
    t = 2.0 * np.einsum("ij,ji->i", Dt, F) - 2.0 * np.einsum("ij,ji->i", W, S)
    grad = _chain(pb, gx, gl, gc, sh, t)
    E = res.energy[0] + res.energy[1]
    return E, grad, res


def uhf_gradient(nuc, pb, config=None, schwarz=1e-12, res=None, gpu_basis=None):
    """E_UHF (electronic + nuclear) and dE/dtheta of the ParamBasis `pb`
    (qbx_uhf_gradient; the chain rule as rhf_gradient with
    t_mu = 2 sum_s (D_s F_s)_mumu - 2 (W S)_mumu).  Runs the GPU UHF SCF
    (the reference's UOHartreeFock) unless a converged UHF result `res` of
    the same basis is given."""
    from .hartree_fock import UOHartreeFock
    gb = gpu_basis or GPUBasis(pb.orbitals())
    if res is None:
        cfg = config or HFconfig(UOHartreeFock(), initial="CoreH",
                                 strategy=SCFconfig(("ADIIS", "DIIS"), (1e-4, 1e-11)), maxStep=200)
        res = runHartreeFock(nuc, gb, cfg, printInfo=False, schwarzThreshold=schwarz)
    if len(res.coeff) != 2:
        raise ValueError("uhf_gradient needs an unrestricted result")
    Ds, W = [], 0.0
    for sector in (0, 1):
        Co, eo = _occupied(res, sector)
        Ds.append(Co @ Co.T)
        W = W + (Co * eo) @ Co.T
    Dt = Ds[0] + Ds[1]
    gx, gl, gc, sh = gb.uhf_gradient(Dt, Ds[0], Ds[1], W, nuc.charges, nuc.coord_array, schwarz)
    S = gb.overlap()
    t = 2.0 * sum(np.einsum("ij,ji->i", Ds[k], res.fock[k]) for k in (0, 1)) - 2.0 * np.einsum("ij,ji->i", W, S)
    grad = _chain(pb, gx, gl, gc, sh, t)
    E = res.energy[0] + res.energy[1]
    return E, grad, res


def optimize_basis(nuc, pb, steps=100, lr=0.05, schwarz=1e-12, config=None, callback=None):
    """Plain gradient descent on theta in log space for exponents and
    coefficients (positive parameters) and in bohr for the centres:
    log x <- log x - lr * x dE/dx, R <- R - lr dE/dR.  Returns the final
    ParamBasis and the energy history.  The SCF of each step after the first
    starts from the previous step's converged coefficients, with DIIS alone to
    the configuration's final threshold (the starting point is already near
    convergence, where ADIIS/EDIIS only add their host-side solves: 100 steps
    of the 32-atom cluster 23.1 -> 22.5 s)."""
    hist = []
    res = None
    nx = sum(len(x) for _, _, x, _ in pb.shells)
    base = config or HFconfig(initial="CoreH", strategy=SCFconfig(("ADIIS", "DIIS"), (1e-4, 1e-10)), maxStep=200)
    st = base.strategy
    warm = SCFconfig(("DIIS",), (st.interval[-1],), secondaryConvRatio=st.secondaryConvRatio,
                     oscillateThreshold=st.oscillateThreshold)
    for it in range(steps):
        gb = GPUBasis(pb.orbitals())
        cfg = base
        if res is not None and res.converged:
            cfg = HFconfig(initial=OrbCoeffInitialConfig(RCHartreeFock(), (res.coeff[0],)), strategy=warm,
                           maxStep=base.maxStep)
        E, g, res = rhf_gradient(nuc, pb, cfg, schwarz, gpu_basis=gb)
        hist.append(E)
        if callback:
            callback(it, E, g)
        th = pb.theta()
        pos = th[:2 * nx]
        step_pos = np.exp(-lr * pos * g[:2 * nx])
        th_new = th.copy()
        th_new[:2 * nx] = pos * step_pos
        th_new[2 * nx:] = th[2 * nx:] - lr * g[2 * nx:]
        pb = pb.with_theta(th_new)
    return pb, hist
