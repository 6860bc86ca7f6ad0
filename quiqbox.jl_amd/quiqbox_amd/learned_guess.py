"""Learned SCF initial guess: a pair-feature network that predicts the RHF
density from the overlap and core-Hamiltonian blocks of every atom pair.

The reference's guesses are the core Hamiltonian, GWH and SAD
(`HartreeFock.jl:215-293`); this is an addition with no counterpart there.
For atoms A, B with function blocks S_AB and H_AB (zero-padded to NMAX x NMAX)
the features are [S_AB, H_AB, element flags of A and B, A == B, R, exp(-R),
1/(1+R)] normalized with the training set's statistics, where H_AB is the pair's
local core Hamiltonian T_AB + V(A)_AB + V(B)_AB: the kinetic block plus the
attraction of the two atoms' own nuclei only (the full core Hamiltonian holds the
attraction of every nucleus, so its blocks would not mean the same in a small
training cluster and a large one).  A 3-layer MLP (width 256, ReLU) maps the
features to the padded density block D_AB.  The
network runs on the GPU in BF16 with FP32 accumulation through the library's
own kernels (qbx_gemm_bf16 on the matrix cores, qbx_mlp_* for the elementwise
parts, Adam on FP32 master weights), for training as well as inference.  torch
only provides device buffers.

Use: a trained `LearnedGuess` turns (nuclei, basis) into initial coefficients:
the predicted blocks form a symmetric density, one Fock build (GPU J/K) gives
F(D_pred), and its eigenvectors in the orthonormal basis are the guess
(`initial_coefficients`); `guess_config` wraps them in the reference's
OrbCoeffInitialConfig(:Direct) form, and falls back to the core-Hamiltonian
guess when the predicted density is not usable (non-finite, or its electron
count tr(D S) off by more than 10%).  Atom pairs farther apart than any pair of
the training set get a zero block instead of an extrapolated one.  (Quiqbox's
SAD, the other plain guess, runs one full-basis UHF per nucleus,
`HartreeFock.jl:256-293`: minutes at 600 functions, so it is not the fallback.)
"""
import ctypes
import os
import time

import numpy as np

from . import _lib
from .engine import GPUBasis
from .hartree_fock import (HFconfig, OrbCoeffInitialConfig, RCHartreeFock, SCFconfig, getC,
                           runHartreeFock)

WIDTH = 256


class PairLayout:
    """Elements the network knows and functions per atom in the padded blocks.
    Features per atom pair: S and local core-Hamiltonian blocks (nmax^2 each),
    element flags of both atoms, A == B, R, exp(-R), 1/(1+R)."""

    def __init__(self, elements, nmax):
        self.elements = tuple(elements)
        self.nmax = int(nmax)
        self.n_feat = 2 * self.nmax * self.nmax + 2 * len(self.elements) + 4
        self.n_in = (self.n_feat + 7) // 8 * 8   # padded to the MFMA K step
        self.n_out = self.nmax * self.nmax
        self.dist_col = 2 * self.nmax * self.nmax + 2 * len(self.elements) + 1

    def __repr__(self):
        return "PairLayout(%s, nmax=%d)" % (self.elements, self.nmax)


# cc-pVDZ water (oxygen: 15 Cartesian functions)
WATER = PairLayout(("H", "O"), 15)
# clusters of the random C/H/O generator with s and p shells on every atom (4 functions)
CHO_SP = PairLayout(("H", "C", "O"), 4)

# the default layout's sizes (water)
NMAX = WATER.nmax
ELEMENTS = WATER.elements
N_FEAT = WATER.n_feat
N_IN = WATER.n_in
N_OUT = WATER.n_out


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the learned guess runs on the GPU (no device found)")
    return torch


def atom_blocks(nuc, gb, layout=WATER):
    """Function indices of every atom (functions whose centre is the nucleus)."""
    cen = np.asarray(gb.flat.prim_cen)
    first = np.asarray(gb.flat.bf_prim)[np.asarray(gb.flat.bf_ptr)[:-1]]
    fc = cen[first]
    R = nuc.coord_array
    owner = np.argmin(((fc[:, None, :] - R[None, :, :]) ** 2).sum(-1), axis=1)
    if np.any(np.sqrt(((fc - R[owner]) ** 2).sum(-1)) > 1e-8):
        raise ValueError("every basis function must sit on a nucleus")
    blocks = [np.nonzero(owner == a)[0] for a in range(len(R))]
    if max(len(b) for b in blocks) > layout.nmax:
        raise ValueError("atoms with more than %d functions" % layout.nmax)
    for s in nuc.symbols:
        if s not in layout.elements:
            raise ValueError("element %s outside the trained set %s" % (s, layout.elements))
    return blocks


def _padded_index(blocks, n, nmax):
    """(n_atoms, nmax) function indices of every atom, padded with n (a zero
    row/column appended to the matrices)."""
    idx = np.full((len(blocks), nmax), n, dtype=np.int64)
    for a, b in enumerate(blocks):
        idx[a, :len(b)] = b
    return idx


def _ext(M):
    n = M.shape[0]
    E = np.zeros((n + 1, n + 1))
    E[:n, :n] = M
    return E


def pair_features(nuc, blocks, S, H, layout=WATER):
    """Raw features (n_atoms^2 x N_FEAT) of all ordered atom pairs.  H is
    (T, O): the kinetic matrix and the own-nucleus attraction O[i][j] (the
    nucleus at function i's centre alone, GPUBasis.nuclear_own), so that the
    pair (A, B) gets T + V(A) + V(B) = T + O + O^T on its block, and the pair
    (A, A) gets T + V(A)."""
    T, O = H
    nm, els = layout.nmax, layout.elements
    na = len(blocks)
    n = S.shape[0]
    idx = _padded_index(blocks, n, nm)
    I, J = idx[:, None, :, None], idx[None, :, None, :]
    Sb = _ext(S)[I, J]                                  # (na, na, nmax, nmax)
    OTb = _ext(O.T)[I, J]
    Hb = _ext(T)[I, J] + _ext(O)[I, J] + OTb
    diag = np.arange(na)
    Hb[diag, diag] -= OTb[diag, diag]                   # V(A) once on the atom's own block
    X = np.zeros((na * na, layout.n_feat))
    X[:, :nm * nm] = Sb.reshape(na * na, -1)
    X[:, nm * nm:2 * nm * nm] = Hb.reshape(na * na, -1)
    o = 2 * nm * nm
    el = np.array([els.index(sym) for sym in nuc.symbols])
    X[np.arange(na * na), o + np.repeat(el, na)] = 1.0
    X[np.arange(na * na), o + len(els) + np.tile(el, na)] = 1.0
    R = nuc.coord_array
    d = np.sqrt(((R[:, None, :] - R[None, :, :]) ** 2).sum(-1)).ravel()
    X[:, o + 2 * len(els)] = np.eye(na).ravel()
    X[:, o + 2 * len(els) + 1] = d
    X[:, o + 2 * len(els) + 2] = np.exp(-d)
    X[:, o + 2 * len(els) + 3] = 1.0 / (1.0 + d)
    return X


def pair_targets(blocks, D, layout=WATER):
    na = len(blocks)
    idx = _padded_index(blocks, D.shape[0], layout.nmax)
    return _ext(D)[idx[:, None, :, None], idx[None, :, None, :]].reshape(na * na, layout.n_out)


def assemble_density(blocks, Y, n, layout=WATER):
    na = len(blocks)
    nm = layout.nmax
    idx = _padded_index(blocks, n, nm)
    D = np.zeros((n + 1, n + 1))
    D[idx[:, None, :, None], idx[None, :, None, :]] = Y.reshape(na, na, nm, nm)
    D = D[:n, :n]
    return 0.5 * (D + D.T)


def system_data(nuc, gb, layout=WATER):
    """S, (T, own-nucleus attraction), atom blocks; one-electron integrals from
    the GPU kernels."""
    S = gb.overlap()
    T = gb.kinetic()
    blocks = atom_blocks(nuc, gb, layout)
    O = gb.nuclear_own(nuc.charges, nuc.coord_array)
    return S, (T, O), blocks


class _Dev:
    """Device buffers of one layer set (FP32 master weights, BF16 copies, Adam
    moments).  Weights start uniform in +-1/sqrt(fan_in), biases at zero: with
    the larger He scale sqrt(2/fan_in) the 225 outputs start at O(1) each and
    Adam stalls at a loss twenty times higher on the water training sets."""

    def __init__(self, torch, rng, sizes):
        self.W, self.Wb, self.b, self.mW, self.vW, self.mb, self.vb = [], [], [], [], [], [], []
        for (fi, fo) in sizes:
            w = (rng.uniform(-1.0, 1.0, size=(fi, fo)) / np.sqrt(fi)).astype(np.float32)
            self.W.append(torch.tensor(w, device="cuda"))
            self.Wb.append(torch.tensor(w, device="cuda").to(torch.bfloat16))
            self.b.append(torch.zeros(fo, dtype=torch.float32, device="cuda"))
            self.mW.append(torch.zeros_like(self.W[-1]))
            self.vW.append(torch.zeros_like(self.W[-1]))
            self.mb.append(torch.zeros_like(self.b[-1]))
            self.vb.append(torch.zeros_like(self.b[-1]))


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class LearnedGuess:
    """The pair-feature MLP, its feature statistics and its GPU training loop."""

    def __init__(self, seed=0, layout=WATER):
        self.torch = _torch()
        self.layout = layout
        self.rng = np.random.default_rng(seed)
        self.sizes = [(layout.n_in, WIDTH), (WIDTH, WIDTH), (WIDTH, layout.n_out)]
        self.dev = _Dev(self.torch, self.rng, self.sizes)
        self.mu = np.zeros(layout.n_feat)
        self.sd = np.ones(layout.n_feat)
        self.step = 0
        self.r_max = np.inf      # largest atom-pair distance of the training set

    # ---- data ----
    def _inputs(self, X):
        Z = np.zeros((X.shape[0], self.layout.n_in), dtype=np.float32)
        Z[:, :self.layout.n_feat] = (X - self.mu) / self.sd
        return Z

    def fit_stats(self, X):
        self.mu = X.mean(0)
        sd = X.std(0)
        self.sd = np.where(sd > 1e-8, sd, 1.0)

    # ---- forward / backward on the GPU ----
    def _forward(self, xb, n, keep=False):
        """xb: (n, N_IN) BF16 device tensor -> output FP32 (n, N_OUT) device tensor."""
        torch = self.torch
        L = _lib.load()
        d = self.dev
        acts_b = [xb]
        acts_f = []
        h = xb
        for li, (fi, fo) in enumerate(self.sizes):
            out = torch.empty((n, fo), dtype=torch.float32, device="cuda")
            last = li == len(self.sizes) - 1
            outb = None if last else torch.empty((n, fo), dtype=torch.bfloat16, device="cuda")
            _lib.check(L.qbx_gemm_bf16(0, 0, n, fo, fi, 1.0, _p(h), fi, _p(d.Wb[li]), fo, 0.0, _p(out), fo,
                                       2 if last else 1, _p(d.b[li]), _p(outb), None))
            acts_f.append(out)
            if not last:
                acts_b.append(outb)
                h = outb
        return (out, acts_b) if keep else out

    def train_step(self, xb, yt, rw, n, lr=1e-3):
        """One Adam step on a batch: xb (n, N_IN) BF16, yt (n, N_OUT) FP32 targets,
        rw (n,) FP32 row weights (1 real, 0 padding).  Returns the batch loss."""
        torch = self.torch
        L = _lib.load()
        d = self.dev
        y, acts_b = self._forward(xb, n, keep=True)
        nreal = float(rw.sum().item())
        dY = torch.empty_like(y)
        dYb = torch.empty((n, self.layout.n_out), dtype=torch.bfloat16, device="cuda")
        loss = torch.zeros(1, dtype=torch.float32, device="cuda")
        cmask = getattr(self, "_cmask", None)
        if cmask is None:
            cmask = self._cmask = torch.ones(self.layout.n_out, dtype=torch.float32, device="cuda")
        _lib.check(L.qbx_mlp_mse(n, self.layout.n_out, _p(y), _p(yt), _p(rw), _p(cmask), 1.0 / max(nreal, 1.0), _p(dY), _p(dYb),
                                 _p(loss), None))
        grads = self.backward(acts_b, dY, dYb, n)
        self.step += 1
        for li in range(len(self.sizes)):
            gW, gb = grads[li]
            _lib.check(L.qbx_mlp_adam(d.W[li].numel(), _p(d.W[li]), _p(d.mW[li]), _p(d.vW[li]), _p(gW), _p(d.Wb[li]),
                                      lr, 0.9, 0.999, 1e-8, self.step, None))
            _lib.check(L.qbx_mlp_adam(d.b[li].numel(), _p(d.b[li]), _p(d.mb[li]), _p(d.vb[li]), _p(gb), None,
                                      lr, 0.9, 0.999, 1e-8, self.step, None))
        return loss

    def backward(self, acts_b, dY, dYb, n):
        """Weight and bias gradients [(dW_l, db_l)] of the loss whose output
        gradient is dY (FP32) / dYb (BF16), given the forward BF16 activations
        acts_b = [x, h1, h2]."""
        torch = self.torch
        L = _lib.load()
        d = self.dev
        grads = [None] * len(self.sizes)
        g_f, g_b = dY, dYb
        for li in reversed(range(len(self.sizes))):
            fi, fo = self.sizes[li]
            gW = torch.empty((fi, fo), dtype=torch.float32, device="cuda")
            gbias = torch.empty(fo, dtype=torch.float32, device="cuda")
            # dW = h^T g, db = column sums of g
            _lib.check(L.qbx_gemm_bf16(1, 0, fi, fo, n, 1.0, _p(acts_b[li]), fi, _p(g_b), fo, 0.0, _p(gW), fo, 0,
                                       None, None, None))
            _lib.check(L.qbx_mlp_colsum(n, fo, _p(g_f), _p(gbias), None))
            grads[li] = (gW, gbias)
            if li > 0:
                # dh = g W^T, then through the ReLU of the layer below
                dh = torch.empty((n, fi), dtype=torch.float32, device="cuda")
                _lib.check(L.qbx_gemm_bf16(0, 1, n, fi, fo, 1.0, _p(g_b), fo, _p(d.Wb[li]), fo, 0.0, _p(dh), fi, 0,
                                           None, None, None))
                nf = torch.empty_like(dh)
                nb = torch.empty((n, fi), dtype=torch.bfloat16, device="cuda")
                _lib.check(L.qbx_mlp_elementwise(1, n * fi, _p(dh), _p(acts_b[li]), _p(nb), _p(nf), None))
                g_f, g_b = nf, nb
        return grads

    def predict(self, X):
        torch = self.torch
        xb = torch.tensor(self._inputs(X), device="cuda").to(torch.bfloat16)
        y = self._forward(xb, X.shape[0])
        torch.cuda.synchronize()
        return y.cpu().numpy().astype(np.float64)

    # ---- training on converged densities ----
    def train(self, data, epochs=200, batch=4096, lr=1e-3, log=None):
        """data: list of (X, Y) per system (raw features, padded density blocks)."""
        torch = self.torch
        X = np.concatenate([x for x, _ in data])
        Y = np.concatenate([y for _, y in data]).astype(np.float32)
        self.fit_stats(X)
        self.r_max = float(X[:, self.layout.dist_col].max())
        Z = self._inputs(X)
        n = len(Z)
        xb_all = torch.tensor(Z, device="cuda").to(torch.bfloat16)
        yt_all = torch.tensor(Y, device="cuda")
        hist = []
        for ep in range(epochs):
            perm = torch.tensor(self.rng.permutation(n), device="cuda")
            tot = torch.zeros(1, device="cuda")
            for s in range(0, n, batch):
                idx = perm[s:s + batch]
                m = len(idx)
                xb = xb_all[idx].contiguous()
                yt = yt_all[idx].contiguous()
                rw = torch.ones(m, dtype=torch.float32, device="cuda")
                tot += self.train_step(xb, yt, rw, m, lr) * m
            hist.append(float(tot.item()) / n)
            if log and (ep % max(1, epochs // 10) == 0 or ep == epochs - 1):
                log("epoch %d loss %.3e" % (ep, hist[-1]))
        return hist

    # ---- persistence ----
    def save(self, path, meta=None):
        """Weights (FP32 masters), feature statistics and layout to an .npz
        file (plain arrays; `meta` as a JSON string)."""
        import json
        arrs = {"W%d" % i: w.cpu().numpy() for i, w in enumerate(self.dev.W)}
        arrs.update({"b%d" % i: b.cpu().numpy() for i, b in enumerate(self.dev.b)})
        np.savez(path, mu=self.mu, sd=self.sd, r_max=np.array(self.r_max), nmax=np.array(self.layout.nmax),
                 elements=np.array(self.layout.elements), meta=np.array(json.dumps(meta or {})), **arrs)

    @classmethod
    def load(cls, path):
        """A network saved by `save` (read with allow_pickle=False)."""
        import json
        z = np.load(path, allow_pickle=False)
        layout = PairLayout(tuple(str(e) for e in z["elements"]), int(z["nmax"]))
        m = cls(seed=0, layout=layout)
        for i in range(len(m.sizes)):
            w = m.torch.tensor(z["W%d" % i], device="cuda")
            m.dev.W[i].copy_(w)
            m.dev.Wb[i].copy_(w.to(m.torch.bfloat16))
            m.dev.b[i].copy_(m.torch.tensor(z["b%d" % i], device="cuda"))
        m.mu, m.sd, m.r_max = z["mu"], z["sd"], float(z["r_max"])
        m.meta = json.loads(str(z["meta"]))
        return m

    # ---- guess ----
    def predict_density(self, nuc, gb, S=None, H=None):
        if S is None:
            S, H, blocks = system_data(nuc, gb, self.layout)
        else:
            blocks = atom_blocks(nuc, gb, self.layout)
        X = pair_features(nuc, blocks, S, H, self.layout)
        Y = self.predict(X)
        Y[X[:, self.layout.dist_col] > self.r_max] = 0.0
        return assemble_density(blocks, Y, gb.n, self.layout)

    def initial_coefficients(self, nuc, gb, S=None, H=None, schwarz=1e-12):
        """Coefficients from F(D_pred) = Hcore + J[D_pred] - K[D_pred]/2 (D_pred
        the predicted total density), diagonalized in the orthonormal basis."""
        if S is None:
            S, H, _ = system_data(nuc, gb, self.layout)
        Dt = self.predict_density(nuc, gb, S, H)
        Hcore = gb.core_hamiltonian(nuc.charges, nuc.coord_array)
        G = gb.jk(Dt, 0.5 * Dt, schwarz)
        X = gb.orthonormalization()
        return getC(X, Hcore + G), Dt


# The C/H/O network shipped with the package: trained by
# `python -m quiqbox_amd.train_guess --systems 4000 --min-atoms 8 --max-atoms 16
# --epochs 150 --batch 2048` on the GPU-converged RHF densities of random C/H/O
# clusters (cho_training_systems, seeds 100000-103999: 2016 of the 4000
# converged, 285,425 atom pairs); the file's `meta` holds the counts and times.
CHO_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "cho_sp_guess.npz")


def load_cho_model():
    """The packaged C/H/O network (CHO_SP layout) on the GPU."""
    return LearnedGuess.load(CHO_MODEL)


def guess_config(model, nuc, gb, strategy=None, maxStep=200, schwarz=1e-12, info=None):
    """HFconfig with the learned coefficients as OrbCoeffInitialConfig(:Direct),
    or with the core-Hamiltonian guess when the predicted density is not usable.
    `info` (a dict), if given, receives the predicted electron count and the choice."""
    strategy = strategy or SCFconfig(("ADIIS", "DIIS"), (1e-4, 1e-8))
    S, H, _ = system_data(nuc, gb, model.layout)
    C, Dt = model.initial_coefficients(nuc, gb, S, H, schwarz=schwarz)
    n_e = float(np.sum(nuc.charges))
    n_pred = float(np.sum(Dt * S))
    usable = bool(np.all(np.isfinite(C)) and abs(n_pred - n_e) <= 0.1 * n_e)
    if info is not None:
        info.update(electrons_predicted=n_pred, electrons=n_e, learned=usable)
    if not usable:
        return HFconfig(initial="CoreH", strategy=strategy, maxStep=maxStep)
    return HFconfig(initial=OrbCoeffInitialConfig(RCHartreeFock(), (C,)), strategy=strategy, maxStep=maxStep)


def training_set(systems, schwarz=1e-12, strategy=None, log=None, layout=WATER, maxStep=150, stats=None):
    """Converged GPU RHF densities of `systems` [(nuc, basis orbitals)] as
    (features, targets) per system; systems that do not converge are skipped.
    `stats` (a dict), if given, receives the counts and the SCF time."""
    out = []
    t0 = time.perf_counter()
    builds = 0
    for k, (nuc, bs) in enumerate(systems):
        gb = GPUBasis(bs)
        S, H, blocks = system_data(nuc, gb, layout)
        cfg = HFconfig(initial="CoreH", strategy=strategy or SCFconfig(("ADIIS", "DIIS"), (1e-4, 1e-9)),
                       maxStep=maxStep)
        res = runHartreeFock(nuc, gb, cfg, printInfo=False, schwarzThreshold=schwarz)
        builds += res.timing.get("fock_builds", 0)
        if not res.converged:
            continue
        Dt = 2.0 * res.density[0]
        out.append((pair_features(nuc, blocks, S, H, layout), pair_targets(blocks, Dt, layout)))
    t = time.perf_counter() - t0
    if log:
        log("training set: %d of %d systems converged, %.1f s" % (len(out), len(systems), t))
    if stats is not None:
        stats.update(systems=len(systems), converged=len(out), scf_s=t, fock_builds=builds,
                     pairs=int(sum(len(x) for x, _ in out)))
    return out


def cho_cluster(n_atoms, seed):
    """A closed-shell C/H/O cluster of the random s/p generator
    (systems.random_cluster: per atom one 3-primitive s and one p shell with
    random exponents and coefficients) at the 64-atom cluster's density: a
    cube of 20 (n/64)^(1/3) bohr, minimum separation 1.8 bohr."""
    from .systems import random_cluster
    return random_cluster(n_atoms, seed, box=20.0 * (n_atoms / 64.0) ** (1.0 / 3.0))


def cho_training_systems(count, seed0=100000, sizes=(8, 48)):
    """`count` clusters of cho_cluster with atom counts drawn uniformly from
    [sizes[0], sizes[1]] (seed seed0 + k for system k)."""
    rng = np.random.default_rng(seed0)
    ns = rng.integers(sizes[0], sizes[1] + 1, size=count)
    return [cho_cluster(int(n), seed0 + k) for k, n in enumerate(ns)]
