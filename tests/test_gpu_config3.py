"""Fock builds at the measured ~600-function workload: the seeded 160-atom
C/H/O cluster with s/p/d shells (592 functions), in FP64 and in mixed
precision (FP32 primitive arithmetic below the Schwarz cut of 1e-5).

- Sampled elements of G = J[2D] - K[D] against the oracle's contracted
  quartets (every (mu nu|l s) and (mu l|s nu) of the sampled row pairs; a
  restatement of the reference's per-component Obara-Saika scheme),
  |dG| <= 1e-11 in FP64 and <= 1e-5 max|G| in mixed precision.
- Full-size properties: exact symmetry, linearity, run-to-run agreement,
  two density sets in one call (the UHF form, whose LDS K rows exceed 64 KB for
  the d bras and therefore run the one-wave-per-chunk kernels) against one set
  per call, the Schwarz threshold's effect, the mixed-precision bound.

The density is the core-Hamiltonian guess of the cluster (HartreeFock.jl:215-226).
"""
import os

import numpy as np
import pytest

import oracle
import quiqbox_amd as qb
from quiqbox_amd.hartree_fock import getC, getD
from quiqbox_amd.systems import random_cluster

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def _say(msg):
    print("[gpu-test] " + msg, flush=True)


_CACHE = {}


def _cluster():
    if "c" not in _CACHE:
        nuc, bs = random_cluster(160, 0, with_d=True)
        gb = qb.GPUBasis(bs)
        S = gb.overlap()
        H = gb.core_hamiltonian(nuc.charges, nuc.coord_array)
        X = qb.getOrthonormalization(S)
        nocc = int(round(sum(nuc.charges))) // 2
        D = getD(getC(X, H), nocc)
        _CACHE["c"] = (nuc, gb, D)
    return _CACHE["c"]


def _function_L(flat, i):
    return int(np.sum(flat.prim_ang[flat.bf_prim[flat.bf_ptr[i]]]))


def _oracle_G(ob, n, D, mu, nu):
    ls = np.array([(l, s) for l in range(n) for s in range(n)], dtype=np.int32)
    qJ = np.column_stack([np.full(len(ls), mu), np.full(len(ls), nu), ls[:, 0], ls[:, 1]]).astype(np.int32)
    qK = np.column_stack([np.full(len(ls), mu), ls[:, 0], ls[:, 1], np.full(len(ls), nu)]).astype(np.int32)
    J = ob.eri_list(qJ, threads=THREADS).reshape(n, n)
    K = ob.eri_list(qK, threads=THREADS).reshape(n, n)
    return float(np.sum(2 * D.T * J) - np.sum(D * K))


def test_config3_size_and_classes():
    nuc, gb, D = _cluster()
    info = gb.info()
    _say("config 3: %d atoms, %d functions, %d shells, max L %d" % (len(nuc.charges), gb.n, info["n_shell"],
                                                                    info["max_l"]))
    assert gb.n == 592 and info["max_l"] == 2
    st = gb.jk_stats(1e-12)
    assert st["shell_quartets"].sum() > 0
    # every one of the 21 s/p/d classes has screened quartets at this size
    assert np.all(st["shell_quartets"] > 0)


def test_config3_fp64_and_mixed_sampled_against_oracle():
    nuc, gb, D = _cluster()
    n = gb.n
    flat = gb.flat
    Ls = [_function_L(flat, i) for i in range(n)]
    d_fn = [i for i in range(n) if Ls[i] == 2]
    p_fn = [i for i in range(n) if Ls[i] == 1]
    s_fn = [i for i in range(n) if Ls[i] == 0]
    pairs = [(s_fn[0], s_fn[0]), (d_fn[3], p_fn[40]), (d_fn[-1], d_fn[-4])]
    _say("config 3 GPU builds: FP64 and mixed, Schwarz 0")
    G64 = gb.jk(2 * D, D, 0.0)
    Gmx = gb.jk(2 * D, D, 0.0, qb.QBX_PREC_MIXED)
    gmax = float(np.max(np.abs(G64)))
    ob = oracle.Basis(flat.as_dict())
    for mu, nu in pairs:
        g = _oracle_G(ob, n, D, mu, nu)
        _say("  G[%d,%d] (L %d,%d): fp64 %.14e mixed %.14e oracle %.14e" % (
            mu, nu, Ls[mu], Ls[nu], G64[mu, nu], Gmx[mu, nu], g))
        assert abs(G64[mu, nu] - g) < 1e-11
        assert abs(Gmx[mu, nu] - g) < 1e-5 * gmax


def test_config3_full_size_properties():
    nuc, gb, D = _cluster()
    n = gb.n
    rng = np.random.default_rng(11)
    A = rng.standard_normal((n, n))
    D2 = 0.01 * (A + A.T)
    _say("config 3 properties: FP64 builds at Schwarz 1e-12")
    G1 = gb.jk(2 * D, D, 1e-12)
    G2 = gb.jk(2 * D2, D2, 1e-12)
    scale = float(np.max(np.abs(G1)))
    # exact symmetry of the symmetrized output
    assert np.array_equal(G1, G1.T)
    # run to run: the J/K atomics sum in varying order
    G1b = gb.jk(2 * D, D, 1e-12)
    err_rep = float(np.max(np.abs(G1b - G1)))
    _say("  repeat max|dG| %.3e (max|G| %.3e)" % (err_rep, scale))
    assert err_rep < 1e-12 * max(1.0, scale)
    # linearity
    G12 = gb.jk(2 * (D + 0.5 * D2), D + 0.5 * D2, 1e-12)
    err_lin = float(np.max(np.abs(G12 - (G1 + 0.5 * G2))))
    _say("  linearity max|dG| %.3e" % err_lin)
    assert err_lin < 1e-10 * max(1.0, scale)
    # two density sets in one call (UHF form) against one set per call
    DJ = np.stack([2 * D, D + D2])
    DK = np.stack([D, D2])
    Gs = gb.jk(DJ, DK, 1e-12)
    for s in range(2):
        Gref = gb.jk(DJ[s], DK[s], 1e-12)
        err = float(np.max(np.abs(Gs[s] - Gref)))
        _say("  n_set 2, set %d vs single-set call max|dG| %.3e" % (s, err))
        assert err < 1e-12 * max(1.0, float(np.max(np.abs(Gref))))
    # screening at 1e-12 against none
    G0 = gb.jk(2 * D, D, 0.0)
    err_scr = float(np.max(np.abs(G0 - G1)))
    _say("  Schwarz 1e-12 vs 0 max|dG| %.3e" % err_scr)
    assert err_scr < 1e-9
    # mixed precision against FP64 at full size
    Gm = gb.jk(2 * D, D, 1e-12, qb.QBX_PREC_MIXED)
    err_mx = float(np.max(np.abs(Gm - G1)))
    _say("  mixed vs FP64 max|dG| %.3e (bound %.3e)" % (err_mx, 1e-5 * scale))
    assert err_mx < 1e-5 * scale
    assert np.array_equal(Gm, Gm.T)
    # the mixed build with every quartet above the cut is the FP64 build
    gb.ctx.set_mixed_cut(0.0)
    try:
        Gm64 = gb.jk(2 * D, D, 1e-12, qb.QBX_PREC_MIXED)
    finally:
        gb.ctx.set_mixed_cut(1e-5)
    assert float(np.max(np.abs(Gm64 - G1))) < 1e-12 * max(1.0, scale)
