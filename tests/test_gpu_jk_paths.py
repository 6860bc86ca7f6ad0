"""J/K kernel paths that the default build reaches only at some sizes, and the
ordering of builds across streams.

- QBX_JK_ROWS=0 (fresh context) turns the LDS K-row and twin kernels off, so
  every class runs the one-wave-per-chunk J/K kernel (the path bases take when
  a bra's K rows exceed 64 KB of LDS, e.g. d bras above ~680 functions, or two
  density sets at ~600).  Checked against the oracle's dense getGcore on a water
  dimer in cc-pVDZ (ket buckets on) and on a small s/p/d cluster whose staged
  classes then take their per-wave kernel: one and two density sets, FP64 and
  mixed precision.
- QBX_STAGED_SPLIT=0 (fresh context) gives (dd|dd), (dd|dp) and (dp|dp)
  back their one-wave-per-chunk staged kernels (by default the waves of a
  block share each chunk); with QBX_JK_ROWS=0 those classes take the shared
  chunks with K straight to global memory.  Both against the oracle.
- A build enqueued on a caller's stream (qbx_jk_device), followed by a build
  with another Schwarz threshold and precision on the context stream (which
  rewrites the task lists): both against the oracle.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle
import quiqbox_amd as qb
from helpers import OracleEngine, random_spd_basis
from quiqbox_amd.engine import Context
from quiqbox_amd.systems import water_box

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[gpu-test] " + msg, flush=True)


def _sym(rng, n, scale=0.05):
    A = rng.standard_normal((n, n))
    return scale * (A + A.T)


def _context(**env):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return Context(0)
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


@pytest.mark.parametrize("system", ["water_dimer_ccpvdz", "spd_cluster"])
def test_one_wave_per_chunk_jk_against_oracle(system):
    if system == "water_dimer_ccpvdz":
        _, bs = water_box(2, 7, basis="cc-pVDZ")
    else:
        _, bs = random_spd_basis(31, n_atoms=4)
    ctx = _context(QBX_JK_ROWS="0")
    gb = qb.GPUBasis(bs, ctx=ctx)
    E = OracleEngine(gb.flat).eri()
    n = gb.n
    rng = np.random.default_rng(3)
    DJ = np.stack([_sym(rng, n), _sym(rng, n)])
    DK = np.stack([_sym(rng, n), _sym(rng, n)])
    Gref = [oracle.gcore(E, DJ[s], DK[s]) for s in range(2)]
    gmax = max(float(np.max(np.abs(g))) for g in Gref)
    _say("%s (%d functions), LDS K rows off: one and two density sets, FP64 and mixed" % (system, n))
    G1 = gb.jk(DJ[0], DK[0], 0.0)
    err1 = float(np.max(np.abs(G1 - Gref[0])))
    G2 = gb.jk(DJ, DK, 0.0)
    err2 = max(float(np.max(np.abs(G2[s] - Gref[s]))) for s in range(2))
    Gm = gb.jk(DJ, DK, 0.0, qb.QBX_PREC_MIXED)
    errm = max(float(np.max(np.abs(Gm[s] - Gref[s]))) for s in range(2))
    _say("  max|G - G_oracle|: n_set 1 %.3e, n_set 2 %.3e, mixed %.3e (max|G| %.3e)" % (err1, err2, errm, gmax))
    assert err1 < 1e-11 and err2 < 1e-11
    assert errm < 1e-5 * max(1.0, gmax)
    # the default context (rows on) gives the same matrices
    gd = qb.GPUBasis(bs)
    Gd = gd.jk(DJ, DK, 0.0)
    assert max(float(np.max(np.abs(Gd[s] - G2[s]))) for s in range(2)) < 1e-12 * max(1.0, gmax)
    del gb
    ctx.close()


@pytest.mark.parametrize("env", [{"QBX_STAGED_SPLIT": "0"}, {"QBX_STAGED_SPLIT": "0", "QBX_JK_ROWS": "0"}])
def test_staged_kernels_without_shared_chunks(env):
    _, bs = random_spd_basis(31, n_atoms=4)
    ctx = _context(**env)
    gb = qb.GPUBasis(bs, ctx=ctx)
    E = OracleEngine(gb.flat).eri()
    n = gb.n
    rng = np.random.default_rng(5)
    DJ = np.stack([_sym(rng, n), _sym(rng, n)])
    DK = np.stack([_sym(rng, n), _sym(rng, n)])
    Gref = [oracle.gcore(E, DJ[s], DK[s]) for s in range(2)]
    G = gb.jk(DJ, DK, 0.0)
    err = max(float(np.max(np.abs(G[s] - Gref[s]))) for s in range(2))
    _say("spd cluster (%d functions), %s: max|G - G_oracle| %.3e" % (n, env, err))
    assert err < 1e-11
    del gb
    ctx.close()


def test_device_build_on_caller_stream_then_context_stream():
    import torch
    _, bs = random_spd_basis(12, n_atoms=3)
    gb = qb.GPUBasis(bs)
    E = OracleEngine(gb.flat).eri()
    n = gb.n
    rng = np.random.default_rng(9)
    D = _sym(rng, n)
    Gref = oracle.gcore(E, 2 * D, D)
    dev = torch.device("cuda", 0)
    dDJ = torch.tensor(2 * D, dtype=torch.float64, device=dev)
    dDK = torch.tensor(D, dtype=torch.float64, device=dev)
    dG = torch.zeros((n, n), dtype=torch.float64, device=dev)
    side = torch.cuda.Stream(device=dev)
    torch.cuda.synchronize()
    _say("device build on a caller stream (n %d), then a host build at another threshold and precision" % n)
    gb.jk_device(dDJ.data_ptr(), dDK.data_ptr(), dG.data_ptr(), 1, 1e-12, qb.QBX_PREC_FP64,
                 stream=ctypes.c_void_p(side.cuda_stream).value)
    # no wait here: the next call rewrites the task lists for (thr 0, mixed)
    Gm = gb.jk(2 * D, D, 0.0, qb.QBX_PREC_MIXED)
    side.synchronize()
    Gdev = dG.cpu().numpy()
    err_dev = float(np.max(np.abs(Gdev - Gref)))
    err_mx = float(np.max(np.abs(Gm - Gref)))
    _say("  caller-stream build max|dG| %.3e, mixed host build %.3e" % (err_dev, err_mx))
    assert err_dev < 1e-11
    assert err_mx < 1e-5 * max(1.0, float(np.max(np.abs(Gref))))
    # and back: a device build on the caller stream right after the host build
    dG.zero_()
    torch.cuda.synchronize()
    gb.jk_device(dDJ.data_ptr(), dDK.data_ptr(), dG.data_ptr(), 1, 0.0, qb.QBX_PREC_FP64,
                 stream=ctypes.c_void_p(side.cuda_stream).value)
    G64 = gb.jk(2 * D, D, 1e-12)
    side.synchronize()
    assert float(np.max(np.abs(dG.cpu().numpy() - Gref))) < 1e-11
    assert float(np.max(np.abs(G64 - Gref))) < 1e-11
