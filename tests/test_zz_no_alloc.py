"""qbx_density_purify_noalloc (device arrays, work matrices in the library's
fixed device array) against qbx_density_purify (host arrays, work buffers the
context allocates): two calls in a row, the second with a smaller system, each
compared with the allocating entry point on the same input; and the size
limit rejected before any work."""
import ctypes

import numpy as np
import pytest

import quiqbox_amd as qb
from helpers import molecule_basis, literals
from quiqbox_amd import _lib
from quiqbox_amd.systems import water_box

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _say(msg):
    print("[gpu-test] " + msg, flush=True)


def _core_system(nuc, bs):
    gb = qb.GPUBasis(bs)
    X = np.ascontiguousarray(gb.orthonormalization())
    F = np.ascontiguousarray(gb.core_hamiltonian(nuc.charges, nuc.coord_array))
    nocc = int(round(sum(nuc.charges))) // 2
    return X, F, nocc


def _noalloc(ctx, X, F, nocc, stream=None):
    n = X.shape[0]
    dev = torch.device("cuda", 0)
    dX = torch.tensor(X, dtype=torch.float64, device=dev)
    dF = torch.tensor(F, dtype=torch.float64, device=dev)
    dD = torch.full((n, n), float("nan"), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    it = ctypes.c_int32()
    idem = ctypes.c_double()
    s = ctypes.c_void_p(stream.cuda_stream if stream is not None else 0)
    with ctx.lock:
        rc = _lib.load().qbx_density_purify_noalloc(ctx.h, n, nocc, ctypes.c_void_p(dX.data_ptr()),
                                                    ctypes.c_void_p(dF.data_ptr()), ctypes.c_void_p(dD.data_ptr()),
                                                    100, ctypes.byref(it), ctypes.byref(idem), s)
    _lib.check(rc)
    return dD.cpu().numpy(), it.value


def test_noalloc_purification_matches_the_allocating_entry_point():
    ctx = qb.default_context()
    nuc1, bs1 = water_box(24, 0, basis="cc-pVDZ")
    L = literals()["h2o2_631g"]
    coords = [tuple(L["angstrom_to_bohr"] * x for x in c) for c in L["coords_angstrom"]]
    nuc2 = qb.NuclearCluster(L["nucs"], coords)
    bs2 = molecule_basis(nuc2.symbols, nuc2.coords, "6-31G")
    side = torch.cuda.Stream()
    for k, (nuc, bs, stream) in enumerate(((nuc1, bs1, None), (nuc2, bs2, side))):
        X, F, nocc = _core_system(nuc, bs)
        n = X.shape[0]
        D_ref, it_ref, _ = ctx.purified_density(X, F, nocc)
        D, it = _noalloc(ctx, X, F, nocc, stream)
        err = float(np.max(np.abs(D - D_ref)))
        _say("call %d: n %d, nocc %d: %d products (allocating entry point %d), max|D - D_ref| %.3e" % (
            k + 1, n, nocc, it, it_ref, err))
        assert np.all(np.isfinite(D))
        assert it == it_ref
        assert err <= 1e-14 * max(1.0, float(np.max(np.abs(D_ref))))


def test_noalloc_rejects_sizes_above_its_array():
    ctx = qb.default_context()
    n = 1025
    rc = _lib.load().qbx_density_purify_noalloc(ctx.h, n, 1, ctypes.c_void_p(16), ctypes.c_void_p(16),
                                                ctypes.c_void_p(16), 100, None, None, None)
    assert rc == -1
    buf = ctypes.create_string_buffer(512)
    _lib.load().qbx_last_error(buf, 512)
    assert b"1024" in buf.value
