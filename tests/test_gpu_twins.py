"""Twin bras in the Fock build: cc-pVDZ writes oxygen's 1s and 2s as two
segmented shells over one set of 9 exponents, so their shell pairs share all
primitive geometry and the J/K kernels evaluate them as one group (one Boys
evaluation per primitive quartet, one accumulator set per member).  Ket
buckets (kets ordered by primitive-pair count, one range per bucket and bra)
are checked the same way against the plain Schwarz order (QBX_KET_BUCKETS=0):
the screened quartet counts must agree exactly and G to 1e-12.  Checked
against the oracle's dense getGcore on one water molecule, and against the same
build with the groups switched off (QBX_TWINS=0 in a fresh context) on a water
dimer (cross-molecule groups of four), for RHF, two density sets (UHF) and the
mixed-precision task split, and through a converged SCF.  A water dimer with
everything on (ket buckets, twin bras, the staged classes' plain-order ket
lists) is checked against the oracle's dense getGcore as well."""
import os

import numpy as np
import pytest

import oracle
import quiqbox_amd as qb
from quiqbox_amd import _lib
from quiqbox_amd.engine import Context
from quiqbox_amd.systems import water_box

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[gpu-test] " + msg, flush=True)


def _density(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    return 0.05 * (A + A.T)


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


def _no_twin_context():
    return _context(QBX_TWINS="0")


def test_single_water_ccpvdz_against_oracle():
    nuc, bs = water_box(1, 3, basis="cc-pVDZ")
    gb = qb.GPUBasis(bs)
    ob = oracle.Basis(gb.flat.as_dict())
    D = _density(gb.n, 1)
    _say("water cc-pVDZ (%d functions): GPU J/K with twin groups vs oracle getGcore" % gb.n)
    G = gb.jk(2 * D, D, 0.0)
    Gref = oracle.gcore(ob.eri_dense(), 2 * D, D)
    err = float(np.abs(G - Gref).max())
    _say("  max |G - G_oracle| = %.3e" % err)
    assert err < 1e-11


@pytest.mark.parametrize("nset,precision", [(1, _lib.QBX_PREC_FP64), (2, _lib.QBX_PREC_FP64),
                                            (1, _lib.QBX_PREC_MIXED)])
def test_water_dimer_twins_on_off(nset, precision):
    nuc, bs = water_box(2, 11, basis="cc-pVDZ")
    on = qb.GPUBasis(bs)
    off = qb.GPUBasis(bs, ctx=_no_twin_context())
    Ds = [_density(on.n, 2 + k) for k in range(nset)]
    if nset == 1:
        DJ, DK = 2 * Ds[0], Ds[0]
    else:
        DJ = np.stack([Ds[0] + Ds[1]] * 2)
        DK = np.stack(Ds)
    _say("water dimer cc-pVDZ: nset %d precision %d, twin groups on vs off" % (nset, precision))
    Gon = on.jk(DJ, DK, 1e-12, precision)
    Goff = off.jk(DJ, DK, 1e-12, precision)
    err = float(np.abs(Gon - Goff).max())
    _say("  max |G_on - G_off| = %.3e (max |G| %.2e)" % (err, float(np.abs(Goff).max())))
    assert err < 1e-12 * max(1.0, float(np.abs(Goff).max()))


def test_water_trimer_scf_twins_on_off():
    nuc, bs = water_box(3, 21, basis="cc-pVDZ")
    cfg = qb.HFconfig(initial="CoreH", strategy=qb.SCFconfig(("ADIIS", "DIIS"), (1e-4, 1e-10)))
    _say("water trimer cc-pVDZ RHF, twin groups on vs off")
    r_on = qb.runHartreeFock(nuc, qb.GPUBasis(bs), cfg, printInfo=False)
    r_off = qb.runHartreeFock(nuc, qb.GPUBasis(bs, ctx=_no_twin_context()), cfg, printInfo=False)
    _say("  E on %.12f off %.12f" % (r_on.energy[0], r_off.energy[0]))
    assert r_on.converged and r_off.converged
    assert abs(r_on.energy[0] - r_off.energy[0]) < 1e-10


@pytest.mark.parametrize("nset,precision,screen", [(1, _lib.QBX_PREC_FP64, 0.0), (2, _lib.QBX_PREC_FP64, 0.0),
                                                   (1, _lib.QBX_PREC_MIXED, 0.0), (1, _lib.QBX_PREC_FP64, 1e-11)])
def test_water_dimer_ket_buckets_on_off(nset, precision, screen):
    nuc, bs = water_box(2, 11, basis="cc-pVDZ")
    on_ctx, off_ctx = Context(0), _context(QBX_KET_BUCKETS="0")
    for c in (on_ctx, off_ctx):
        c.set_density_screening(screen)
    on = qb.GPUBasis(bs, ctx=on_ctx)
    off = qb.GPUBasis(bs, ctx=off_ctx)
    Ds = [_density(on.n, 5 + k) for k in range(nset)]
    if nset == 1:
        DJ, DK = 2 * Ds[0], Ds[0]
    else:
        DJ = np.stack([Ds[0] + Ds[1]] * 2)
        DK = np.stack(Ds)
    _say("water dimer cc-pVDZ: nset %d precision %d screen %g, ket buckets on vs off" % (nset, precision, screen))
    for thr in (1e-12, 1e-9):
        s_on, s_off = on.jk_stats(thr), off.jk_stats(thr)
        for key in s_on:
            assert np.array_equal(np.asarray(s_on[key]), np.asarray(s_off[key])), key
        Gon = on.jk(DJ, DK, thr, precision)
        Goff = off.jk(DJ, DK, thr, precision)
        err = float(np.abs(Gon - Goff).max())
        _say("  Schwarz %g: max |G_on - G_off| = %.3e (max |G| %.2e)" % (thr, err, float(np.abs(Goff).max())))
        assert err < 1e-12 * max(1.0, float(np.abs(Goff).max()))


def test_dense_eri_ket_buckets_on_off():
    """The dense tensor (every quartet, no screening) through the bucketed task
    lists equals the plain order's, on a water dimer in cc-pVDZ (contraction
    lengths 1, 3 and 9 within one pair class)."""
    nuc, bs = water_box(2, 3, basis="cc-pVDZ")
    on = qb.GPUBasis(bs)
    off = qb.GPUBasis(bs, ctx=_context(QBX_KET_BUCKETS="0"))
    E_on, E_off = on.eri_dense(), off.eri_dense()
    err = float(np.abs(E_on - E_off).max())
    _say("water dimer cc-pVDZ (%d functions): max |ERI_on - ERI_off| = %.3e" % (on.n, err))
    assert err < 1e-13


def test_water_dimer_ccpvdz_against_oracle():
    """Ket buckets, twin bras and the staged classes' plain-order ket lists all
    active (50 functions, d shells on oxygen) against the oracle's dense
    getGcore, RHF and one UHF pair."""
    nuc, bs = water_box(2, 11, basis="cc-pVDZ")
    gb = qb.GPUBasis(bs)
    ob = oracle.Basis(gb.flat.as_dict())
    E = ob.eri_dense()
    Da, Db = _density(gb.n, 21), _density(gb.n, 22)
    _say("water dimer cc-pVDZ (%d functions): GPU J/K vs oracle getGcore" % gb.n)
    G = gb.jk(2 * Da, Da, 0.0)
    err = float(np.abs(G - oracle.gcore(E, 2 * Da, Da)).max())
    _say("  RHF max |G - G_oracle| = %.3e" % err)
    assert err < 1e-11
    Gu = gb.jk(np.stack([Da + Db] * 2), np.stack([Da, Db]), 0.0)
    for s, Ds in enumerate((Da, Db)):
        err = float(np.abs(Gu[s] - oracle.gcore(E, Da + Db, Ds)).max())
        _say("  UHF set %d max |G - G_oracle| = %.3e" % (s, err))
        assert err < 1e-11
