"""The reference's own SCF-driver pins, run on the GPU engine.

- H2O2 in 6-31G, all 13 configurations of HartreeFock-test.jl:292-353: the
  default HFconfig, and SAD/CoreH/GWH guesses each with the default strategy
  and with DIIS, EDIIS and ADIIS alone (threshold 5e-10, secondary ratio 5):
  |E - E_ref| <= 5 t1 and converged, as the reference asserts.
- H2 in 3-21G over the 100 bond lengths 0.1, 0.3, .., 19.9 bohr
  (HartreeFock-test.jl:220-289), RHF and UHF, CoreH and SAD guesses, maxStep
  300: every energy within the reference's comparator (7.5e-7 Eh).
"""
import numpy as np
import pytest

import quiqbox_amd as qb
from helpers import literals, molecule_basis

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[gpu-test] " + msg, flush=True)


def test_h2o2_631g_all_thirteen_configurations():
    L = literals()["h2o2_631g"]
    coords = [tuple(L["angstrom_to_bohr"] * x for x in c) for c in L["coords_angstrom"]]
    nuc = qb.NuclearCluster(L["nucs"], coords)
    t1 = L["threshold"]
    ratio = tuple(L["secondaryConvRatio"])
    bs = molecule_basis(nuc.symbols, nuc.coords, "6-31G")
    gb = qb.GPUBasis(bs)
    scf1 = qb.SCFconfig(threshold=t1, secondaryConvRatio=ratio)
    scf2 = qb.SCFconfig(("DIIS",), (t1,), secondaryConvRatio=ratio)
    scf3 = qb.SCFconfig(("EDIIS",), (t1,), secondaryConvRatio=ratio)
    scf4 = qb.SCFconfig(("ADIIS",), (t1,), secondaryConvRatio=ratio)
    cfgs = [("default", qb.HFconfig())]
    for scf_name, scf in (("default strategy", scf1), ("DIIS", scf2), ("EDIIS", scf3), ("ADIIS", scf4)):
        for guess in ("SAD", "CoreH", "GWH"):
            cfgs.append(("%s + %s" % (guess, scf_name), qb.HFconfig(initial=guess, strategy=scf, maxStep=200)))
    assert len(cfgs) == 13
    failed = []
    for name, cfg in cfgs:
        res = qb.runHartreeFock(nuc, gb, cfg, printInfo=False)
        err = res.energy[0] - L["E_elec"]
        _say("H2O2 6-31G %-26s E %.10f  dE %.2e  converged %s" % (name, res.energy[0], err, res.converged))
        if not (abs(err) <= 5 * t1 and res.converged):
            failed.append(name)
    assert not failed, "configurations off the reference energy or not converged: %s" % failed


@pytest.mark.parametrize("guess", ["CoreH", "SAD"])
def test_h2_321g_full_potential_curve(guess):
    L = literals()["h2_321g_curve"]
    tol = 10 * 7.5e-8
    worst = {"rhf": 0.0, "uhf": 0.0}
    bad = []
    for k, R in enumerate(L["bond_lengths"]):
        coords = [(0.0, 0.0, 0.0), (R, 0.0, 0.0)]
        nuc = qb.NuclearCluster(["H", "H"], coords)
        gb = qb.GPUBasis(molecule_basis(["H", "H"], coords, "3-21G"))
        for HFT, key in ((qb.RCHartreeFock(), "rhf"), (qb.UOHartreeFock(), "uhf")):
            res = qb.runHartreeFock(nuc, gb, qb.HFconfig(HFT, initial=guess, maxStep=300), printInfo=False)
            e = sum(res.energy) - L[key][k]
            worst[key] = max(worst[key], abs(e))
            if abs(e) > tol:
                bad.append((key, R, e))
        if k % 20 == 19:
            _say("H2 3-21G %s: %d of 100 bond lengths, worst |dE| RHF %.2e UHF %.2e" % (
                guess, k + 1, worst["rhf"], worst["uhf"]))
    assert not bad, "points off the reference curve: %s" % bad[:10]
