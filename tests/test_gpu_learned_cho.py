"""Learned initial guess on C/H/O clusters of the per-atom random s+p
generator (learned_guess.cho_cluster: one 3-primitive s and one p shell per
atom, exponents log-uniform in [0.1, 50], coefficients uniform in [0.1, 1]).

The pipeline runs end to end on the GPU: RHF densities of 8-, 12- and
16-atom clusters (20 seeds each), the pair network with the C/H/O layout
(CHO_SP) trained on the converged ones, and its guess on unseen 16- and
24-atom clusters next to the core-Hamiltonian guess.  These clusters do not
give a stable benchmark: about 40% of the small ones converge in 150 steps
(none at 48 atoms), they have several SCF solutions within a few 1e-6 Eh,
and their SCF paths are chaotic -- the rounding order of the J/K sums alone
changes which training clusters converge (23 and 26 of 60 in two runs) and
whether the SCF from the learned guess converges.  The test therefore
asserts what is stable (enough training clusters converge, the network
trains, its guess is finite and within 30% of the electron count on every
unseen cluster) and prints the SCF outcomes."""
import numpy as np
import pytest

import quiqbox_amd as qb
from quiqbox_amd import learned_guess as lg

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

HELD_OUT = [(16, 16010), (16, 16012), (16, 16013), (16, 16015), (24, 24010), (24, 24012)]


def _say(msg):
    print("[gpu-test] " + msg, flush=True)


def test_cho_learned_guess_on_unseen_clusters():
    train = [lg.cho_cluster(n, 1000 * n + s) for n in (8, 12, 16) for s in range(20)]
    stats = {}
    data = lg.training_set(train, layout=lg.CHO_SP, stats=stats)
    _say("training: %d of %d clusters converged, %d atom pairs, %.1f s" % (
        stats["converged"], stats["systems"], stats["pairs"], stats["scf_s"]))
    assert stats["converged"] >= 15
    model = lg.LearnedGuess(seed=0, layout=lg.CHO_SP)
    hist = model.train(data, epochs=300, batch=512, lr=1e-3)
    _say("loss %.3e -> %.3e" % (hist[0], hist[-1]))
    assert hist[-1] < 0.2 * hist[0]
    strat = qb.SCFconfig(("ADIIS", "DIIS"), (1e-4, 1e-8))
    outcome = []
    for n, seed in HELD_OUT:
        nuc, bs = lg.cho_cluster(n, seed)
        gb = qb.GPUBasis(bs)
        r0 = qb.runHartreeFock(nuc, gb, qb.HFconfig(initial="CoreH", strategy=strat, maxStep=150), printInfo=False)
        S, H, _ = lg.system_data(nuc, gb, model.layout)
        C, Dt = model.initial_coefficients(nuc, gb, S, H)
        assert np.all(np.isfinite(C)) and np.all(np.isfinite(Dt))
        n_e = float(np.sum(nuc.charges))
        n_pred = float(np.sum(Dt * S))
        assert abs(n_pred - n_e) < 0.3 * n_e
        info = {}
        r1 = qb.runHartreeFock(nuc, gb, lg.guess_config(model, nuc, gb, strat, 150, info=info), printInfo=False)
        outcome.append((r0.converged, r1.converged, info["learned"]))
        _say("%d atoms seed %d: core guess %s in %d builds; learned guess (%s, %.1f of %.0f electrons) %s in %d "
             "builds, dE %.2e" % (n, seed, "converged" if r0.converged else "not converged", r0.timing["fock_builds"],
                                  "used" if info["learned"] else "fallback", n_pred, n_e,
                                  "converged" if r1.converged else "not converged",
                                  r1.timing["fock_builds"], r1.energy[0] - r0.energy[0]))
    _say("converged (core guess, learned guess, learned used): %s" % outcome)


def test_packaged_cho_network_on_unseen_clusters():
    """The C/H/O network shipped with the package (quiqbox_amd/data/cho_sp_guess.npz,
    trained by quiqbox_amd.train_guess on 4000 generated 8-16-atom clusters) loads
    with its training record and gives a finite guess within 30% of the electron
    count on every unseen cluster; the SCF outcomes next to the core guess are
    printed (they are chaotic, as above)."""
    model = lg.load_cho_model()
    _say("packaged network: %s" % model.meta)
    assert model.layout.elements == lg.CHO_SP.elements and model.layout.nmax == lg.CHO_SP.nmax
    assert model.meta["systems"] == 4000 and model.meta["converged"] >= 1000
    # the loss after the first epoch is already 0.18 over 285k pairs; it levels off at 0.11
    assert model.meta["loss_last"] < 0.8 * model.meta["loss_first"] and model.meta["loss_last"] < 0.15
    strat = qb.SCFconfig(("ADIIS", "DIIS"), (1e-4, 1e-8))
    for n, seed in HELD_OUT:
        nuc, bs = lg.cho_cluster(n, seed)
        gb = qb.GPUBasis(bs)
        S, H, _ = lg.system_data(nuc, gb, model.layout)
        C, Dt = model.initial_coefficients(nuc, gb, S, H)
        assert np.all(np.isfinite(C)) and np.all(np.isfinite(Dt))
        n_e = float(np.sum(nuc.charges))
        n_pred = float(np.sum(Dt * S))
        assert abs(n_pred - n_e) < 0.3 * n_e
        info = {}
        r1 = qb.runHartreeFock(nuc, gb, lg.guess_config(model, nuc, gb, strat, 150, info=info), printInfo=False)
        assert np.isfinite(r1.energy[0])
        _say("%d atoms seed %d: packaged guess (%s, %.1f of %.0f electrons) %s in %d builds" % (
            n, seed, "used" if info["learned"] else "fallback", n_pred, n_e,
            "converged" if r1.converged else "not converged", r1.timing["fock_builds"]))
