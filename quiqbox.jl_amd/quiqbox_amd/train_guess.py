"""Train the learned initial guess on GPU-converged densities of random C/H/O
clusters (the s/p cluster generator, 8-48 atoms) and save the network.

    python -m quiqbox_amd.train_guess --systems 2000 --out model.npz
    python -m quiqbox_amd.train_guess --load model.npz --eval 10 --sad

Every system runs a GPU RHF (core guess, ADIIS then DIIS to 1e-9 Eh); the
converged ones supply the (pair features, density block) training pairs, and
the network trains on them with the library's BF16 matrix-core kernels.  The
saved file records how many systems converged, the SCF and training times and
the loss history (`LearnedGuess.load(path).meta`).

With --eval N the network (trained here or loaded) starts RHF (ADIIS then
DIIS to 1e-8 Eh) on N unseen clusters per size of --eval-sizes, next to the
core-Hamiltonian guess and, with --sad, Quiqbox's SAD guess (one UHF per
atom, HartreeFock.jl:256-293); each run's convergence, Fock builds, wall time
(guess included) and energy are printed, then the totals.
"""
import argparse
import sys
import time

import numpy as np

from . import learned_guess as lg


def log(msg):
    print("[train_guess] " + msg, flush=True)


def build_training_data(count, seed0, sizes, chunk=100):
    data, stats = [], {"systems": 0, "converged": 0, "scf_s": 0.0, "fock_builds": 0, "pairs": 0}
    t0 = time.perf_counter()
    for start in range(0, count, chunk):
        systems = lg.cho_training_systems(min(chunk, count - start), seed0 + start, sizes)
        st = {}
        data += lg.training_set(systems, layout=lg.CHO_SP, stats=st)
        for k in ("systems", "converged", "fock_builds", "pairs"):
            stats[k] += st[k]
        stats["scf_s"] += st["scf_s"]
        log("%d of %d systems, %d converged, %.0f s" % (stats["systems"], count, stats["converged"],
                                                      time.perf_counter() - t0))
    return data, stats


def _run(nuc, gb, initial, strategy, max_step, model=None):
    from .hartree_fock import HFconfig, runHartreeFock
    t0 = time.perf_counter()
    info = {}
    if model is None:
        cfg = HFconfig(initial=initial, strategy=strategy, maxStep=max_step)
    else:
        cfg = lg.guess_config(model, nuc, gb, strategy, max_step, info=info)
    res = runHartreeFock(nuc, gb, cfg, printInfo=False)
    return dict(converged=bool(res.converged), builds=int(res.timing.get("fock_builds") or 0),
                wall_s=time.perf_counter() - t0, energy=float(res.energy[0]), learned=info.get("learned"))


def evaluate_clusters(model, clusters, sad=False, max_step=150, say=None):
    """RHF (ADIIS then DIIS to 1e-8 Eh, at most max_step steps) of every
    cho_cluster(n, seed) in `clusters` from the core guess, the learned guess
    and optionally SAD.  Returns one record per cluster: per guess the
    convergence, Fock builds, wall time (guess included) and energy."""
    from .engine import GPUBasis
    from .hartree_fock import SCFconfig
    say = say or log
    strategy = SCFconfig(("ADIIS", "DIIS"), (1e-4, 1e-8))
    guesses = ["CoreH", "learned"] + (["SAD"] if sad else [])
    rows = []
    for n, seed in clusters:
        nuc, bs = lg.cho_cluster(n, seed)
        gb = GPUBasis(bs)
        row = dict(atoms=n, seed=seed, nbf=gb.n)
        for g in guesses:
            row[g] = _run(nuc, gb, g, strategy, max_step, model if g == "learned" else None)
        rows.append(row)
        say("%d atoms seed %d (%d functions): " % (n, seed, gb.n) + "; ".join(
            "%s %s %d builds %.2f s E %.9f%s" % (
                g, "conv" if row[g]["converged"] else "NOT conv", row[g]["builds"], row[g]["wall_s"],
                row[g]["energy"], "" if g != "learned" else (" (used)" if row[g]["learned"] else " (fallback)"))
            for g in guesses))
    return rows


def summarize(rows, guesses):
    """Per guess: clusters converged, and the Fock builds and wall time summed
    over the clusters on which every guess converged."""
    both = [r for r in rows if all(r[g]["converged"] for g in guesses)]
    out = {"clusters": len(rows), "all_converged": len(both)}
    for g in guesses:
        out[g] = {"converged": sum(r[g]["converged"] for r in rows),
                  "builds_all_converged": sum(r[g]["builds"] for r in both),
                  "wall_s_all_converged": sum(r[g]["wall_s"] for r in both)}
    return out


def evaluate(model, sizes, count, seed0, sad=False, max_step=150):
    """evaluate_clusters on cho_cluster(n, seed0 + 1000 n + k), k < count, for
    every n in sizes, with a summary per size."""
    guesses = ["CoreH", "learned"] + (["SAD"] if sad else [])
    rows = evaluate_clusters(model, [(n, seed0 + 1000 * n + k) for n in sizes for k in range(count)], sad, max_step)
    for n in sizes:
        sm = summarize([r for r in rows if r["atoms"] == n], guesses)
        log("%d atoms: %s; totals over the %d clusters on which every guess converged: %s" % (
            n, ", ".join("%s converged %d/%d" % (g, sm[g]["converged"], sm["clusters"]) for g in guesses),
            sm["all_converged"], ", ".join("%s %d builds %.1f s" % (g, sm[g]["builds_all_converged"],
                                                                   sm[g]["wall_s_all_converged"])
                                           for g in guesses)))
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--systems", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=100000)
    ap.add_argument("--min-atoms", type=int, default=8)
    ap.add_argument("--max-atoms", type=int, default=48)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=None)
    ap.add_argument("--load", default=None, help="evaluate a saved network instead of training one")
    ap.add_argument("--eval", type=int, default=0, help="unseen clusters per size to evaluate on")
    ap.add_argument("--eval-sizes", default="16,24")
    ap.add_argument("--eval-seed", type=int, default=900000)
    ap.add_argument("--sad", action="store_true", help="also run the SAD guess in the evaluation")
    args = ap.parse_args(argv)
    sizes = [int(x) for x in args.eval_sizes.split(",") if x]
    if args.load:
        model = lg.LearnedGuess.load(args.load)
        log("loaded %s: %s" % (args.load, model.meta))
        if args.eval:
            evaluate(model, sizes, args.eval, args.eval_seed, sad=args.sad)
        return 0
    if not args.out:
        ap.error("--out is required when training")
    data, stats = build_training_data(args.systems, args.seed, (args.min_atoms, args.max_atoms))
    if not data:
        log("no system converged: nothing to train on")
        return 1
    model = lg.LearnedGuess(seed=0, layout=lg.CHO_SP)
    t0 = time.perf_counter()
    hist = model.train(data, epochs=args.epochs, batch=args.batch, lr=args.lr, log=log)
    t_train = time.perf_counter() - t0
    meta = dict(stats, train_s=t_train, epochs=args.epochs, batch=args.batch, lr=args.lr,
                loss_first=hist[0], loss_last=hist[-1], seed0=args.seed,
                atoms=[args.min_atoms, args.max_atoms],
                generator="random C/H/O clusters, s+p shells per atom (3 primitives), 20 (n/64)^(1/3) bohr cube")
    model.save(args.out, meta)
    log("saved %s: %d of %d systems converged, %d atom pairs, SCF %.0f s, training %.0f s, loss %.3e -> %.3e"
        % (args.out, stats["converged"], stats["systems"], stats["pairs"], stats["scf_s"], t_train, hist[0],
           hist[-1]))
    if args.eval:
        evaluate(model, sizes, args.eval, args.eval_seed, sad=args.sad)
    return 0


if __name__ == "__main__":
    sys.exit(main())
