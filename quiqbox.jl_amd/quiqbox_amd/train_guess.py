"""Train the learned initial guess on GPU-converged densities of random C/H/O
clusters (the s/p cluster generator, 8-48 atoms) and save the network.

    python -m quiqbox_amd.train_guess --systems 2000 --out model.npz

Every system runs a GPU RHF (core guess, ADIIS then DIIS to 1e-9 Eh); the
converged ones supply the (pair features, density block) training pairs, and
the network trains on them with the library's BF16 matrix-core kernels.  The
saved file records how many systems converged, the SCF and training times and
the loss history (`LearnedGuess.load(path).meta`).
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


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--systems", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=100000)
    ap.add_argument("--min-atoms", type=int, default=8)
    ap.add_argument("--max-atoms", type=int, default=48)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", required=True)
    args = ap.parse_args(argv)
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
    return 0


if __name__ == "__main__":
    sys.exit(main())
