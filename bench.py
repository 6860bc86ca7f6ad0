#!/usr/bin/env python3
"""Benchmark: RHF Fock build (direct, Schwarz-screened ERIs + J/K digest) and
RHF SCF on one MI355X.

Workload (default): seeded 160-atom C/H/O cluster in a 20-bohr cube, H s /
C s,p / O s,p,d shells (3-primitive s and p, 2-primitive d), 592 Cartesian
basis functions, Schwarz threshold 1e-12.  A step is one RHF two-electron
Fock build G = J[2D] - K[D] with the density already resident in HBM
(qbx_jk_device).  `value` is unique contracted Cartesian ERI quartets (i>=j,
k>=l, ij>=kl) surviving screening, per second, over the K timed steps.

After the timed steps (outside them) the script times the CPU oracle (a
restatement of the reference's per-component Obara-Saika algorithm) on a
bounded sample of the same quartets, and runs the full RHF to |dE| < 1e-8 Eh
in plain FP64 and in mixed precision with incremental builds.  The RHF runs
on a 24-molecule water cluster in cc-pVDZ (600 Cartesian functions): the
random-exponent basis of the cluster above gives an SCF that Quiqbox's
strategies (DD/ADIIS/DIIS from core or SAD guesses) do not converge.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-scf] [--no-cpu]
       python bench.py --copy-calibration   (a plain device copy of known size,
       profiled with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE to scale the
       counters to bytes)
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "quiqbox.jl_amd"))

METRIC = "RHF wall-time to |ΔE|<1e-8 Eh at ~600 basis fns; ERI quartets/s (% peak)"
PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0
# arithmetic per primitive quartet outside the generated VRR: geometry (p+q,
# 1/sqrt, rho, PQ, T, prefactor: ~22), Boys table Taylor (2*7 + 8*2 for exp(-dT)),
# downward recurrence (3 per order), scaling of F (1 per order), W-P / W-Q (6)
PRIM_OVERHEAD_FLOPS = 22 + 14 + 16 + 6


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


class Hip:
    """Minimal HIP runtime access for device buffers (libamdhip64)."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipFree.argtypes = [ctypes.c_void_p]

    def check(self, e, what):
        if e != 0:
            raise RuntimeError("%s failed with HIP error %d" % (what, e))

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        self.check(self.L.hipMalloc(ctypes.byref(p), nbytes), "hipMalloc")
        return p.value

    def h2d(self, dst, arr):
        arr = np.ascontiguousarray(arr)
        self.check(self.L.hipMemcpy(ctypes.c_void_p(dst), arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes, 1), "H2D")

    def d2h(self, arr, src):
        self.check(self.L.hipMemcpy(arr.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(src), arr.nbytes, 2), "D2H")

    def sync(self):
        self.check(self.L.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def free(self, p):
        self.L.hipFree(ctypes.c_void_p(p))


def copy_calibration(n_bytes=1 << 30, reps=3):
    """A plain elementwise device copy (y = -x over n_bytes of doubles): reads
    and writes exactly n_bytes per launch, the yardstick for the PMC counters."""
    import torch
    x = torch.ones(n_bytes // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read": n_bytes, "bytes_written": n_bytes, "launches": reps,
                                           "kernel": "elementwise neg (float64)"}}), flush=True)


def kernel_from_profile(tag):
    """Symbol of the class-`tag` FP64 J/K kernel in the committed PMC summary."""
    path = os.path.join(ROOT, "profiles", "eri_jk64_hbm_traffic_592bf.txt")
    if os.path.exists(path):
        with open(path) as f:
            for line in f:
                w = line.split()
                if len(w) >= 3 and w[0] == "kernel" and w[1] == tag:
                    return w[2]
    return "J/K kernel of class (%s), FP64" % tag


def traffic_from_profile(tag):
    """HBM bytes per launch of the class-`tag` JK kernel from the committed
    PMC summary (FETCH_SIZE + WRITE_SIZE scaled by the copy calibration), or
    None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", "eri_jk64_hbm_traffic_592bf.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            w = line.split()
            if len(w) >= 3 and w[0] == "traffic" and w[1] == tag:
                return float(w[2])
    return None


def build_traffic_from_profile():
    """Sum over the classes of the committed PMC summary's bytes per launch:
    the L2-miss (HBM-side) bytes of one FP64 Fock build's J/K kernels, or None."""
    path = os.path.join(ROOT, "profiles", "eri_jk64_hbm_traffic_592bf.txt")
    if not os.path.exists(path):
        return None
    tot, seen = 0.0, False
    with open(path) as f:
        for line in f:
            w = line.split()
            if len(w) >= 3 and w[0] == "traffic":
                tot += float(w[2])
                seen = True
    return tot if seen else None


def core_guess_density(gb, nuc, nocc):
    import quiqbox_amd as qb
    from quiqbox_amd.hartree_fock import getC, getD
    S = gb.overlap()
    H = gb.core_hamiltonian(nuc.charges, nuc.coord_array)
    X = qb.getOrthonormalization(S)
    return getD(getC(X, H), nocc)


def class_flops(flops_table, classes, stats):
    """Algorithmic FP64 flops per class for one Fock build."""
    out = []
    for c, l4 in enumerate(classes):
        tag = "%d%d%d%d" % tuple(l4)
        f = flops_table[tag]
        na, nb, nc, nd = [(l + 1) * (l + 2) // 2 for l in l4]
        nint = na * nb * nc * nd
        lt = int(sum(l4))
        per_prim = f["vrr_per_prim_quartet"] + PRIM_OVERHEAD_FLOPS + 4 * lt
        # digest: scaling (3/int), J (4/int), K (8/int)
        per_quartet = f["hrr_per_quartet"] + 15 * nint
        out.append(stats["prim_quartets"][c] * per_prim + stats["shell_quartets"][c] * per_quartet)
    return np.array(out, dtype=np.float64)


def host_cpu():
    """The host CPUs this process may use: model, logical CPUs of the machine,
    CPUs in this process's affinity set, and the OpenMP thread setting."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    return {"model": model, "machine_logical_cpus": os.cpu_count(), "affinity_cpus": affinity,
            "omp_num_threads": int(omp) if omp and omp.isdigit() else None}


_NATIVE_ORACLE = []


def native_oracle():
    """The oracle module, built once per process for this host (-march=native,
    OpenMP) in a temporary directory: the CPU baseline's timing build."""
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    if not _NATIVE_ORACLE:
        d = tempfile.mkdtemp(prefix="qbx_oracle_")
        oracle.use_native_build(d)
        _NATIVE_ORACLE.append(d)
    return oracle


def cpu_baseline(flat, n_quartets_total, seconds=12.0, threads=None):
    """Time the oracle (reference algorithm restated in C, compiled here for
    this host with -march=native and OpenMP) on random unique contracted
    quartets of the same basis, on every CPU this process is given (the
    affinity set, capped by OMP_NUM_THREADS where the machine sets it);
    returns quartets/s and the extrapolated time of one full screened Fock
    build's ERIs."""
    oracle = native_oracle()
    cpu = host_cpu()
    if threads is None:
        threads = cpu["affinity_cpus"]
        if cpu["omp_num_threads"]:
            threads = min(threads, cpu["omp_num_threads"])
    ob = oracle.Basis(flat.as_dict())
    n = flat.n_bf
    rng = np.random.default_rng(0)
    done, t_used, batch = 0, 0.0, 64
    while t_used < seconds:
        q = rng.integers(0, n, size=(batch, 4)).astype(np.int32)
        q[:, :2] = np.sort(q[:, :2], axis=1)[:, ::-1]
        q[:, 2:] = np.sort(q[:, 2:], axis=1)[:, ::-1]
        t0 = time.perf_counter()
        ob.eri_list(q, threads=threads)
        t_used += time.perf_counter() - t0
        done += batch
        batch = min(batch * 2, 100000)
    rate = done / t_used
    return dict(value=rate, unit="quartets/s", cores=threads, kind="port",
                sample="%d random unique contracted Cartesian quartets of the %d-function basis, %.1f s on %d "
                       "threads (oracle built -O3 -march=native on this host); one screened Fock build's %d "
                       "quartets extrapolate to %.0f s"
                       % (done, n, t_used, threads, n_quartets_total, n_quartets_total / rate),
                fock_build_s_extrapolated=n_quartets_total / rate, host=cpu)


def run_scf(nuc, gb, precision, thr=1e-8, incremental=False, density_solver="eigen"):
    import quiqbox_amd as qb
    cfg = qb.HFconfig(initial="CoreH", strategy=qb.SCFconfig(("ADIIS", "DIIS"), (1e-4, thr)), maxStep=120)
    t0 = time.perf_counter()
    res = qb.runHartreeFock(nuc, gb, cfg, printInfo=False, schwarzThreshold=1e-12, precision=precision,
                            incremental=incremental, densitySolver=density_solver)
    return res, time.perf_counter() - t0


def build_log(res):
    """One token per Fock build: F/i (full/incremental), d/m (FP64/mixed), milliseconds."""
    return " ".join("%s%s:%.0f" % ("F" if f else "i", "m" if p else "d", 1e3 * t) for f, p, t in res.timing["builds"])


def learned_section(nuc, gb, res64, t64, n_train, epochs):
    """Learned initial guess: train the pair-feature network on GPU-converged
    densities of small water clusters (cc-pVDZ, 2-6 molecules, seeds disjoint
    from the evaluation system), then start the 600-function water SCF from
    it, in FP64 and in mixed precision with incremental builds and purified
    densities."""
    import quiqbox_amd as qb
    from quiqbox_amd import learned_guess as lg
    from quiqbox_amd.systems import water_box
    t0 = time.perf_counter()
    systems = [water_box(2 + k % 5, 1000 + k, basis="cc-pVDZ") for k in range(n_train)]
    data = lg.training_set(systems, log=log)
    t_data = time.perf_counter() - t0
    model = lg.LearnedGuess(seed=0)
    t0 = time.perf_counter()
    hist = model.train(data, epochs=epochs, batch=512, lr=1e-3, log=log)
    t_train = time.perf_counter() - t0
    strat = qb.SCFconfig(("ADIIS", "DIIS"), (1e-4, 1e-8))
    out = {"training_systems": len(data), "training_pairs": int(sum(len(x) for x, _ in data)),
           "data_s": t_data, "train_s": t_train, "epochs": epochs, "loss_first": hist[0], "loss_last": hist[-1]}
    for tag, prec, inc, ds in (("fp64", qb.QBX_PREC_FP64, False, "eigen"),
                               ("fast", qb.QBX_PREC_MIXED, True, "purify"),
                               ("fast_fp64", qb.QBX_PREC_FP64, True, "purify")):
        t0 = time.perf_counter()
        info = {}
        cfg = lg.guess_config(model, nuc, gb, strat, 120, info=info)
        log("learned guess: %.2f predicted electrons of %.0f, %s (%.2f s)" % (
            info["electrons_predicted"], info["electrons"], "used" if info["learned"] else "core-guess fallback",
            time.perf_counter() - t0))
        out["electrons_predicted"] = info["electrons_predicted"]
        out["learned_used"] = info["learned"]
        res = qb.runHartreeFock(nuc, gb, cfg, printInfo=False, schwarzThreshold=1e-12, precision=prec,
                                incremental=inc, densitySolver=ds)
        t = time.perf_counter() - t0
        log("RHF %s from the learned guess: converged=%s in %.2f s, %d Fock builds, dE vs FP64 %.2e" % (
            tag, res.converged, t, res.timing["fock_builds"], res.energy[0] - res64.energy[0]))
        log("  builds: " + build_log(res))
        out[tag + "_wall_s"] = t
        out[tag + "_fock_builds"] = res.timing["fock_builds"]
        out[tag + "_converged"] = bool(res.converged)
        out[tag + "_energy_error_vs_fp64"] = res.energy[0] - res64.energy[0]
    out["speedup_fast_learned_vs_fp64_core"] = t64 / out["fast_wall_s"]
    out["speedup_fast_fp64_learned_vs_fp64_core"] = t64 / out["fast_fp64_wall_s"]
    return out


# Unseen clusters of the random C/H/O generator (s+p shells, cho_cluster) for
# the packaged network: the six held-out clusters of its GPU test and the
# three of the 24 evaluation clusters (seeds 916000-916007, 924000-924007,
# 932000-932007) on which any starting guess converged in 150 steps.
CHO_EVAL = [(16, 16010), (16, 16012), (16, 16013), (16, 16015), (24, 24010), (24, 24012),
            (16, 916002), (16, 916006), (24, 924004)]


def learned_cho_section():
    """The packaged C/H/O network (trained on GPU-converged densities of
    8-16-atom clusters of the random C/H/O generator) against the
    core-Hamiltonian guess and Quiqbox's SAD guess on unseen clusters of that
    generator: convergence in 150 steps, Fock builds and wall time."""
    from quiqbox_amd import learned_guess as lg
    from quiqbox_amd import train_guess as tg
    t0 = time.perf_counter()
    model = lg.load_cho_model()
    guesses = ["CoreH", "learned", "SAD"]
    rows = tg.evaluate_clusters(model, CHO_EVAL, sad=True, say=log)
    out = tg.summarize(rows, guesses)
    out["model"] = {k: model.meta[k] for k in ("systems", "converged", "pairs", "atoms", "epochs", "loss_last")}
    out["per_cluster"] = [{"atoms": r["atoms"], "seed": r["seed"],
                           **{g: [r[g]["converged"], r[g]["builds"], round(r[g]["wall_s"], 3)] for g in guesses}}
                          for r in rows]
    out["section_s"] = time.perf_counter() - t0
    log("C/H/O learned guess: %s; over the %d clusters on which every guess converged: %s" % (
        ", ".join("%s converged %d/%d" % (g, out[g]["converged"], out["clusters"]) for g in guesses),
        out["all_converged"], ", ".join("%s %d builds %.2f s" % (g, out[g]["builds_all_converged"],
                                                                out[g]["wall_s_all_converged"]) for g in guesses)))
    return out


def reference_benchmark_section(no_cpu):
    """The reference's own integral benchmarks (benchmark/benchmarks.jl:29-30,
    46, 70-72, 203-212): N2 with the nuclei at (0, 0, -1) and (0, 0, 1) bohr in
    cc-pVTZ (70 Cartesian functions, f shells included).  elecRepulsions (the
    dense ERI tensor, returned to the host) and nucAttractions through the
    public API on the GPU, each the median of five calls after one warm-up
    call; beside them the oracle's dense tensor on the host's threads (the CPU
    baseline) and the largest difference between the two tensors."""
    import quiqbox_amd as qb
    coords = [(0.0, 0.0, -1.0), (0.0, 0.0, 1.0)]
    bs = []
    for sym, c in zip(("N", "N"), coords):
        bs += qb.genGaussTypeOrbSeq(c, sym, "cc-pVTZ")

    def med(f):
        f()
        ts = []
        for _ in range(5):
            t0 = time.perf_counter()
            r = f()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)), r

    t_eri, E = med(lambda: qb.elecRepulsions(bs))
    t_nuc, _ = med(lambda: qb.nucAttractions(["N", "N"], coords, bs))
    n = E.shape[0]
    out = {"system": "N2 cc-pVTZ, nuclei at (0, 0, -1) and (0, 0, 1) bohr", "basis_functions": n,
           "elecRepulsions_s": t_eri, "nucAttractions_s": t_nuc,
           "unique_quartets": n * (n + 1) // 2 * (n * (n + 1) // 2 + 1) // 2}
    log("reference benchmark (N2 cc-pVTZ, %d functions): elecRepulsions %.1f ms (dense %d^4 tensor to the host), "
        "nucAttractions %.2f ms" % (n, 1e3 * t_eri, n, 1e3 * t_nuc))
    if not no_cpu:
        oracle = native_oracle()
        cpu = host_cpu()
        threads = cpu["affinity_cpus"]
        if cpu["omp_num_threads"]:
            threads = min(threads, cpu["omp_num_threads"])
        ob = oracle.Basis(qb.flatten_basis(bs).as_dict())
        ii, jj = np.tril_indices(n)
        P, Q = np.tril_indices(len(ii))
        q = np.stack([ii[P], jj[P], ii[Q], jj[Q]], axis=1).astype(np.int32)
        t0 = time.perf_counter()
        v = ob.eri_list(q, threads=threads)
        t_cpu = time.perf_counter() - t0
        dmax = float(np.max(np.abs(E[q[:, 0], q[:, 1], q[:, 2], q[:, 3]] - v)))
        out["cpu_baseline"] = {"oracle_unique_quartets_s": t_cpu, "threads": threads, "kind": "port",
                               "max_abs_difference": dmax}
        log("  oracle: the %d unique quartets in %.2f s on %d threads (GPU %.0fx faster); max |GPU - oracle| %.1e"
            % (len(q), t_cpu, threads, t_cpu / t_eri, dmax))
    return out


def scf_trace(atoms, seeds, steps, with_d=True):
    """RHF on the Fock-build workload's random C/H/O clusters (s, p and d
    shells; s and p only with with_d=False) from the core guess with three of Quiqbox's strategies, at most
    `steps` SCF steps each: per seed and strategy, convergence, the number of
    steps, the energies of the last ten steps, and the HOMO-LUMO gap of the
    last Fock matrix."""
    import quiqbox_amd as qb
    from quiqbox_amd.systems import random_cluster
    out = []
    for seed in seeds:
        nuc, bs = random_cluster(atoms, seed, with_d=with_d)
        gb = qb.GPUBasis(bs)
        S = gb.overlap()
        w, U = np.linalg.eigh(S)
        X = U / np.sqrt(w)
        nocc = int(round(sum(nuc.charges))) // 2
        for name, strat in (("ADIIS+DIIS", qb.SCFconfig(("ADIIS", "DIIS"), (1e-4, 1e-8))),
                            ("DIIS", qb.SCFconfig(("DIIS",), (1e-8,))),
                            ("EDIIS+DIIS", qb.SCFconfig(("EDIIS", "DIIS"), (1e-4, 1e-8)))):
            t0 = time.perf_counter()
            res = qb.runHartreeFock(nuc, gb, qb.HFconfig(initial="CoreH", strategy=strat, maxStep=steps),
                                    printInfo=False)
            t = time.perf_counter() - t0
            E = [float(e) for e in res.memory[0].shared.Etots]
            eps = np.linalg.eigvalsh(X.T @ np.asarray(res.fock[0]) @ X)
            gap = float(eps[nocc] - eps[nocc - 1])
            tail = E[-10:]
            rec = {"atoms": atoms, "seed": seed, "functions": gb.n, "strategy": name, "converged": bool(res.converged),
                   "steps": len(E) - 1, "wall_s": t, "last_energies": tail, "last10_spread": max(tail) - min(tail),
                   "homo_lumo_gap": gap}
            out.append(rec)
            log("scf trace: %d atoms seed %d (%d functions) %s: %s after %d steps (%.1f s); last ten energies "
                "%s; spread %.2e Eh; HOMO-LUMO gap %.4f Eh" % (
                    atoms, seed, gb.n, name, "converged" if res.converged else "not converged", len(E) - 1, t,
                    " ".join("%.6f" % e for e in tail), rec["last10_spread"], gap))
    return out


def gradient_section(steps):
    """Basis-parameter gradient on the 32-atom H cluster with floating
    4-primitive s and p shells (256 exponents, 256 coefficients, 96 centre
    coordinates): one adjoint gradient (SCF + kernels) timed alone, then
    `steps` gradient-descent steps of basis optimization."""
    import quiqbox_amd as qb
    from quiqbox_amd.gradient import ParamBasis, optimize_basis, rhf_gradient
    from quiqbox_amd.systems import h_cluster_sp
    nuc, _ = h_cluster_sp(32, 0)
    pb = ParamBasis.random_h_cluster(32, 0)
    gb = qb.GPUBasis(pb.orbitals())
    cfg = qb.HFconfig(initial="CoreH", strategy=qb.SCFconfig(("ADIIS", "DIIS"), (1e-4, 1e-10)), maxStep=200)
    t0 = time.perf_counter()
    res = qb.runHartreeFock(nuc, gb, cfg, printInfo=False)
    t_scf = time.perf_counter() - t0
    t0 = time.perf_counter()
    E, g, _ = rhf_gradient(nuc, pb, res=res, gpu_basis=gb)
    t_grad = time.perf_counter() - t0
    log("gradient: %d parameters, SCF %.3f s, adjoint gradient %.3f s, |g| %.3e" % (pb.n_theta, t_scf, t_grad,
                                                                                     np.linalg.norm(g)))
    t0 = time.perf_counter()
    _, hist = optimize_basis(nuc, pb, steps=steps, lr=0.02, config=cfg,
                             callback=lambda it, E, g: log("  step %d: E %.10f |g| %.3e" % (it, E, np.linalg.norm(g))))
    t_opt = time.perf_counter() - t0
    log("basis optimization: %d steps in %.2f s, E %.8f -> %.8f" % (steps, t_opt, hist[0], hist[-1]))
    return {"workload": "32 H atoms, floating 4-primitive s and p shells (seed 0), 128 functions, %d parameters"
                        % pb.n_theta,
            "scf_s": t_scf, "gradient_s": t_grad, "grad_norm": float(np.linalg.norm(g)),
            "opt_steps": steps, "opt_wall_s": t_opt, "opt_s_per_step": t_opt / steps,
            "energy_first": hist[0], "energy_last": hist[-1]}


def mo_section(n, reps=3):
    """changeOrbitalBasis (Interface.jl:376-380) of a dense n^4 two-body tensor
    with an n x n coefficient matrix: four quarter transformations, each one
    FP64 matrix-core GEMM of 2 n^5 flops.  Random values (the transform needs no
    index symmetry); host arrays in and out, as the reference's interface hands
    them over, so the wall time includes the PCIe copies."""
    import quiqbox_amd as qb
    rng = np.random.default_rng(0)
    A = rng.standard_normal((n,) * 4)
    C = rng.standard_normal((n, n))
    ctx = qb.default_context()
    ctx.mo_transform(A, C)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        ctx.mo_transform(A, C)
        ts.append(time.perf_counter() - t0)
    t = min(ts)
    flops = 4 * 2.0 * n ** 5
    log("mo transform n %d: %.1f ms wall (%.2f GB each way over PCIe), %.2f TFLOP/s including copies"
        % (n, 1e3 * t, 8.0 * n ** 4 / 1e9, flops / t / 1e12))
    return {"workload": "two-body tensor %d^4 FP64, square coefficient matrix" % n, "wall_ms": 1e3 * t,
            "gemm_flops": flops, "tflops_incl_pcie": flops / t / 1e12, "host_bytes_each_way": 8 * n ** 4}


def rhf_section(hip, cpu_seconds, no_cpu):
    """RHF wall time to |dE| < 1e-8 Eh on the 600-function water cluster: plain
    FP64 and mixed precision + incremental builds, plus that system's FP64 Fock
    build rate and the CPU oracle's extrapolated time for the same SCF."""
    import quiqbox_amd as qb
    from quiqbox_amd.systems import water_box
    nuc, bs = water_box(24, 0, basis="cc-pVDZ")
    gb = qb.GPUBasis(bs)
    n = gb.n
    log("RHF workload: 24 H2O in cc-pVDZ, %d basis functions" % n)
    nocc = int(round(sum(nuc.charges))) // 2
    D = core_guess_density(gb, nuc, nocc)
    nb = n * n * 8
    dDJ, dDK, dG = hip.alloc(nb), hip.alloc(nb), hip.alloc(nb)
    hip.h2d(dDJ, 2 * D)
    hip.h2d(dDK, D)
    stats = gb.jk_stats(1e-12)
    n_cart = int(stats["cart_quartets"].sum())
    gb.jk_device(dDJ, dDK, dG, 1, 1e-12)
    hip.sync()
    t0 = time.perf_counter()
    for _ in range(3):
        gb.jk_device(dDJ, dDK, dG, 1, 1e-12)
    hip.sync()
    t_build = (time.perf_counter() - t0) / 3
    for p in (dDJ, dDK, dG):
        hip.free(p)
    log("water Fock build: %.2f ms, %.3e quartets/s" % (1e3 * t_build, n_cart / t_build))
    res64, t64 = run_scf(nuc, gb, qb.QBX_PREC_FP64)
    log("RHF FP64: E=%.10f converged=%s in %.2f s, %d Fock builds" % (
        res64.energy[0] + res64.energy[1], res64.converged, t64, res64.timing["fock_builds"]))
    rhf_section.last = (nuc, gb, res64, t64)
    respu, tpu = run_scf(nuc, gb, qb.QBX_PREC_FP64, density_solver="purify")
    log("RHF FP64 + purified densities: E=%.10f converged=%s in %.2f s, %d Fock builds" % (
        respu.energy[0] + respu.energy[1], respu.converged, tpu, respu.timing["fock_builds"]))
    resmx, tmx = run_scf(nuc, gb, qb.QBX_PREC_MIXED, incremental=True, density_solver="purify")
    log("RHF mixed+incremental+purified: E=%.10f converged=%s in %.2f s (%s of %s builds mixed)" % (
        resmx.energy[0] + resmx.energy[1], resmx.converged, tmx, resmx.timing["mixed_builds"],
        resmx.timing["fock_builds"]))
    log("  builds: " + build_log(resmx))
    out = {"workload": "24 H2O (liquid density, seeded orientations), cc-pVDZ, %d Cartesian functions, "
                       "core-Hamiltonian guess, ADIIS then DIIS, |dE| < 1e-8 Eh, Schwarz 1e-12" % n,
           "fock_build_ms": 1e3 * t_build, "fock_build_quartets_per_s": n_cart / t_build,
           "screened_cart_quartets": n_cart,
           "fp64_wall_s": t64, "fp64_energy": res64.energy[0] + res64.energy[1],
           "fp64_converged": bool(res64.converged), "fp64_fock_builds": res64.timing["fock_builds"],
           "fp64_purify_wall_s": tpu, "fp64_purify_fock_builds": respu.timing["fock_builds"],
           "fp64_purify_energy_error_vs_fp64": (respu.energy[0] - res64.energy[0]),
           "fast_wall_s": tmx, "fast_mode": "mixed precision (FP32 below Schwarz 1e-5, FP64 rebuilds near "
                                           "convergence) + density-screened incremental builds + "
                                           "densities by purification on the GPU",
           "fast_energy_error_vs_fp64": (resmx.energy[0] - res64.energy[0]),
           "fast_converged": bool(resmx.converged), "fast_fock_builds": resmx.timing["fock_builds"],
           "fast_mixed_builds": resmx.timing["mixed_builds"]}
    if not no_cpu:
        cb = cpu_baseline(gb.flat, n_cart, cpu_seconds)
        out["cpu_fock_build_s_extrapolated"] = cb["fock_build_s_extrapolated"]
        out["cpu_scf_s_extrapolated"] = cb["fock_build_s_extrapolated"] * res64.timing["fock_builds"]
        out["cpu_sample"] = cb["sample"]
        log("cpu oracle on the water basis: %.3e quartets/s -> %.0f s per Fock build" % (
            cb["value"], cb["fock_build_s_extrapolated"]))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--atoms", type=int, default=160)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-scf", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--learned-systems", type=int, default=40,
                    help="training systems of the learned-guess section (0: skip it)")
    ap.add_argument("--learned-epochs", type=int, default=300)
    ap.add_argument("--scf-trace", default=None, metavar="SEEDS",
                    help="only run RHF on the workload's clusters of these seeds (comma list) and print the traces")
    ap.add_argument("--scf-steps", type=int, default=100)
    ap.add_argument("--scf-sp", action="store_true", help="--scf-trace on s/p clusters (no d shells)")
    ap.add_argument("--no-refbench", action="store_true",
                    help="skip the reference's N2 cc-pVTZ integral benchmarks")
    ap.add_argument("--no-cho", action="store_true",
                    help="skip the packaged C/H/O network's evaluation on unseen C/H/O clusters")
    ap.add_argument("--grad-steps", type=int, default=100,
                    help="basis-optimization steps of the gradient section (0: skip it)")
    ap.add_argument("--mo-size", type=int, default=128,
                    help="orbital-basis change of an n^4 two-body tensor (0: skip it)")
    ap.add_argument("--serial", action="store_true",
                    help="run the class kernels of each build one after another on one stream (for profiling)")
    ap.add_argument("--timed-only", action="store_true",
                    help="skip the serialized per-class builds after the timed region (profiling runs: every "
                         "J/K dispatch is then one of the concurrent builds)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.scf_trace is not None:
        recs = scf_trace(args.atoms, [int(x) for x in args.scf_trace.split(",") if x], args.scf_steps,
                         with_d=not args.scf_sp)
        print(json.dumps({"scf_trace": recs}), flush=True)
        return

    import quiqbox_amd as qb
    from quiqbox_amd.systems import random_cluster

    nuc, bs = random_cluster(args.atoms, args.seed, with_d=True)
    nocc = int(round(sum(nuc.charges))) // 2
    log("workload: %d atoms, %d basis functions, %d occupied" % (args.atoms, len(bs), nocc))
    gb = qb.GPUBasis(bs)
    n = gb.n
    thr = 1e-12
    D = core_guess_density(gb, nuc, nocc)
    hip = Hip()
    nb = n * n * 8
    dDJ, dDK, dG = hip.alloc(nb), hip.alloc(nb), hip.alloc(nb)
    hip.h2d(dDJ, 2 * D)
    hip.h2d(dDK, D)
    stats = gb.jk_stats(thr)
    classes = qb.class_table()
    with open(os.path.join(ROOT, "quiqbox.jl_amd", "csrc", "gen", "eri_flops.json")) as f:
        flops_table = json.load(f)
    cflops = class_flops(flops_table, classes, stats)
    n_cart = int(stats["cart_quartets"].sum())
    log("screened: %d shell quartets, %d unique Cartesian quartets, %d primitive shell quartets"
        % (stats["shell_quartets"].sum(), n_cart, stats["prim_quartets"].sum()))

    for _ in range(args.warmup):
        gb.jk_device(dDJ, dDK, dG, 1, thr)
    hip.sync()
    # timed region: K Fock builds, inputs resident in HBM; each class kernel
    # bracketed by HIP events on the stream it runs on (concurrent timing: the
    # build runs as it does untimed, the durations are those of the
    # concurrent build)
    gb.ctx.set_timing(1 if args.serial else 2)
    hip.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gb.jk_device(dDJ, dDK, dG, 1, thr)
    hip.sync()
    t_steps = time.perf_counter() - t0
    ms_step = 1000.0 * t_steps / args.steps
    quartets_per_s = n_cart / (t_steps / args.steps)
    conc, conc_total = gb.ctx.last_timing()
    log("Fock build: %.2f ms/step, %.3e quartets/s (event-timed build %.2f ms)" % (ms_step, quartets_per_s,
                                                                                 conc_total))

    # the same classes one at a time (serialized builds, outside the timed region)
    ser = np.full(len(classes), np.nan)
    if not args.timed_only:
        gb.ctx.set_timing(1)
        ser[:] = 0.0
        reps = max(1, min(3, args.steps))
        for _ in range(reps):
            gb.jk_device(dDJ, dDK, dG, 1, thr)
            ms, tot = gb.ctx.last_timing()
            ser += ms
        ser /= reps
    gb.ctx.set_timing(0)
    # dominant kernel: the largest share of the concurrent build
    dom = int(np.argmax(conc))
    tag = "%d%d%d%d" % tuple(classes[dom])
    achieved = cflops[dom] / (conc[dom] * 1e-3) / 1e12
    achieved_ser = cflops[dom] / (ser[dom] * 1e-3) / 1e12
    build_tflops = cflops.sum() / (ms_step * 1e-3) / 1e12
    log("dominant kernel (concurrent build): class (%s) %.2f ms -> %.2f TFLOP/s; serialized %.2f ms -> %.2f TFLOP/s"
        % (tag, conc[dom], achieved, ser[dom], achieved_ser))
    log("whole build: %.2f TFLOP/s algorithmic (%.3e flops per build); serialized class kernels sum %.2f ms"
        % (build_tflops, cflops.sum(), ser.sum()))
    for c in np.argsort(-conc):
        log("  class %s: concurrent %.3f ms, serialized %.3f ms, %d shell quartets, %.2f TFLOP/s serialized" % (
            "%d%d%d%d" % tuple(classes[c]), conc[c], ser[c], stats["shell_quartets"][c],
            cflops[c] / max(ser[c], 1e-9) / 1e9))

    out = {
        "metric": METRIC,
        "value": quartets_per_s,
        "unit": "quartets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "RHF Fock build, seeded 160-atom C/H/O cluster, s/p/d Cartesian shells, "
                               "%d basis functions, Schwarz 1e-12, FP64" % n,
                   "atoms": args.atoms, "basis_functions": n, "seed": args.seed, "schwarz": thr,
                   "screened_shell_quartets": int(stats["shell_quartets"].sum()),
                   "screened_cart_quartets": n_cart,
                   "primitive_shell_quartets": int(stats["prim_quartets"].sum())},
        "roofline": {"bound": "valu", "pipe": "fp64-valu", "kernel": kernel_from_profile(tag),
                     "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic_from_profile(tag),
                     "kernel_ms": float(conc[dom]), "kernel_flops": float(cflops[dom]),
                     "timing": "average duration in the concurrent builds of the timed region (HIP events "
                               "on the kernel's stream)",
                     "serialized_ms": None if args.timed_only else float(ser[dom]),
                     "serialized_tflops": None if args.timed_only else achieved_ser,
                     "serialized_frac": None if args.timed_only else achieved_ser / PEAK_FP64_TFLOPS,
                     "build_tflops": build_tflops, "build_frac": build_tflops / PEAK_FP64_TFLOPS},
    }
    from quiqbox_amd import _lib
    prov = _lib.provenance()
    out["library"] = {"source_digest": prov["library"], "tree_source_digest": prov["sources"],
                      "built_from_these_sources": prov["match"]}
    # J/K bytes per second: the profile's bytes per launch over the measured times
    tr, btr = traffic_from_profile(tag), build_traffic_from_profile()
    out["roofline"]["kernel_hbm_gbs"] = tr / (1e-3 * float(conc[dom])) / 1e9 if tr else None
    out["roofline"]["build_hbm_bytes"] = btr
    out["roofline"]["build_hbm_gbs"] = btr / (1e-3 * ms_step) / 1e9 if btr else None
    out["roofline"]["hbm_peak_gbs"] = 8000.0
    if btr:
        log("J/K traffic: %.1f GB of L2 misses per build (PMC summary) -> %.0f GB/s over the build (%.1f%% of the "
            "8 TB/s HBM peak)" % (btr / 1e9, out["roofline"]["build_hbm_gbs"], out["roofline"]["build_hbm_gbs"] / 80))

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(gb.flat, n_cart, args.cpu_seconds)
        log("cpu baseline: %.3e quartets/s on %d threads (%s)" % (out["cpu_baseline"]["value"],
                                                                 out["cpu_baseline"]["cores"],
                                                                 out["cpu_baseline"]["host"]))
    for p in (dDJ, dDK, dG):
        hip.free(p)
    if not args.no_scf:
        out["rhf"] = rhf_section(hip, min(args.cpu_seconds, 8.0), args.no_cpu)
        if args.learned_systems > 0:
            out["learned_guess"] = learned_section(*rhf_section.last, args.learned_systems, args.learned_epochs)
        if not args.no_cho:
            out["learned_guess_cho"] = learned_cho_section()
        if args.grad_steps > 0:
            out["gradient"] = gradient_section(args.grad_steps)
    if not args.no_refbench:
        out["reference_benchmark"] = reference_benchmark_section(args.no_cpu)
    if args.mo_size > 0:
        out["mo_transform"] = mo_section(args.mo_size)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
