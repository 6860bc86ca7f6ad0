#!/usr/bin/env python3
"""Generate the per-angular-class ERI code for the HIP kernels.

For every class (la lb | lc ld) with la >= lb, lc >= ld and
(la, lb) >= (lc, ld) this emits a straight-line device function

    template <typename R> struct EriClass<la, lb, lc, ld, R> {
        static __device__ void vrr(...);       // per primitive quartet, accumulates [e0|f0]
        static __device__ void hrr_ket(...);   // [e0|f0] -> [e|cd)
        static __device__ void hrr_bra(...);   // [e|cd) -> (ab|cd) for one component a
    };

Scheme (Head-Gordon & Pople, J. Chem. Phys. 89, 5777 (1988), built on the
Obara-Saika recurrences, J. Chem. Phys. 84, 3963 (1986)):
  1. per primitive quartet, the Boys values F_m(T) scaled by the prefactor
     2 pi^{5/2} K_AB K_CD / (p q sqrt(p+q)) seed [00|00]^(m);
  2. the vertical recurrence builds [e0|f0]^(0) for |e| in [la, la+lb] and
     |f| in [lc, lc+ld]; only intermediates reachable from those targets are
     emitted;
  3. [e0|f0] is summed over the primitive quartets of the shell quartet
     (contraction happens before any horizontal transfer);
  4. the horizontal recurrence moves angular momentum to the second centre
     of the ket (hrr_ket, once), then of the bra (hrr_bra, one Cartesian
     component of shell A at a time, so the caller can consume each slice of
     the integral block before the next is formed):
     (a,b+1_i| = (a+1_i,b| + (A-B)_i (a,b|.

Cartesian components follow the order used by Quiqbox (Lexicons.jl:17-36):
for i = L..0, for j = L-i..0: (i, j, L-i-j).

Staged classes (the largest: their [e0|f0] accumulators do not fit the
register file) are emitted in groups of bra components e: vrr_grp<g> sums the
targets of group g only, hrr_ket_grp<g> transfers that group to the ket, and
the caller stores each group's [e|cd) to a per-wave workspace before the next
group's pass over the primitive quartets.  hrr_bra_ab then forms one
(a b | c d) row for all cd from the workspace, the bra transfer written out
as polynomials in (A-B):  (a,b| = sum_e P_e(A-B) (e,0|.

For the RHF energy gradient (classes with s and p shells) it also emits
EriAdj<la, lb, lc, ld>: the reverse-mode (adjoint) form of the same
recurrences, written out statement by statement from the recurrence graph:
  hrr_adj   seeds on the contracted (ab|cd) -> seeds on the contracted
            [e0|f0] plus the derivatives with respect to A-B and C-D;
  vrr_adj   per primitive quartet, the forward recurrence, then its adjoint
            with the [e0|f0] seeds -> derivatives with respect to the Boys
            values F_m and the geometric inputs (PA, WP, QC, WQ, 1/2p, 1/2q,
            1/2(p+q), q/(p+q), p/(p+q)), and the seeded sum itself.

The generator also writes eri_flops.json: arithmetic operations per primitive
quartet (VRR + Boys seed) and per contracted quartet (HRR) for each class,
counted from the emitted expressions (add/sub/mul = 1, fma pattern a*b+c
counted as 2).
"""
import functools
import itertools
import json
import os
import sys

LMAX = 2
# shell pairs are classified up to this angular momentum (f and g shells);
# quartets with a shell above LMAX run the generic kernel (qbx_eri_generic.hip)
PAIR_LMAX = 4
# classes with an adjoint (gradient) kernel: every shell s or p
GRAD_LMAX = 1
# classes evaluated in groups of bra components through a workspace
STAGED = {(2, 1, 1, 1), (2, 1, 2, 1), (2, 2, 1, 1), (2, 2, 2, 0), (2, 2, 2, 1), (2, 2, 2, 2)}
STAGE_ACC = 32     # accumulators per group
BRA_BATCH = 24     # workspace values per batch of loads in the bra transfer of a staged class
COST_PRIM0 = 80    # launch-order cost model: geometry and Boys values per primitive quartet
COST_DIGEST = 8    # launch-order cost model: J/K digest per Cartesian quartet


def ncart_prod(c):
    return functools.reduce(lambda x, l: x * (l + 1) * (l + 2) // 2, c, 1)
# Classes whose staged kernel runs at one wave per SIMD with larger groups: the
# register file of a lone wave (VGPRs plus AGPRs as spill space) holds 62
# accumulators (two bra components e), so the class makes 16 passes over its
# primitive quartets instead of 31.  Measured on the 592-function cluster:
# (dd|dd) 4.1 -> 2.7 ms at 21 passes (one wave per chunk); the other staged
# classes lose more latency hiding than they save.  16 passes also split best
# over the 4 waves that share a chunk (largest per-wave cost 6848 against 7386
# at 21 passes, both with the assignment below).
STAGE_ONE_WAVE = {(2, 2, 2, 2): 62}
# Staged classes whose 64-ket chunks are each shared by the waves of a block
# (value: waves per block; groups, then digest rows, spread over them; the
# bra's K rows in LDS).  Measured serialized on the 592-function cluster,
# against one wave per chunk: (dd|dd) 2.78 -> 0.82 ms (12470 shell quartets,
# too few tasks to give every SIMD more than one wave), (dd|dp) 3.89 -> 2.66
# ms, (dp|dp) 2.50 -> 2.14 ms; (dp|pp) 3.72 -> 4.29, (dd|pp) 1.47 -> 1.61 and
# (dd|ds) 1.13 -> 1.27 ms lost (many tasks already fill the chip, and the
# extra barriers cost).
STAGE_SPLIT = {(2, 2, 2, 2): 4, (2, 2, 2, 1): 4, (2, 1, 2, 1): 4}
# 64-ket chunks per task (one bra, its K rows flushed once per task) of the
# classes of STAGE_SPLIT.  Each chunk already occupies a whole block, so small
# tasks balance best.  Measured on the 592-function cluster against 8 chunks a
# task, serialized: (dd|dd) 0.75 -> 0.60 ms at one chunk (0.72 at two), (dd|dp)
# 2.55 -> 1.60 ms at two (1.45 at one), (dp|dp) 2.11 -> 1.87 ms at two (2.00 at
# one); the concurrent build was 0.5-1 ms slower with one-chunk tasks for all
# three than with two (more K-row flushes among the other classes' atomics).
STAGE_SPLIT_CHUNKS = {(2, 2, 2, 2): 1, (2, 2, 2, 1): 2, (2, 1, 2, 1): 2}
# Classes whose K-row J/K kernel is compiled for two waves per SIMD (256
# registers, the rest spilled to scratch) instead of the compiler's choice of
# one wave with AGPRs.  Measured on the 592-function cluster, serialized:
# (ds|pp) 1.93 -> 1.75 ms, (dp|ps) 3.34 -> 3.30 ms, build 32.3-32.5 -> 32.1 ms.
# The classes that need 400+ registers ((pp|pp), (dp|ds), (dd|ps)) would
# spill 640-940 bytes per lane.
ROWS_TWO_WAVES = {(2, 1, 1, 0), (2, 0, 1, 1)}
# Classes launched after all others in a Fock build.  (dp|ds)'s K-row kernel
# needs 512 registers a wave (one wave per SIMD, a 4-wave block per CU) and
# spills to scratch, so it runs on the main stream, and its blocks start only
# on CUs that have drained completely.  Queued behind (dp|ps) and (ds|pp) it
# starts when the staged classes are done and holds CUs for a shorter stretch:
# its average duration in the concurrent build 18.2 -> 14.2 ms, the build
# 32.30 -> 32.42 ms (five A/B pairs on the 592-function cluster, within the
# run-to-run spread).
LAUNCH_LAST = {(2, 1, 2, 0)}
TWIN_MAX_ACC = 9   # must match qbx_eri_api.h


def cart(L):
    return [(i, j, L - i - j) for i in range(L, -1, -1) for j in range(L - i, -1, -1)]


def ncart(L):
    return (L + 1) * (L + 2) // 2


def sub1(t, i):
    l = list(t)
    l[i] -= 1
    return tuple(l)


def add1(t, i):
    l = list(t)
    l[i] += 1
    return tuple(l)


def pick_axis(t):
    """axis to lower: the largest component, ties to the lowest axis"""
    best = -1
    bi = -1
    for i in range(3):
        if t[i] > best:
            best = t[i]
            bi = i
    return bi


class Emitter:
    def __init__(self):
        self.lines = []
        self.names = {}
        self.count = 0
        self.flops = 0

    def fresh(self, prefix):
        self.count += 1
        return "%s%d" % (prefix, self.count)


AX = "xyz"


def gen_vrr(em, targets):
    """Emit VRR code; returns dict (e, f) -> variable name for m = 0."""
    memo = {}

    def V(e, f, m):
        key = (e, f, m)
        if key in memo:
            return memo[key]
        if sum(e) == 0 and sum(f) == 0:
            memo[key] = "F[%d]" % m
            return memo[key]
        terms = []
        if sum(f) > 0:
            i = pick_axis(f)
            f1 = sub1(f, i)
            a = V(e, f1, m)
            b = V(e, f1, m + 1)
            terms.append(("QC%s" % AX[i], a))
            terms.append(("WQ%s" % AX[i], b))
            if f1[i] > 0:
                f2 = sub1(f1, i)
                c = V(e, f2, m)
                d = V(e, f2, m + 1)
                terms.append(("(R)%d * inv2q" % f1[i], "(%s - rho_q * %s)" % (c, d)))
                em.flops += 2
            if e[i] > 0:
                e1 = sub1(e, i)
                g = V(e1, f1, m + 1)
                terms.append(("(R)%d * inv2pq" % e[i], g))
        else:
            i = pick_axis(e)
            e1 = sub1(e, i)
            a = V(e1, f, m)
            b = V(e1, f, m + 1)
            terms.append(("PA%s" % AX[i], a))
            terms.append(("WP%s" % AX[i], b))
            if e1[i] > 0:
                e2 = sub1(e1, i)
                c = V(e2, f, m)
                d = V(e2, f, m + 1)
                terms.append(("(R)%d * inv2p" % e1[i], "(%s - rho_p * %s)" % (c, d)))
                em.flops += 2
        name = em.fresh("v")
        expr = " + ".join("%s * %s" % t for t in terms)
        em.flops += 2 * len(terms) - 1
        em.lines.append("const R %s = %s;" % (name, expr))
        memo[key] = name
        return name

    out = {}
    for (e, f) in targets:
        out[(e, f)] = V(e, f, 0)
    return out


def bra_expansion(a, b):
    """(a b| as {e: {(px, py, pz): integer coefficient}}: the bra transfer
    (a, b+1_i| = (a+1_i, b| + (A-B)_i (a, b| unrolled down to b = 0."""
    if sum(b) == 0:
        return {a: {(0, 0, 0): 1}}
    i = pick_axis(b)
    b1 = sub1(b, i)
    res = {e: dict(p) for e, p in bra_expansion(add1(a, i), b1).items()}
    for e, poly in bra_expansion(a, b1).items():
        tgt = res.setdefault(e, {})
        for mono, c in poly.items():
            m2 = add1(mono, i)
            tgt[m2] = tgt.get(m2, 0) + c
    return res


def mono_expr(mono):
    f = []
    for ax, pw in zip(AX, mono):
        f += ["AB%s" % ax] * pw
    return " * ".join(f) if f else "(R)1"


def gen_staged(la, lb, lc, ld, es, fs, eidx, fidx, C_, D_):
    """Grouped VRR / ket HRR and the workspace bra HRR of a staged class."""
    nE, nF = len(es), len(fs)
    nC, nD = len(C_), len(D_)
    NCD = nC * nD
    acc_per_group = STAGE_ONE_WAVE.get((la, lb, lc, ld), STAGE_ACC)
    G = -(-nE * nF // acc_per_group)
    G = min(G, nE)
    bounds = [k * nE // G for k in range(G + 1)]
    code = []
    flops_vrr_exec = 0
    code.append("    static constexpr bool STAGED = true;")
    code.append("    static constexpr int WAVES_PER_SIMD = %d;" % (1 if (la, lb, lc, ld) in STAGE_ONE_WAVE else 2))
    code.append("    static constexpr int NGRP = %d;" % G)
    code.append("    static constexpr int SPLIT = %d;" % STAGE_SPLIT.get((la, lb, lc, ld), 1))
    code.append("    static constexpr int NCD = %d;" % NCD)
    code.append("    // group g holds bra components e in [ge0(g), ge0(g + 1))")
    sel = " : ".join("g == %d ? %d" % (k, bounds[k]) for k in range(G)) + " : %d" % nE
    code.append("    static constexpr int ge0(int g) { return %s; }" % sel)
    code.append("    static constexpr int GMAXACC = %d;" % (max(bounds[k + 1] - bounds[k] for k in range(G)) * nF))
    gcost = []
    for g in range(G):
        eg = es[bounds[g]:bounds[g + 1]]
        em = Emitter()
        targets = [(e, f) for e in eg for f in fs]
        vv = gen_vrr(em, targets)
        lines = list(em.lines)
        for (e, f) in targets:
            lines.append("acc[%d] += %s;" % ((eidx[e] - bounds[g]) * nF + fidx[f], vv[(e, f)]))
        flops_vrr_exec += em.flops + len(targets)
        gcost.append(em.flops + len(targets) + COST_PRIM0)
        code.append("    static __device__ __forceinline__ void vrr_g%d(const R* __restrict__ F, R* __restrict__ acc," % g)
        code.append("        R PAx, R PAy, R PAz, R WPx, R WPy, R WPz, R QCx, R QCy, R QCz,")
        code.append("        R WQx, R WQy, R WQz, R inv2p, R inv2q, R inv2pq, R rho_p, R rho_q) {")
        code += ["        " + l for l in lines]
        code.append("    }")
        # ket transfer of this group: kout[(e - ge0) * NCD + cd]
        hm = Emitter()
        memo = {}

        def K(e, c, d):
            key = (e, c, d)
            if key in memo:
                return memo[key]
            if sum(d) == 0:
                r = "acc[%d]" % ((e - bounds[g]) * nF + fidx[c])
                memo[key] = r
                return r
            i = pick_axis(d)
            d1 = sub1(d, i)
            x = K(e, add1(c, i), d1)
            y = K(e, c, d1)
            name = hm.fresh("k")
            hm.lines.append("const R %s = %s + CD%s * %s;" % (name, x, AX[i], y))
            hm.flops += 2
            memo[key] = name
            return name

        outs = []
        for ie in range(bounds[g], bounds[g + 1]):
            for ic, c in enumerate(C_):
                for idd, d in enumerate(D_):
                    outs.append("ws.st(%d, %s);" % (ie * NCD + ic * nD + idd, K(ie, c, d)))
        code.append("    template <class W> static __device__ __forceinline__ void hrr_ket_g%d(const R* __restrict__ acc, "
                    "const W& ws, R CDx, R CDy, R CDz) {" % g)
        code += ["        " + l for l in hm.lines + outs]
        code.append("    }")
    # Groups over the SPLIT waves of a block that share a chunk: longest
    # processing time first on the per-primitive-quartet cost of each group's
    # pass (its VRR plus the geometry and Boys values every pass repeats)
    W = STAGE_SPLIT.get((la, lb, lc, ld), 1)
    wave_of = [0] * G
    load = [0] * W
    for g in sorted(range(G), key=lambda g: -gcost[g]):
        w = min(range(W), key=lambda w: load[w])
        wave_of[g] = w
        load[w] += gcost[g]
    rr = [sum(gcost[g] for g in range(G) if g % W == w) for w in range(W)]
    code.append("    // wave of a block (of SPLIT sharing a chunk) that evaluates group g; per-wave cost %s "
                "(round robin: %s)" % (load, rr))
    sel = " : ".join("g == %d ? %d" % (g, wave_of[g]) for g in range(G - 1)) + (" : %d" % wave_of[G - 1])
    code.append("    static constexpr int grp_wave(int g) { return %s; }" % sel)
    args = "F, acc, PAx, PAy, PAz, WPx, WPy, WPz, QCx, QCy, QCz, WQx, WQy, WQz, inv2p, inv2q, inv2pq, rho_p, rho_q"
    code.append("    template <int G> static __device__ __forceinline__ void vrr_grp(const R* __restrict__ F, "
                "R* __restrict__ acc,")
    code.append("        R PAx, R PAy, R PAz, R WPx, R WPy, R WPz, R QCx, R QCy, R QCz,")
    code.append("        R WQx, R WQy, R WQz, R inv2p, R inv2q, R inv2pq, R rho_p, R rho_q) {")
    for g in range(G):
        code.append("        %sif constexpr (G == %d) vrr_g%d(%s);" % ("" if g == 0 else "else ", g, g, args))
    code.append("    }")
    code.append("    // workspace value e * NCD + cd = [e | c d) for the bra components e of group G")
    code.append("    template <int G, class W> static __device__ __forceinline__ void hrr_ket_grp(const R* __restrict__ acc, "
                "const W& ws, R CDx, R CDy, R CDz) {")
    for g in range(G):
        code.append("        %sif constexpr (G == %d) hrr_ket_g%d(acc, ws, CDx, CDy, CDz);" % (
            "" if g == 0 else "else ", g, g))
    code.append("    }")
    # bra transfer rows from the workspace: ws[(e * NCD + cd) * WS_STRIDE]
    A_, B_ = cart(la), cart(lb)
    nB = len(B_)
    cases = []
    flops_bra = 0
    for ia, a0 in enumerate(A_):
        for ib, b0 in enumerate(B_):
            ex = bra_expansion(a0, b0)
            terms = []
            pre = []
            for e in sorted(ex, key=lambda t: eidx[t]):
                for mono, c in sorted(ex[e].items()):
                    if c == 0:
                        continue
                    terms.append((eidx[e], mono, c))
            body = []
            names = {}
            for (ei, mono, c) in terms:
                if (mono, c) not in names and not (sum(mono) == 0 and c == 1):
                    nm = "c%d" % len(names)
                    names[(mono, c)] = nm
                    expr = mono_expr(mono)
                    pre.append("const R %s = (R)%d * %s;" % (nm, c, expr) if c != 1 else "const R %s = %s;" % (nm, expr))
            # The row's workspace values are loaded in batches of about
            # BRA_BATCH values (CB kets cd x the row's terms), each batch
            # issued before any of it is used: left to itself the scheduler
            # of these register-bound kernels waits for every load before
            # issuing the next, one memory latency per value.
            nt = len(terms)
            cb = max(1, min(NCD, BRA_BATCH // nt))
            sum_terms = []
            for t, (ei, mono, c) in enumerate(terms):
                w = "w[%d][j]" % t
                sum_terms.append(w if (sum(mono) == 0 and c == 1) else "%s * %s" % (names[(mono, c)], w))
            flops_bra += NCD * (2 * len(terms) - 1)
            body += pre
            body.append("#pragma unroll")
            body.append("for (int cq = 0; cq < NCD; cq += %d) {" % cb)
            body.append("    R w[%d][%d];" % (nt, cb))
            body.append("#pragma unroll")
            body.append("    for (int j = 0; j < %d; ++j) {" % cb)
            body.append("        if (cq + j < NCD) {")
            for t, (ei, mono, c) in enumerate(terms):
                body.append("            w[%d][j] = ws.ld(%d * NCD + cq + j);" % (t, ei))
            body.append("        }")
            body.append("    }")
            body.append("    __builtin_amdgcn_sched_barrier(0);")
            body.append("#pragma unroll")
            body.append("    for (int j = 0; j < %d; ++j)" % cb)
            body.append("        if (cq + j < NCD) out[cq + j] = %s;" % " + ".join(sum_terms))
            body.append("}")
            cases.append("        case %d: {\n%s\n            break;\n        }" % (
                ia * nB + ib, "\n".join("            " + l for l in body)))
    code.append("    // out[cd] = (a b | c d) for bra components (ia, ib), iab = ia * NB + ib, from the")
    code.append("    // workspace values e * NCD + cd = [e | c d)")
    code.append("    template <class W> static __device__ __forceinline__ void hrr_bra_ab(int iab, "
                "const W& ws, R* __restrict__ out, R ABx, R ABy, R ABz) {")
    code.append("        switch (iab) {")
    code += cases
    code.append("        default: break;")
    code.append("        }")
    code.append("    }")
    return code, flops_vrr_exec, flops_bra


VRR_INPUTS = ["PAx", "PAy", "PAz", "WPx", "WPy", "WPz", "QCx", "QCy", "QCz",
              "WQx", "WQy", "WQz", "inv2p", "inv2q", "inv2pq", "rho_p", "rho_q"]


def vrr_dag(targets):
    """The vertical recurrence of gen_vrr as a graph: nodes [(name, terms)] in
    evaluation order and {(e, f): name} for the m = 0 targets.  Terms:
      ("lin", coef, x)            coef * x
      ("ho", k, inv, rho, c, d)   k * inv * (c - rho * d)
      ("tr", k, g)                k * inv2pq * g"""
    nodes = []
    memo = {}

    def V(e, f, m):
        key = (e, f, m)
        if key in memo:
            return memo[key]
        if sum(e) == 0 and sum(f) == 0:
            memo[key] = "F[%d]" % m
            return memo[key]
        terms = []
        if sum(f) > 0:
            i = pick_axis(f)
            f1 = sub1(f, i)
            terms.append(("lin", "QC%s" % AX[i], V(e, f1, m)))
            terms.append(("lin", "WQ%s" % AX[i], V(e, f1, m + 1)))
            if f1[i] > 0:
                f2 = sub1(f1, i)
                terms.append(("ho", f1[i], "inv2q", "rho_q", V(e, f2, m), V(e, f2, m + 1)))
            if e[i] > 0:
                terms.append(("tr", e[i], V(sub1(e, i), f1, m + 1)))
        else:
            i = pick_axis(e)
            e1 = sub1(e, i)
            terms.append(("lin", "PA%s" % AX[i], V(e1, f, m)))
            terms.append(("lin", "WP%s" % AX[i], V(e1, f, m + 1)))
            if e1[i] > 0:
                e2 = sub1(e1, i)
                terms.append(("ho", e1[i], "inv2p", "rho_p", V(e2, f, m), V(e2, f, m + 1)))
        name = "v%d" % (len(nodes) + 1)
        nodes.append((name, terms))
        memo[key] = name
        return name

    out = {}
    for (e, f) in targets:
        out[(e, f)] = V(e, f, 0)
    return nodes, out


def _adj(x):
    if x.startswith("F["):
        return "Fb[%s" % x[2:]
    if x.startswith("acc["):
        return "accb[%s" % x[4:]
    if x in VRR_INPUTS or x in ("ABx", "ABy", "ABz", "CDx", "CDy", "CDz"):
        return "b_" + x
    return "b" + x


def gen_adjoint(la, lb, lc, ld):
    """EriAdj<la, lb, lc, ld> (see the module docstring) and its flop counts."""
    es = [e for L in range(la, la + lb + 1) for e in cart(L)]
    fs = [f for L in range(lc, lc + ld + 1) for f in cart(L)]
    eidx = {e: n for n, e in enumerate(es)}
    fidx = {f: n for n, f in enumerate(fs)}
    nF = len(fs)
    targets = [(e, f) for e in es for f in fs]
    nodes, out = vrr_dag(targets)
    Lt = la + lb + lc + ld
    fwd, rev = [], []
    flops = 0
    for name, terms in nodes:
        ex = []
        for t in terms:
            if t[0] == "lin":
                ex.append("%s * %s" % (t[1], t[2]))
            elif t[0] == "ho":
                ex.append("%d.0 * %s * (%s - %s * %s)" % (t[1], t[2], t[4], t[3], t[5]))
            else:
                ex.append("%d.0 * inv2pq * %s" % (t[1], t[2]))
        fwd.append("const double %s = %s;" % (name, " + ".join(ex)))
    decl = ["double %s = 0.0;" % _adj(n) for n, _ in nodes]
    decl += ["double b_%s = 0.0;" % x for x in VRR_INPUTS]
    seed = []
    lsum = []
    for k, t in enumerate(targets):
        seed.append("%s += g[%d];" % (_adj(out[t]), k))
        lsum.append("g[%d] * %s" % (k, out[t]))
    for name, terms in reversed(nodes):
        b = _adj(name)
        for t in terms:
            if t[0] == "lin":
                rev.append("%s += %s * %s;" % (_adj(t[2]), t[1], b))
                rev.append("b_%s += %s * %s;" % (t[1], t[2], b))
                flops += 4
            elif t[0] == "ho":
                k, inv, rho, c, d = t[1:]
                rev.append("%s += %d.0 * %s * %s;" % (_adj(c), k, inv, b))
                rev.append("%s -= %d.0 * %s * %s * %s;" % (_adj(d), k, inv, rho, b))
                rev.append("b_%s += %d.0 * (%s - %s * %s) * %s;" % (inv, k, c, rho, d, b))
                rev.append("b_%s -= %d.0 * %s * %s * %s;" % (rho, k, inv, d, b))
                flops += 14
            else:
                rev.append("%s += %d.0 * inv2pq * %s;" % (_adj(t[2]), t[1], b))
                rev.append("b_inv2pq += %d.0 * %s * %s;" % (t[1], t[2], b))
                flops += 6
    code = []
    code.append("template <> struct EriAdj<%d, %d, %d, %d> {" % (la, lb, lc, ld))
    code.append("    static constexpr int NACC = %d;" % len(targets))
    code.append("    // forward recurrence of one primitive quartet, then its adjoint with seeds")
    code.append("    // g[NACC] on [e0|f0]: Fb[m] = d/dF_m, bin[] = d/d(%s)," % ", ".join(VRR_INPUTS[:4]) + " ...")
    code.append("    // in the order PA, WP, QC, WQ (x, y, z each), inv2p, inv2q, inv2pq, rho_p, rho_q;")
    code.append("    // returns sum_k g[k] [e0|f0]_k")
    code.append("    static __device__ __forceinline__ double vrr_adj(const double* __restrict__ F, "
                "const double* __restrict__ g,")
    code.append("        double PAx, double PAy, double PAz, double WPx, double WPy, double WPz,")
    code.append("        double QCx, double QCy, double QCz, double WQx, double WQy, double WQz,")
    code.append("        double inv2p, double inv2q, double inv2pq, double rho_p, double rho_q,")
    code.append("        double* __restrict__ Fb, double* __restrict__ bin) {")
    code += ["        " + l for l in fwd]
    code.append("#pragma unroll")
    code.append("        for (int m = 0; m <= %d; ++m) Fb[m] = 0.0;" % Lt)
    code += ["        " + l for l in decl + seed + rev]
    code.append("        " + " ".join("bin[%d] = b_%s;" % (k, x) for k, x in enumerate(VRR_INPUTS)))
    code.append("        return %s;" % " + ".join(lsum))
    code.append("    }")

    # full horizontal recurrence (ket, then bra) as a graph of n = x + coef * y
    C_, D_, A_, B_ = cart(lc), cart(ld), cart(la), cart(lb)
    nC, nD, nB = len(C_), len(D_), len(B_)
    NCD = nC * nD
    hn = []
    kmemo, bmemo = {}, {}

    def K(e, c, d):
        key = (e, c, d)
        if key in kmemo:
            return kmemo[key]
        if sum(d) == 0:
            r = "acc[%d]" % (eidx[e] * nF + fidx[c])
            kmemo[key] = r
            return r
        i = pick_axis(d)
        d1 = sub1(d, i)
        x = K(e, add1(c, i), d1)
        y = K(e, c, d1)
        name = "k%d" % (len(hn) + 1)
        hn.append((name, x, "CD%s" % AX[i], y))
        kmemo[key] = name
        return name

    def Bv(a, b, c, d):
        key = (a, b, c, d)
        if key in bmemo:
            return bmemo[key]
        if sum(b) == 0:
            r = K(a, c, d)
            bmemo[key] = r
            return r
        i = pick_axis(b)
        b1 = sub1(b, i)
        x = Bv(add1(a, i), b1, c, d)
        y = Bv(a, b1, c, d)
        name = "h%d" % (len(hn) + 1)
        hn.append((name, x, "AB%s" % AX[i], y))
        bmemo[key] = name
        return name

    outs = []
    for a0 in A_:
        for b0 in B_:
            for c0 in C_:
                for d0 in D_:
                    outs.append(Bv(a0, b0, c0, d0))
    hf = ["const double %s = %s + %s * %s;" % (n, x, c, y) for (n, x, c, y) in hn]
    hdecl = ["double %s = 0.0;" % _adj(n) for (n, _, _, _) in hn]
    hdecl += ["double b_%s = 0.0;" % x for x in ("ABx", "ABy", "ABz", "CDx", "CDy", "CDz")]
    hseed = ["%s += G[%d];" % (_adj(o), k) for k, o in enumerate(outs)]
    hrev = []
    hflops = 0
    for (n, x, c, y) in reversed(hn):
        b = _adj(n)
        hrev.append("%s += %s;" % (_adj(x), b))
        hrev.append("%s += %s * %s;" % (_adj(y), c, b))
        hrev.append("b_%s += %s * %s;" % (c, y, b))
        hflops += 5
    code.append("    static constexpr int NINT = %d;" % len(outs))
    code.append("    // adjoint of the horizontal recurrence: seeds G[NINT] on the contracted (ab|cd)")
    code.append("    // (index ((ia * NB + ib) * NC + ic) * ND + id) -> accb[NACC] on the contracted [e0|f0],")
    code.append("    // bAB[3], bCD[3] = derivatives with respect to A-B and C-D; returns sum_k G[k] (ab|cd)_k")
    code.append("    static __device__ __forceinline__ double hrr_adj(const double* __restrict__ acc, "
                "const double* __restrict__ G,")
    code.append("        double* __restrict__ accb, double* __restrict__ bAB, double* __restrict__ bCD,")
    code.append("        double ABx, double ABy, double ABz, double CDx, double CDy, double CDz) {")
    code += ["        " + l for l in hf]
    code.append("#pragma unroll")
    code.append("        for (int n = 0; n < NACC; ++n) accb[n] = 0.0;")
    code += ["        " + l for l in hdecl + hseed + hrev]
    code.append("        bAB[0] = b_ABx; bAB[1] = b_ABy; bAB[2] = b_ABz;")
    code.append("        bCD[0] = b_CDx; bCD[1] = b_CDy; bCD[2] = b_CDz;")
    code.append("        return %s;" % " + ".join("G[%d] * %s" % (k, o) for k, o in enumerate(outs)))
    code.append("    }")
    code.append("};")
    return "\n".join(code), flops, hflops


def stage_ws(la, lb, lc, ld):
    """[e|cd) values per lane that a staged class keeps in its workspace."""
    if (la, lb, lc, ld) not in STAGED:
        return 0
    nE = sum(ncart(L) for L in range(la, la + lb + 1))
    return nE * ncart(lc) * ncart(ld)


def class_list():
    """ERI classes over the pair classes up to LMAX, and every pair class up
    to PAIR_LMAX (those up to LMAX first, in the same order)."""
    low = [(a, b) for a in range(LMAX + 1) for b in range(a + 1)]
    order = {c: n for n, c in enumerate(low)}
    res = []
    for X in low:
        for Y in low:
            if order[X] >= order[Y]:
                res.append((X[0], X[1], Y[0], Y[1]))
    pcl = low + [(a, b) for a in range(LMAX + 1, PAIR_LMAX + 1) for b in range(a + 1)]
    return res, pcl


def gen_class(la, lb, lc, ld):
    """Return (code, flops_prim, flops_contracted, n_acc)."""
    Lt = la + lb + lc + ld
    es = [e for L in range(la, la + lb + 1) for e in cart(L)]
    fs = [f for L in range(lc, lc + ld + 1) for f in cart(L)]
    eidx = {e: n for n, e in enumerate(es)}
    fidx = {f: n for n, f in enumerate(fs)}
    nE, nF = len(es), len(fs)

    # ---- contraction step (VRR inside the primitive loops) ----
    em = Emitter()
    targets = [(e, f) for e in es for f in fs]
    vv = gen_vrr(em, targets)
    acc_lines = []
    for (e, f) in targets:
        acc_lines.append("acc[%d] += %s;" % (eidx[e] * nF + fidx[f], vv[(e, f)]))
    flops_vrr = em.flops + len(targets)
    vrr_code = "\n".join("        " + l for l in em.lines + acc_lines)

    # ---- HRR, ket: [e | c d) from [e0|f0] for every e; kacc[e * NCD + cd] ----
    C_, D_ = cart(lc), cart(ld)
    nC, nD = len(C_), len(D_)
    NCD = nC * nD
    hm = Emitter()
    ket_memo = {}

    def K(e, c, d):  # [e | c d) with e an index into es
        key = (e, c, d)
        if key in ket_memo:
            return ket_memo[key]
        if sum(d) == 0:
            r = "acc[%d]" % (e * nF + fidx[c])
            ket_memo[key] = r
            return r
        i = pick_axis(d)
        d1 = sub1(d, i)
        a = K(e, add1(c, i), d1)
        b = K(e, c, d1)
        name = hm.fresh("k")
        hm.lines.append("const R %s = %s + CD%s * %s;" % (name, a, AX[i], b))
        hm.flops += 2
        ket_memo[key] = name
        return name

    kouts = []
    for ie, e in enumerate(es):
        for ic, c in enumerate(C_):
            for idd, d in enumerate(D_):
                kouts.append("kacc[%d] = %s;" % (ie * NCD + ic * nD + idd, K(ie, c, d)))
    ket_code = "\n".join("        " + l for l in hm.lines + kouts)
    flops_hrr = hm.flops

    # ---- HRR, bra, one Cartesian component a of shell A at a time:
    #      out[ib * NCD + cd] = (a b | cd) from kacc ----
    A_, B_ = cart(la), cart(lb)
    nB = len(B_)
    cases = []
    sum_bra_flops = [0]
    for ia, a0 in enumerate(A_):
        bm = Emitter()
        memo = {}

        def Bv(a, b, cd):
            key = (a, b, cd)
            if key in memo:
                return memo[key]
            if sum(b) == 0:
                r = "kacc[%d]" % (eidx[a] * NCD + cd)
                memo[key] = r
                return r
            i = pick_axis(b)
            b1 = sub1(b, i)
            x = Bv(add1(a, i), b1, cd)
            y = Bv(a, b1, cd)
            name = bm.fresh("h")
            bm.lines.append("const R %s = %s + AB%s * %s;" % (name, x, AX[i], y))
            bm.flops += 2
            memo[key] = name
            return name

        outs = []
        for ib, b0 in enumerate(B_):
            for cd in range(NCD):
                outs.append("out[%d] = %s;" % (ib * NCD + cd, Bv(a0, b0, cd)))
        flops_hrr += bm.flops
        sum_bra_flops[0] += bm.flops
        body = "\n".join("            " + l for l in bm.lines + outs)
        cases.append("        case %d: {\n%s\n            break;\n        }" % (ia, body))
    bra_code = "        switch (ia) {\n" + "\n".join(cases) + "\n        default: break;\n        }"

    tag = "%d%d%d%d" % (la, lb, lc, ld)
    code = []
    code.append("template <typename R> struct EriClass<%d, %d, %d, %d, R> {" % (la, lb, lc, ld))
    code.append("    static constexpr int NACC = %d;" % (nE * nF))
    code.append("    static constexpr int NE = %d;" % nE)
    code.append("    static constexpr int LT = %d;" % Lt)
    code.append("    // FP64 K-row J/K kernel compiled for two waves per SIMD")
    code.append("    static constexpr bool ROWS_TWO_WAVES = %s;" % (
        "std::is_same<R, double>::value" if (la, lb, lc, ld) in ROWS_TWO_WAVES else "false"))
    flops_exec = None
    if (la, lb, lc, ld) in STAGED:
        scode, fve, fbra = gen_staged(la, lb, lc, ld, es, fs, eidx, fidx, C_, D_)
        code += scode
        flops_exec = (fve, flops_hrr - sum_bra_flops[0] + fbra)
    else:
        code.append("    static constexpr bool STAGED = false;")
        code.append("    static constexpr int SPLIT = 1;")
    code.append("    static __device__ __forceinline__ void vrr(const R* __restrict__ F, R* __restrict__ acc,")
    code.append("        R PAx, R PAy, R PAz, R WPx, R WPy, R WPz, R QCx, R QCy, R QCz,")
    code.append("        R WQx, R WQy, R WQz, R inv2p, R inv2q, R inv2pq, R rho_p, R rho_q) {")
    code.append(vrr_code)
    code.append("    }")
    code.append("    // kacc[e * NCD + cd] = [e | c d)")
    code.append("    static __device__ __forceinline__ void hrr_ket(const R* __restrict__ acc, R* __restrict__ kacc,")
    code.append("        R CDx, R CDy, R CDz) {")
    code.append(ket_code)
    code.append("    }")
    code.append("    // out[ib * NCD + cd] = (a b | c d) for Cartesian component ia of shell A")
    code.append("    static __device__ __forceinline__ void hrr_bra(int ia, const R* __restrict__ kacc, R* __restrict__ out,")
    code.append("        R ABx, R ABy, R ABz) {")
    code.append(bra_code)
    code.append("    }")
    code.append("};")
    return "\n".join(code), flops_vrr, flops_hrr, nE * nF, flops_exec


def main(outdir):
    os.makedirs(outdir, exist_ok=True)
    classes, pcl = class_list()
    flops = {}
    twin_ok = []
    for (la, lb, lc, ld) in classes:
        code, fv, fh, nacc, fexec = gen_class(la, lb, lc, ld)
        twin_ok.append(fexec is None and nacc <= TWIN_MAX_ACC)
        tag = "%d%d%d%d" % (la, lb, lc, ld)
        flops[tag] = {"vrr_per_prim_quartet": fv, "hrr_per_quartet": fh, "n_acc": nacc,
                      "staged": fexec is not None}
        if fexec is not None:
            # executed counts of the grouped evaluation (the VRR intermediates
            # shared between groups are recomputed); the figures above stay the
            # algorithmic counts of the ungrouped recurrence
            flops[tag]["vrr_executed_per_prim_quartet"] = fexec[0]
            flops[tag]["hrr_executed_per_quartet"] = fexec[1]
        variants = [("double", "MODE_JK", "jk64"), ("float", "MODE_JK", "jk32"), ("double", "MODE_DENSE", "dense")]
        if (la, lb) == (lc, ld):
            variants.append(("double", "MODE_SCHWARZ", "schwarz"))
        if max(la, lb, lc, ld) <= GRAD_LMAX:
            acode, af, ah = gen_adjoint(la, lb, lc, ld)
            flops[tag]["adjoint_vrr_per_prim_quartet"] = af
            flops[tag]["adjoint_hrr_per_quartet"] = ah
            tu = ["// Generated by gen_eri.py -- do not edit.",
                  "// ERI class (%d %d | %d %d): adjoint recurrences and gradient kernel instantiation."
                  % (la, lb, lc, ld),
                  '#include "qbx_grad.h"', "", code, "", acode, "", "namespace qbx {",
                  "template void launch_eri_grad<%d, %d, %d, %d>(const EriArgs&, hipStream_t);" % (la, lb, lc, ld),
                  "}  // namespace qbx"]
            with open(os.path.join(outdir, "eri_%s_grad.hip" % tag), "w") as f:
                f.write("\n".join(tu) + "\n")
        for R, mode, vtag in variants:
            tu = ["// Generated by gen_eri.py -- do not edit.",
                  "// ERI class (%d %d | %d %d), %s %s: recurrence code and kernel instantiation."
                  % (la, lb, lc, ld, R, mode),
                  '#include "qbx_eri.h"', "", code, "", "namespace qbx {",
                  "template void launch_eri<%d, %d, %d, %d, %s, %s>(const EriArgs&, hipStream_t);"
                  % (la, lb, lc, ld, R, mode),
                  "}  // namespace qbx"]
            with open(os.path.join(outdir, "eri_%s_%s.hip" % (tag, vtag)), "w") as f:
                f.write("\n".join(tu) + "\n")
    # class table for the host dispatcher
    lines = ["// Generated by gen_eri.py -- do not edit.", "#pragma once",
             "#define QBX_GEN_LMAX %d" % LMAX,
             "#define QBX_N_PAIR_CLASSES %d" % len(pcl),
             "// pair classes 0 .. QBX_N_LOW_PAIR_CLASSES-1 (both shells up to QBX_GEN_LMAX) form the"
             " generated ERI classes",
             "#define QBX_N_LOW_PAIR_CLASSES %d" % ((LMAX + 1) * (LMAX + 2) // 2),
             "#define QBX_PAIR_LMAX %d" % PAIR_LMAX,
             "#define QBX_N_CLASSES %d" % len(classes),
             "static const int qbx_pair_class_l[%d][2] = {%s};" % (
                 len(pcl), ", ".join("{%d, %d}" % c for c in pcl)),
             "static const int qbx_class_l[%d][4] = {%s};" % (
                 len(classes), ", ".join("{%d, %d, %d, %d}" % c for c in classes)),
             "// classes with a gradient kernel (all shells s or p)",
             "static const int qbx_class_grad[%d] = {%s};" % (
                 len(classes), ", ".join("1" if max(c) <= GRAD_LMAX else "0" for c in classes)),
             "// workspace values per lane of the staged classes (0: not staged)",
             "static const int qbx_class_ws[%d] = {%s};" % (
                 len(classes), ", ".join(str(stage_ws(*c)) for c in classes)),
             "// staged classes whose 64-ket chunks are shared by the waves of a block (waves per block; 1: not shared)",
             "static const int qbx_class_split[%d] = {%s};" % (
                 len(classes), ", ".join(str(STAGE_SPLIT.get(c, 1) if c in STAGED else 1) for c in classes)),
             "// chunks per K-row task of the classes whose chunks are shared by a block's waves (0: not shared)",
             "static const int qbx_class_split_chunks[%d] = {%s};" % (
                 len(classes), ", ".join(str(STAGE_SPLIT_CHUNKS[c] if c in STAGED and STAGE_SPLIT.get(c, 1) > 1
                                             else 0) for c in classes)),
             "// classes launched after all others in a Fock build",
             "static const int qbx_class_last[%d] = {%s};" % (
                 len(classes), ", ".join("1" if c in LAUNCH_LAST else "0" for c in classes)),
             "// classes with a twin-bra kernel (not staged, at most %d accumulators)" % TWIN_MAX_ACC,
             "static const int qbx_class_twin[%d] = {%s};" % (
                 len(classes), ", ".join("1" if t else "0" for t in twin_ok)),
             "// launch-order cost model: per primitive shell quartet (geometry and Boys"
             " values, about %d operations, plus the executed VRR) and per contracted"
             " shell quartet (executed HRR plus %d per Cartesian quartet for the digest)" % (COST_PRIM0, COST_DIGEST),
             "static const double qbx_class_cost_prim[%d] = {%s};" % (
                 len(classes), ", ".join(str(COST_PRIM0 + flops[t].get("vrr_executed_per_prim_quartet",
                                                                      flops[t]["vrr_per_prim_quartet"]))
                                         for t in ["%d%d%d%d" % c for c in classes])),
             "static const double qbx_class_cost_quartet[%d] = {%s};" % (
                 len(classes), ", ".join(str(flops[t].get("hrr_executed_per_quartet", flops[t]["hrr_per_quartet"])
                                             + COST_DIGEST * ncart_prod(c))
                                         for t, c in zip(["%d%d%d%d" % c for c in classes], classes))),
             "#define QBX_FOR_EACH_CLASS(X) \\"]
    lines += ["    X(%d, %d, %d, %d) \\" % c for c in classes]
    lines.append("")
    with open(os.path.join(outdir, "eri_table.inc"), "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(os.path.join(outdir, "eri_flops.json"), "w") as f:
        json.dump(flops, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(__file__), "gen"))
