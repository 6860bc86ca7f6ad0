"""The generated adjoint recurrences (EriAdj::vrr_adj, EriAdj::hrr_adj) against
central finite differences of their own forward parts, evaluated on the CPU
by translating the emitted statements to Python.  The forward parts are
polynomials in their inputs, so the central differences are exact up to
O(h^2) truncation; a dropped, mis-signed or mis-indexed adjoint term moves a
derivative by O(1)."""
import os
import re
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "quiqbox.jl_amd", "csrc"))
import gen_eri  # noqa: E402

CLASSES = [c for c in gen_eri.class_list()[0] if max(c) <= gen_eri.GRAD_LMAX]


def _body(code, sig):
    i = code.index(sig)
    j = code.index("{\n", i) + 2
    k = code.index("\n    }\n", j)
    return code[j:k]


def _to_python(body):
    out = []
    for line in body.strip().splitlines():
        line = line.strip()
        if not line or line.startswith("#pragma"):
            continue
        m = re.match(r"for \(int (\w) = 0; \w <=? (\w+); \+\+\w\) (.*);$", line)
        if m:
            var, bound, stmt = m.groups()
            rng = "range(%s + 1)" % bound if "<=" in line else "range(%s)" % bound
            out.append("for %s in %s: %s" % (var, rng, stmt))
            continue
        if line.startswith("return "):
            out.append("__ret = " + line[7:].rstrip(";"))
            continue
        for stmt in line.rstrip(";").split(";"):
            stmt = stmt.strip().replace("const double ", "").replace("double ", "")
            if stmt:
                out.append(stmt)
    return compile("\n".join(out), "adj", "exec")


def _run(prog, env):
    env = dict(env)
    exec(prog, env)
    return env


@pytest.mark.parametrize("cls", CLASSES, ids=["".join(map(str, c)) for c in CLASSES])
def test_vrr_adjoint_matches_central_differences(cls):
    code = gen_eri.gen_adjoint(*cls)[0]
    prog = _to_python(_body(code, "double vrr_adj("))
    nacc = int(re.search(r"NACC = (\d+)", code).group(1))
    Lt = sum(cls)
    rng = np.random.default_rng(sum(cls) + 7)
    inputs = {k: float(rng.uniform(-1.0, 1.0)) for k in gen_eri.VRR_INPUTS}
    for k in ("inv2p", "inv2q", "inv2pq", "rho_p", "rho_q"):
        inputs[k] = float(rng.uniform(0.1, 0.9))
    F = [float(x) for x in rng.uniform(0.1, 1.0, Lt + 1)]
    g = [float(x) for x in rng.normal(size=nacc)]

    def L(F_, inp):
        env = dict(inp, F=list(F_), g=g, Fb=[0.0] * (Lt + 1), bin=[0.0] * len(gen_eri.VRR_INPUTS), NACC=nacc)
        return _run(prog, env)

    base = L(F, inputs)
    h = 1e-5
    for m in range(Lt + 1):
        Fp, Fm = list(F), list(F)
        Fp[m] += h
        Fm[m] -= h
        fd = (L(Fp, inputs)["__ret"] - L(Fm, inputs)["__ret"]) / (2 * h)
        assert base["Fb"][m] == pytest.approx(fd, rel=1e-7, abs=1e-9), ("F", m)
    for k, name in enumerate(gen_eri.VRR_INPUTS):
        ip, im = dict(inputs), dict(inputs)
        ip[name] += h
        im[name] -= h
        fd = (L(F, ip)["__ret"] - L(F, im)["__ret"]) / (2 * h)
        assert base["bin"][k] == pytest.approx(fd, rel=1e-7, abs=1e-9), name


@pytest.mark.parametrize("cls", CLASSES, ids=["".join(map(str, c)) for c in CLASSES])
def test_hrr_adjoint_matches_central_differences(cls):
    code = gen_eri.gen_adjoint(*cls)[0]
    prog = _to_python(_body(code, "double hrr_adj("))
    nacc = int(re.search(r"NACC = (\d+)", code).group(1))
    nint = int(re.search(r"NINT = (\d+)", code).group(1))
    rng = np.random.default_rng(sum(cls) + 11)
    acc = [float(x) for x in rng.normal(size=nacc)]
    G = [float(x) for x in rng.normal(size=nint)]
    geo = {k: float(rng.uniform(-1.5, 1.5)) for k in ("ABx", "ABy", "ABz", "CDx", "CDy", "CDz")}

    def L(acc_, geo_):
        env = dict(geo_, acc=list(acc_), G=G, accb=[0.0] * nacc, bAB=[0.0] * 3, bCD=[0.0] * 3, NACC=nacc)
        return _run(prog, env)

    base = L(acc, geo)
    h = 1e-5
    for n in range(nacc):
        ap, am = list(acc), list(acc)
        ap[n] += h
        am[n] -= h
        fd = (L(ap, geo)["__ret"] - L(am, geo)["__ret"]) / (2 * h)
        assert base["accb"][n] == pytest.approx(fd, rel=1e-7, abs=1e-9), ("acc", n)
    for k, name in enumerate(("ABx", "ABy", "ABz", "CDx", "CDy", "CDz")):
        gp, gm = dict(geo), dict(geo)
        gp[name] += h
        gm[name] -= h
        fd = (L(acc, gp)["__ret"] - L(acc, gm)["__ret"]) / (2 * h)
        got = base["bAB"][k] if k < 3 else base["bCD"][k - 3]
        assert got == pytest.approx(fd, rel=1e-7, abs=1e-9), name
