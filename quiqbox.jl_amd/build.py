"""Build libqbx_hip.so for gfx950 in-tree.

Steps: run csrc/gen_eri.py (one translation unit per ERI angular class into
csrc/gen/), compile csrc/qbx.hip, csrc/qbx_mlp.hip, csrc/qbx_linalg.hip, csrc/qbx_boys_probe.hip, csrc/qbx_eri_generic.hip and every
csrc/gen/eri_*.hip to objects in
parallel with hipcc, link quiqbox_amd/libqbx_hip.so.  Objects whose sources and
headers are unchanged are reused.
"""
import concurrent.futures as cf
import glob
import hashlib
import os
import subprocess
import sys
import time

PKG_ROOT = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_ROOT, "csrc")
GEN = os.path.join(CSRC, "gen")
OBJ = os.path.join(PKG_ROOT, "_build")
OUT = os.path.join(PKG_ROOT, "quiqbox_amd", "libqbx_hip.so")
INCLUDE = os.path.join(os.path.dirname(PKG_ROOT), "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("QBX_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC",
    "--offload-arch=" + ARCH,
    "-munsafe-fp-atomics",
    # packed FP32 instructions stay off for all HIP code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-pass-failed",
    "-I" + INCLUDE, "-I" + CSRC,
]


def source_digest():
    """sha256 (first 16 hex digits) over the sources the library is built
    from: csrc/*.h, csrc/*.hip (not the generated ones, which follow from
    gen_eri.py), csrc/gen_eri.py, include/qbx.h and this script.  The library
    reports the digest it was built from (qbx_source_digest), so a run can
    tell whether it loaded a library built from the sources beside it."""
    paths = sorted(glob.glob(os.path.join(CSRC, "*.h")) + glob.glob(os.path.join(CSRC, "*.hip")) +
                   [os.path.join(CSRC, "gen_eri.py"), os.path.join(INCLUDE, "qbx.h"), os.path.abspath(__file__)])
    h = hashlib.sha256()
    for p in paths:
        h.update(os.path.relpath(p, os.path.dirname(PKG_ROOT)).encode() + b"\0")
        with open(p, "rb") as f:
            h.update(f.read())
        h.update(b"\0")
    return h.hexdigest()[:16]


def generate():
    for old in glob.glob(os.path.join(GEN, "eri_*.hip")):
        os.remove(old)
    subprocess.run([sys.executable, os.path.join(CSRC, "gen_eri.py"), GEN], check=True)


def _digest(paths, extra):
    h = hashlib.sha256()
    for p in paths:
        with open(p, "rb") as f:
            h.update(f.read())
    h.update(" ".join(FLAGS + list(extra)).encode())
    return h.hexdigest()[:16]


def _compile(src, extra, headers, verbose, force):
    os.makedirs(OBJ, exist_ok=True)
    tag = _digest([src] + headers, extra)
    obj = os.path.join(OBJ, os.path.basename(src).replace(".hip", "") + "." + tag + ".o")
    if os.path.exists(obj) and not force:
        return obj, 0.0, ""
    t0 = time.time()
    cmd = [HIPCC] + FLAGS + list(extra) + ["-c", src, "-o", obj + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, r.stderr[-6000:]))
    os.replace(obj + ".tmp", obj)
    msg = r.stderr if verbose else ""
    return obj, time.time() - t0, msg


def build(verbose=True, extra=(), jobs=None, force=False):
    """Generate the ERI sources and link libqbx_hip.so.  Objects are cached by
    a digest of their sources, headers and flags; force=True compiles every
    source regardless."""
    t0 = time.time()
    generate()
    headers = sorted(glob.glob(os.path.join(CSRC, "*.h")) + [os.path.join(GEN, "eri_table.inc"),
                                                              os.path.join(INCLUDE, "qbx.h")])
    srcs = [os.path.join(CSRC, "qbx.hip"), os.path.join(CSRC, "qbx_mlp.hip"),
            os.path.join(CSRC, "qbx_linalg.hip"), os.path.join(CSRC, "qbx_boys_probe.hip"),
            os.path.join(CSRC, "qbx_eri_generic.hip")] + sorted(glob.glob(os.path.join(GEN, "eri_*.hip")),
                                                     key=lambda p: -os.path.getsize(p))
    jobs = jobs or max(1, min(len(srcs), os.cpu_count() or 4))
    objs = {}
    digest_flag = ['-DQBX_SOURCE_DIGEST="%s"' % source_digest()]
    with cf.ThreadPoolExecutor(jobs) as ex:
        futs = {ex.submit(_compile, s, list(extra) + (digest_flag if s.endswith("qbx.hip") else []), headers,
                          verbose, force): s for s in srcs}
        for fu in cf.as_completed(futs):
            obj, dt, msg = fu.result()
            objs[futs[fu]] = obj
            if verbose and dt > 0:
                print("[qbx build] %-14s %6.1f s" % (os.path.basename(futs[fu]), dt), flush=True)
            if msg.strip() and verbose:
                print(msg, flush=True)
    rocm_lib = os.path.join(os.path.dirname(os.path.dirname(os.path.realpath(HIPCC))), "lib")
    cmd = ([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", OUT + ".tmp"] + [objs[s] for s in srcs] +
           ["-L" + rocm_lib, "-Wl,-rpath," + rocm_lib, "-lrocsolver", "-lrocblas"])
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    if verbose:
        print("[qbx build] %s in %.1f s" % (OUT, time.time() - t0), flush=True)
    return OUT


if __name__ == "__main__":
    build(extra=sys.argv[1:])
