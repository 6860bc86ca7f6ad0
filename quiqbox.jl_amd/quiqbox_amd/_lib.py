"""ctypes binding of libqbx_hip.so (include/qbx.h).

The library is the only compute path: if it is missing or no GPU is present,
the calls raise; there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libqbx_hip.so")

QBX_SHELL_MAXC = 15     # Cartesian components of the largest shell (g), qbx.h
QBX_MAX_L = 4
QBX_PRIM_LMAX = 10          # qbx_prim_one_e, qbx_prim_eri: total Cartesian exponent per primitive
QBX_PRIM_ERI_LTOT = 24      # qbx_prim_eri: total over the quartet
QBX_PRIM_OVERLAP, QBX_PRIM_KINETIC, QBX_PRIM_COULOMB, QBX_PRIM_MULTIPOLE = 0, 1, 2, 3
QBX_PREC_FP64 = 0
QBX_PREC_MIXED = 1

_d = ctypes.c_double
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_pd = ctypes.POINTER(ctypes.c_double)
_pf = ctypes.POINTER(ctypes.c_float)
_pi = ctypes.POINTER(ctypes.c_int32)
_pl = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p

# every symbol declared in include/qbx.h with (restype, argtypes)
SIGNATURES = {
    "qbx_init": (_i32, [_i32, ctypes.POINTER(_vp)]),
    "qbx_finalize": (None, [_vp]),
    "qbx_basis_create": (_i32, [_vp, _i32, _pd, _pd, _pi, _i32, _pi, _pi, _pd, ctypes.POINTER(_vp)]),
    "qbx_basis_destroy": (None, [_vp]),
    "qbx_basis_info": (_i32, [_vp, _pi, _pi, _pi, _pi]),
    "qbx_overlap": (_i32, [_vp, _vp, _pd]),
    "qbx_kinetic": (_i32, [_vp, _vp, _pd]),
    "qbx_nuclear": (_i32, [_vp, _vp, _i32, _pd, _pd, _pd]),
    "qbx_nuclear_own": (_i32, [_vp, _vp, _i32, _pd, _pd, _pd]),
    "qbx_core_hamiltonian": (_i32, [_vp, _vp, _i32, _pd, _pd, _pd]),
    "qbx_multipole": (_i32, [_vp, _vp, _pd, _pi, _pd]),
    "qbx_eri_dense": (_i32, [_vp, _vp, _pd]),
    "qbx_jk": (_i32, [_vp, _vp, _i32, _pd, _pd, _pd, _d, _i32]),
    "qbx_jk_device": (_i32, [_vp, _vp, _i32, _vp, _vp, _vp, _d, _i32, _vp]),
    "qbx_n_classes": (_i32, []),
    "qbx_class_table": (_i32, [_pi]),
    "qbx_jk_stats": (_i32, [_vp, _vp, _d, _d, _pl, _pl, _pl, _pl]),
    "qbx_set_timing": (_i32, [_vp, _i32]),
    "qbx_last_timing": (_i32, [_vp, _pf, _pf]),
    "qbx_set_mixed_cut": (_i32, [_vp, _d]),
    "qbx_set_density_screening": (_i32, [_vp, _d]),
    "qbx_boys_device": (_i32, [_vp, _i32, _i32, _pd, _pd]),
    "qbx_prim_one_e": (_i32, [_vp, _i32, _i32, _pd, _pd, _pi, _pd, _pd, _pi, _pd, _pi, _pd]),
    "qbx_prim_eri": (_i32, [_vp, _i32, _pd, _pd, _pi, _pd]),
    "qbx_basis_shells": (_i32, [_vp, _pi, _pi, _pi, _pi, _pi, _pi, _pd, _pd, _pd, _pd]),
    "qbx_rhf_gradient": (_i32, [_vp, _vp, _pd, _pd, _i32, _pd, _pd, _d, _pd, _pd, _pd]),
    "qbx_uhf_gradient": (_i32, [_vp, _vp, _pd, _pd, _pd, _pd, _i32, _pd, _pd, _d, _pd, _pd, _pd]),
    "qbx_source_digest": (ctypes.c_char_p, []),
    "qbx_density_purify": (_i32, [_vp, _i32, _i32, _pd, _pd, _pd, _i32, _pi, _pd]),
    "qbx_fock_solve": (_i32, [_vp, _i32, _i32, _pd, _pd, _pd, _pd, _pd]),
    "qbx_density_purify_noalloc": (_i32, [_vp, _i32, _i32, _vp, _vp, _vp, _i32, _pi, _pd, _vp]),
    "qbx_mo_transform": (_i32, [_vp, _i32, _i32, _i32, _pd, _pd, _pd]),
    "qbx_mo_coulomb_ab": (_i32, [_vp, _i32, _i32, _i32, _pd, _pd, _pd, _pd]),
    "qbx_gemm_f64": (_i32, [_i32, _i32, _i32, _i32, _i32, _d, _vp, _i32, _vp, _i32, _d, _vp, _i32, _vp]),
    "qbx_gemm_bf16": (_i32, [_i32, _i32, _i32, _i32, _i32, ctypes.c_float, _vp, _i32, _vp, _i32, ctypes.c_float,
                             _vp, _i32, _i32, _vp, _vp, _vp]),
    "qbx_mlp_elementwise": (_i32, [_i32, _i64, _vp, _vp, _vp, _vp, _vp]),
    "qbx_mlp_mse": (_i32, [_i32, _i32, _vp, _vp, _vp, _vp, ctypes.c_float, _vp, _vp, _vp, _vp]),
    "qbx_mlp_colsum": (_i32, [_i32, _i32, _vp, _vp, _vp]),
    "qbx_mlp_adam": (_i32, [_i64, _vp, _vp, _vp, _vp, _vp, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                            ctypes.c_float, _i32, _vp]),
    "qbx_last_error": (_i32, [ctypes.c_char_p, ctypes.c_size_t]),
}

_lib = None


class QbxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (qbx code %d)" % (msg, code))
        self.code = code


def load(path=LIB_PATH):
    """Load libqbx_hip.so and attach the C-ABI signatures.  Raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libqbx_hip.so is not built (expected at %s); run __graft_entry__.build()" % path)
    # One HIP runtime per process.  torch (device buffers of the learned guess)
    # ships its own libamdhip64 and makes its symbols global when imported; a
    # library loaded after that binds to the same runtime.  Loaded the other
    # way round, the process holds two runtimes and torch finds no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def provenance():
    """The library's source digest (qbx_source_digest), the digest of the
    sources beside it (build.source_digest, None where they are absent) and
    whether the two agree."""
    lib = load().qbx_source_digest().decode()
    src = None
    build_py = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "build.py")
    if os.path.exists(build_py):
        import importlib.util
        spec = importlib.util.spec_from_file_location("qbx_build_script", build_py)
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        src = mod.source_digest()
    return dict(library=lib, sources=src, match=(src is not None and lib == src))


def last_error():
    L = load()
    buf = ctypes.create_string_buffer(2048)
    L.qbx_last_error(buf, 2048)
    return buf.value.decode(errors="replace")


def check(code):
    if code != 0:
        msg = last_error()
        if code == -1:
            raise ValueError(msg)
        raise QbxError(code, msg)


def dp(a):
    return a.ctypes.data_as(_pd)


def ip(a):
    return a.ctypes.data_as(_pi)


def f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))
