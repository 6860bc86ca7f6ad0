"""ctypes binding of the CPU oracle (liboracle.so built from qbx_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product package never
imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_d = ctypes.c_double
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int32)


class _Basis(ctypes.Structure):
    _fields_ = [("n_prim", _i32), ("cen", _pd), ("xpn", _pd), ("ang", _pi),
                ("n_bf", _i32), ("bf_ptr", _pi), ("bf_prim", _pi), ("bf_w", _pd)]


def build(force=False):
    """Compile liboracle.so with the committed Makefile (force: even when it is
    up to date)."""
    subprocess.run(["make", "-s"] + (["-B"] if force else []) + ["-C", _HERE], check=True)


def use_native_build(out_dir):
    """Compile the oracle for the host CPU it runs on (-march=native, OpenMP)
    into out_dir and load that copy instead of liboracle.so.  Used by the CPU
    baseline timing only: the committed build stays portable."""
    global _lib, _LIB_PATH
    path = os.path.join(out_dir, "liboracle_native.so")
    subprocess.run(["gcc", "-O3", "-fPIC", "-fopenmp", "-ffp-contract=off", "-march=native", "-std=c11",
                    "-D_GNU_SOURCE", "-shared", "-o", path, os.path.join(_HERE, "qbx_oracle.c"), "-lm"], check=True)
    _LIB_PATH = path
    _lib = None
    return path


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.qor_boys.restype = _d
        L.qor_boys.argtypes = [_d, ctypes.c_int]
        L.qor_boys_seq.argtypes = [_d, ctypes.c_int, _pd]
        L.qor_gamma_p.restype = _d
        L.qor_gamma_p.argtypes = [_d, _d]
        prim = [_pd, _d, _pi]
        L.qor_prim_overlap.restype = _d
        L.qor_prim_overlap.argtypes = prim * 2
        L.qor_prim_kinetic.restype = _d
        L.qor_prim_kinetic.argtypes = prim * 2
        L.qor_prim_coulomb1.restype = _d
        L.qor_prim_coulomb1.argtypes = prim * 2 + [_pd]
        L.qor_prim_eri.restype = _d
        L.qor_prim_eri.argtypes = prim * 4
        L.qor_prim_multipole.restype = _d
        L.qor_prim_multipole.argtypes = prim * 2 + [_pd, _pi]
        bp = ctypes.POINTER(_Basis)
        L.qor_overlap_matrix.argtypes = [bp, _pd]
        L.qor_kinetic_matrix.argtypes = [bp, _pd]
        L.qor_nuclear_matrix.argtypes = [bp, _i32, _pd, _pd, _pd]
        L.qor_multipole_matrix.argtypes = [bp, _pd, _pi, _pd]
        L.qor_eri_quartet.restype = _d
        L.qor_eri_quartet.argtypes = [bp, _i32, _i32, _i32, _i32]
        L.qor_eri_dense.argtypes = [bp, _pd]
        L.qor_eri_list.argtypes = [bp, _i64, _pi, _pd, _i32]
        L.qor_gcore.argtypes = [_i32, _pd, _pd, _pd, _pd]
        L.qor_set_reference_orientation.argtypes = [_i32]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_pd)


def _ip(a):
    return a.ctypes.data_as(_pi)


def _f64(x, n=None):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    return a


def _i32a(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


def set_reference_orientation(on):
    """True: primitive quartets in the reference's canonical orientation
    (bit-level restatement); False (default): numerically stable orientation."""
    lib().qor_set_reference_orientation(1 if on else 0)


def boys(x, n):
    return lib().qor_boys(float(x), int(n))


def boys_seq(x, n):
    out = np.zeros(n + 1)
    lib().qor_boys_seq(float(x), int(n), _dp(out))
    return out


def _prim_args(p):
    cen, xpn, ang = p
    c = _f64(cen)
    a = _i32a(ang)
    return [c, a], [_dp(c), float(xpn), _ip(a)]


def _call_prims(fn, prims, extra=()):
    keep, args = [], []
    for p in prims:
        k, a = _prim_args(p)
        keep += k
        args += a
    extra_keep = [np.ascontiguousarray(e) for e in extra]
    for e in extra_keep:
        args.append(_dp(e) if e.dtype == np.float64 else _ip(e))
    return fn(*args)


def prim_overlap(p1, p2):
    return _call_prims(lib().qor_prim_overlap, (p1, p2))


def prim_kinetic(p1, p2):
    return _call_prims(lib().qor_prim_kinetic, (p1, p2))


def prim_coulomb1(p1, p2, point):
    return _call_prims(lib().qor_prim_coulomb1, (p1, p2), (_f64(point),))


def prim_eri(p1, p2, p3, p4):
    return _call_prims(lib().qor_prim_eri, (p1, p2, p3, p4))


def prim_multipole(p1, p2, center, degree):
    return _call_prims(lib().qor_prim_multipole, (p1, p2), (_f64(center), _i32a(degree)))


class Basis:
    """Flattened basis in the C-ABI layout (primitives + CSR weights)."""

    def __init__(self, flat):
        self.cen = _f64(flat["prim_cen"]).reshape(-1)
        self.xpn = _f64(flat["prim_xpn"])
        self.ang = _i32a(flat["prim_ang"]).reshape(-1)
        self.bf_ptr = _i32a(flat["bf_ptr"])
        self.bf_prim = _i32a(flat["bf_prim"])
        self.bf_w = _f64(flat["bf_w"])
        self.n_bf = len(self.bf_ptr) - 1
        self.s = _Basis(len(self.xpn), _dp(self.cen), _dp(self.xpn), _ip(self.ang),
                        self.n_bf, _ip(self.bf_ptr), _ip(self.bf_prim), _dp(self.bf_w))

    def _ref(self):
        return ctypes.byref(self.s)

    def overlap(self):
        S = np.zeros((self.n_bf, self.n_bf))
        lib().qor_overlap_matrix(self._ref(), _dp(S))
        return S

    def kinetic(self):
        T = np.zeros((self.n_bf, self.n_bf))
        lib().qor_kinetic_matrix(self._ref(), _dp(T))
        return T

    def nuclear(self, Z, R):
        Z = _f64(Z)
        R = _f64(R).reshape(-1)
        V = np.zeros((self.n_bf, self.n_bf))
        lib().qor_nuclear_matrix(self._ref(), len(Z), _dp(Z), _dp(R), _dp(V))
        return V

    def multipole(self, center, degree):
        c = _f64(center)
        d = _i32a(degree)
        M = np.zeros((self.n_bf, self.n_bf))
        lib().qor_multipole_matrix(self._ref(), _dp(c), _ip(d), _dp(M))
        return M

    def eri(self, i, j, k, l):
        return lib().qor_eri_quartet(self._ref(), i, j, k, l)

    def eri_dense(self):
        n = self.n_bf
        E = np.zeros((n, n, n, n))
        lib().qor_eri_dense(self._ref(), _dp(E))
        return E

    def eri_list(self, quartets, threads=1):
        q = _i32a(quartets).reshape(-1, 4)
        out = np.zeros(len(q))
        lib().qor_eri_list(self._ref(), len(q), _ip(q), _dp(out), int(threads))
        return out


def gcore(eri, DJ, DK):
    n = DJ.shape[0]
    e = _f64(eri)
    dj = _f64(DJ)
    dk = _f64(DK)
    G = np.zeros((n, n))
    lib().qor_gcore(n, _dp(e), _dp(dj), _dp(dk), _dp(G))
    return G


def change_orbital_basis(ints, C):
    """changeOrbitalBasis (Interface.jl:376-380): every index of the one- or
    two-body integrals contracted with the coefficient matrix C, index order
    kept (out[i, j(, k, l)])."""
    A = np.asarray(ints, dtype=np.float64)
    c = np.asarray(C, dtype=np.float64)
    for _ in range(A.ndim):
        A = np.tensordot(A, c, axes=([0], [0]))   # contracted index leaves the front, new one goes last
    return A


def coulomb_ab(eri, C1, C2):
    """getJab (Interface.jl:382-393): Jab[m, n] = sum_abcd eri[a, b, c, d]
    C1[a, m] C1[b, m] C2[c, n] C2[d, n]."""
    e = np.asarray(eri, dtype=np.float64)
    n = e.shape[0]
    k1 = np.einsum("am,bm->abm", C1, C1).reshape(n * n, -1)
    k2 = np.einsum("cn,dn->cdn", C2, C2).reshape(n * n, -1)
    return k1.T @ e.reshape(n * n, n * n) @ k2
