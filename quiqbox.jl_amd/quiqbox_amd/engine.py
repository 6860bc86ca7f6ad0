"""Device context and basis handles over the C-ABI.

`Context` owns one GPU and one stream (qbx_init); `GPUBasis` owns a
qbx_basis (qbx_basis_create) and exposes the integral and Fock-build entry
points.  Bases whose functions mix centres or Cartesian exponents inside one
contraction are evaluated in an auxiliary one-primitive-per-function basis and
transformed with phi = chi T.
"""
import ctypes
import threading

import numpy as np

from . import _lib
from .basis import FlatBasis, expand_to_standard, flatten_basis

_default_ctx = None
_ctx_lock = threading.Lock()


class Context:
    def __init__(self, device=0):
        L = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(L.qbx_init(int(device), ctypes.byref(h)))
        self.h = h
        self.device = device
        self.lock = threading.Lock()

    def close(self):
        if self.h:
            _lib.load().qbx_finalize(self.h)
            self.h = None

    def set_timing(self, mode=1):
        """Per-class J/K kernel timing (qbx_set_timing): 1 serialized builds,
        2 concurrent builds (averaged until the next last_timing), 0 off."""
        _lib.check(_lib.load().qbx_set_timing(self.h, int(mode)))

    def last_timing(self):
        n = _lib.load().qbx_n_classes()
        ms = (ctypes.c_float * n)()
        tot = ctypes.c_float()
        _lib.check(_lib.load().qbx_last_timing(self.h, ms, ctypes.byref(tot)))
        return np.array(ms[:], dtype=np.float64), float(tot.value)

    def boys(self, order, T):
        """F_0..F_order at the arguments T from the ERI kernels' Boys function
        on the device (qbx_boys_device): array (len(T), order + 1)."""
        T = np.ascontiguousarray(np.atleast_1d(T), dtype=np.float64)
        F = np.zeros((len(T), order + 1))
        with self.lock:
            _lib.check(_lib.load().qbx_boys_device(self.h, int(order), len(T), _lib.dp(T), _lib.dp(F)))
        return F

    def prim_one_e(self, kind, A, a, la, B, b, lb, C=None, degree=None):
        """One-electron integrals of primitive pairs on the device
        (qbx_prim_one_e): kind QBX_PRIM_OVERLAP / KINETIC / COULOMB (one-body
        repulsion at C, no charge) / MULTIPOLE (about C, degree per axis);
        A, B (n, 3) centres, a, b (n,) exponents, la, lb (n, 3) Cartesian
        exponents.  Returns an (n,) array."""
        A = np.ascontiguousarray(np.atleast_2d(A), dtype=np.float64)
        B = np.ascontiguousarray(np.atleast_2d(B), dtype=np.float64)
        a = np.ascontiguousarray(np.atleast_1d(a), dtype=np.float64)
        b = np.ascontiguousarray(np.atleast_1d(b), dtype=np.float64)
        la = np.ascontiguousarray(np.atleast_2d(la), dtype=np.int32)
        lb = np.ascontiguousarray(np.atleast_2d(lb), dtype=np.int32)
        n = len(a)
        if not (A.shape == B.shape == la.shape == lb.shape == (n, 3) and len(b) == n):
            raise ValueError("DimensionMismatch: primitive-pair arrays of different lengths")
        Cv = None if C is None else np.ascontiguousarray(C, dtype=np.float64)
        Cp = None if Cv is None else _lib.dp(Cv)
        dv = None if degree is None else np.ascontiguousarray(degree, dtype=np.int32)
        out = np.zeros(n)
        with self.lock:
            _lib.check(_lib.load().qbx_prim_one_e(self.h, int(kind), n, _lib.dp(A), _lib.dp(a), _lib.ip(la),
                                                  _lib.dp(B), _lib.dp(b), _lib.ip(lb), Cp,
                                                  None if dv is None else _lib.ip(dv), _lib.dp(out)))
        return out

    def prim_eri(self, quartets):
        """(ab|cd) of primitive quartets with unit coefficients on the device
        (qbx_prim_eri): quartets is a list of four (centre, exponent,
        (i, j, k)) primitives each.  Returns an (n,) array."""
        n = len(quartets)
        cen = np.ascontiguousarray([[p[0] for p in q] for q in quartets], dtype=np.float64).reshape(n, 4, 3)
        xpn = np.ascontiguousarray([[p[1] for p in q] for q in quartets], dtype=np.float64).reshape(n, 4)
        ang = np.ascontiguousarray([[p[2] for p in q] for q in quartets], dtype=np.int32).reshape(n, 4, 3)
        out = np.zeros(n)
        with self.lock:
            _lib.check(_lib.load().qbx_prim_eri(self.h, n, _lib.dp(cen), _lib.dp(xpn), _lib.ip(ang), _lib.dp(out)))
        return out

    def set_mixed_cut(self, cut):
        _lib.check(_lib.load().qbx_set_mixed_cut(self.h, float(cut)))

    def set_density_screening(self, thr):
        _lib.check(_lib.load().qbx_set_density_screening(self.h, float(thr)))

    def purified_density(self, X, F, nocc, max_iter=100):
        """D = C[:, :nocc] C[:, :nocc]^T of eigen(X' F X) (C = X C') by
        trace-correcting purification on the GPU (qbx_density_purify).
        Returns (D, products taken, final |tr P - tr P^2|); raises QbxError
        when it does not settle (no gap at the Fermi level)."""
        X = np.ascontiguousarray(X, dtype=np.float64)
        F = np.ascontiguousarray(F, dtype=np.float64)
        n = X.shape[0]
        if X.shape != (n, n) or F.shape != (n, n):
            raise ValueError("X and F must be square matrices of one size")
        D = np.empty((n, n), dtype=np.float64)
        it = ctypes.c_int32()
        idem = ctypes.c_double()
        pd = ctypes.POINTER(ctypes.c_double)
        with self.lock:
            _lib.check(_lib.load().qbx_density_purify(self.h, n, int(nocc), X.ctypes.data_as(pd),
                                                      F.ctypes.data_as(pd), D.ctypes.data_as(pd), int(max_iter),
                                                      ctypes.byref(it), ctypes.byref(idem)))
        return D, int(it.value), float(idem.value)


    def fock_solve(self, X, F, nocc):
        """One SCF step's orbitals on the GPU (qbx_fock_solve): eigen(X' F X)
        ascending, C = X C' with C[0, j] >= 0, and D = C[:, :nocc] C[:, :nocc]^T.
        Returns (C, orbital energies, D)."""
        X = np.ascontiguousarray(X, dtype=np.float64)
        F = np.ascontiguousarray(F, dtype=np.float64)
        n = X.shape[0]
        if X.shape != (n, n) or F.shape != (n, n):
            raise ValueError("X and F must be square matrices of one size")
        C = np.empty((n, n))
        e = np.empty(n)
        D = np.empty((n, n))
        pd = ctypes.POINTER(ctypes.c_double)
        with self.lock:
            _lib.check(_lib.load().qbx_fock_solve(self.h, n, int(nocc), X.ctypes.data_as(pd), F.ctypes.data_as(pd),
                                                  C.ctypes.data_as(pd), e.ctypes.data_as(pd), D.ctypes.data_as(pd)))
        return C, e, D

    def mo_transform(self, ints, C):
        """changeOrbitalBasis (Interface.jl:376-380) of a one-body (2-index)
        or two-body (4-index) AO tensor with the n_ao x n_mo matrix C: one
        FP64 matrix-core GEMM per index on the device (qbx_mo_transform)."""
        A = np.ascontiguousarray(ints, dtype=np.float64)
        c = np.ascontiguousarray(C, dtype=np.float64)
        if c.ndim != 2:
            raise ValueError("C must be a matrix")
        nao, nmo = c.shape
        if A.ndim not in (2, 4) or any(d != nao for d in A.shape):
            raise ValueError("integrals must be n_ao^2 or n_ao^4 with n_ao = C.shape[0]")
        out = np.empty((nmo,) * A.ndim, dtype=np.float64)
        pd = ctypes.POINTER(ctypes.c_double)
        with self.lock:
            _lib.check(_lib.load().qbx_mo_transform(self.h, A.ndim, nao, nmo, A.ctypes.data_as(pd),
                                                    c.ctypes.data_as(pd), out.ctypes.data_as(pd)))
        return out

    def mo_coulomb_ab(self, eri, C1, C2):
        """getJab (Interface.jl:382-393): Coulomb matrix between the orbitals
        of C1 and those of C2 (qbx_mo_coulomb_ab)."""
        E = np.ascontiguousarray(eri, dtype=np.float64)
        c1 = np.ascontiguousarray(C1, dtype=np.float64)
        c2 = np.ascontiguousarray(C2, dtype=np.float64)
        nao = E.shape[0]
        if E.shape != (nao,) * 4 or c1.ndim != 2 or c2.ndim != 2 or c1.shape[0] != nao or c2.shape[0] != nao:
            raise ValueError("eri must be n_ao^4 and C1, C2 have n_ao rows")
        J = np.empty((c1.shape[1], c2.shape[1]), dtype=np.float64)
        pd = ctypes.POINTER(ctypes.c_double)
        with self.lock:
            _lib.check(_lib.load().qbx_mo_coulomb_ab(self.h, nao, c1.shape[1], c2.shape[1], E.ctypes.data_as(pd),
                                                     c1.ctypes.data_as(pd), c2.ctypes.data_as(pd),
                                                     J.ctypes.data_as(pd)))
        return J


def default_context():
    global _default_ctx
    with _ctx_lock:
        if _default_ctx is None:
            _default_ctx = Context(0)
        return _default_ctx


def class_table():
    L = _lib.load()
    n = L.qbx_n_classes()
    t = np.zeros(4 * n, dtype=np.int32)
    _lib.check(L.qbx_class_table(_lib.ip(t)))
    return t.reshape(n, 4)


class GPUBasis:
    """A basis set resident on the GPU."""

    def __init__(self, orbs_or_flat, ctx=None):
        self.ctx = ctx or default_context()
        flat = orbs_or_flat if isinstance(orbs_or_flat, FlatBasis) else flatten_basis(orbs_or_flat)
        self.flat = flat
        self.n = flat.n_bf
        if flat.is_standard():
            self.T = None
            dev = flat
        else:
            dev, self.T = expand_to_standard(flat)
        self.n_dev = dev.n_bf
        L = _lib.load()
        self._keep = [_lib.f64(dev.prim_cen).reshape(-1), _lib.f64(dev.prim_xpn), _lib.i32(dev.prim_ang).reshape(-1),
                      _lib.i32(dev.bf_ptr), _lib.i32(dev.bf_prim), _lib.f64(dev.bf_w)]
        c, x, a, p, q, w = self._keep
        h = ctypes.c_void_p()
        _lib.check(L.qbx_basis_create(self.ctx.h, dev.n_prim, _lib.dp(c), _lib.dp(x), _lib.ip(a), dev.n_bf,
                                      _lib.ip(p), _lib.ip(q), _lib.dp(w), ctypes.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib.load().qbx_basis_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def info(self):
        v = [ctypes.c_int32() for _ in range(4)]
        _lib.check(_lib.load().qbx_basis_info(self.h, *[ctypes.byref(x) for x in v]))
        return dict(n_bf=v[0].value, n_shell=v[1].value, n_pair=v[2].value, max_l=v[3].value)

    # ---- helpers for the auxiliary-basis transform ----
    def _mat_out(self, fn, *args):
        M = np.zeros((self.n_dev, self.n_dev))
        with self.ctx.lock:
            _lib.check(fn(self.ctx.h, self.h, *args, _lib.dp(M)))
        if self.T is not None:
            M = self.T.T @ M @ self.T
        return M

    def _cached(self, name, key, make):
        """A copy of the one-electron matrix `name` for `key`, computed once:
        the basis is fixed, so S, T and the core Hamiltonian of the last
        nuclei asked for are kept (one entry per matrix)."""
        c = self.__dict__.setdefault("_one_e", {})
        hit = c.get(name)
        if hit is None or hit[0] != key:
            hit = (key, make())
            c[name] = hit
        return hit[1].copy()

    def overlap(self):
        return self._cached("S", None, lambda: self._mat_out(_lib.load().qbx_overlap))

    def orthonormalization(self):
        """X = S^{-1/2} of this basis (HartreeFock.jl:39-42), computed once per
        basis (read-only array; every SCF and guess on the basis shares it)."""
        if getattr(self, "_X", None) is None:
            from .hartree_fock import getOrthonormalization
            X = getOrthonormalization(self.overlap())
            X.flags.writeable = False
            self._X = X
        return self._X

    def kinetic(self):
        return self._cached("T", None, lambda: self._mat_out(_lib.load().qbx_kinetic))

    def nuclear(self, Z, R):
        Z = _lib.f64(Z)
        R = _lib.f64(R).reshape(-1)
        return self._mat_out(_lib.load().qbx_nuclear, len(Z), _lib.dp(Z), _lib.dp(R))

    def nuclear_own(self, Z, R):
        """V[i][j] = attraction of the nucleus at function i's centre alone
        (qbx_nuclear_own; zero rows where no nucleus sits).  Standard bases only:
        the per-function attribution does not survive a basis transform."""
        if self.T is not None:
            raise ValueError("own-nucleus attraction needs a standard (single-centre) basis")
        Z = _lib.f64(Z)
        R = _lib.f64(R).reshape(-1)
        return self._mat_out(_lib.load().qbx_nuclear_own, len(Z), _lib.dp(Z), _lib.dp(R))

    def core_hamiltonian(self, Z, R):
        Z = _lib.f64(Z)
        R = _lib.f64(R).reshape(-1)
        return self._cached("H", (Z.tobytes(), R.tobytes()),
                            lambda: self._mat_out(_lib.load().qbx_core_hamiltonian, len(Z), _lib.dp(Z), _lib.dp(R)))

    def multipole(self, center, degree):
        c = _lib.f64(center)
        d = _lib.i32(degree)
        return self._mat_out(_lib.load().qbx_multipole, _lib.dp(c), _lib.ip(d))

    def eri_dense(self):
        n = self.n_dev
        E = np.zeros((n, n, n, n))
        with self.ctx.lock:
            _lib.check(_lib.load().qbx_eri_dense(self.ctx.h, self.h, _lib.dp(E)))
        if self.T is not None:
            T = self.T
            E = np.einsum("abcd,ai,bj,ck,dl->ijkl", E, T, T, T, T, optimize=True)
        return E

    def jk(self, DJ, DK, thr=1e-12, precision=_lib.QBX_PREC_FP64):
        """G[s] = J[DJ[s]] - K[DK[s]] for stacked (n_set, n, n) densities."""
        DJ = np.asarray(DJ, dtype=np.float64)
        DK = np.asarray(DK, dtype=np.float64)
        single = DJ.ndim == 2
        if single:
            DJ = DJ[None]
            DK = DK[None]
        if DJ.shape != DK.shape or DJ.shape[1:] != (self.n, self.n):
            raise ValueError("density shape %s / %s does not match the basis size %d" % (DJ.shape, DK.shape, self.n))
        if self.T is not None:
            T = self.T
            DJ = np.einsum("ai,sij,bj->sab", T, DJ, T)
            DK = np.einsum("ai,sij,bj->sab", T, DK, T)
        DJ = np.ascontiguousarray(DJ)
        DK = np.ascontiguousarray(DK)
        G = np.zeros_like(DJ)
        with self.ctx.lock:
            _lib.check(_lib.load().qbx_jk(self.ctx.h, self.h, DJ.shape[0], _lib.dp(DJ), _lib.dp(DK), _lib.dp(G),
                                          float(thr), int(precision)))
        if self.T is not None:
            G = np.einsum("ai,sab,bj->sij", self.T, G, self.T)
        return G[0] if single else G

    def two_body(self, threshold=1e-12, precision=_lib.QBX_PREC_FP64, **kw):
        """The Fock-build operator of this basis (direct, Schwarz-screened)."""
        from .hartree_fock import DirectJK
        return DirectJK(self, threshold, precision, **kw)

    def jk_device(self, dDJ_ptr, dDK_ptr, dG_ptr, n_set, thr=1e-12, precision=_lib.QBX_PREC_FP64, stream=None):
        if self.T is not None:
            raise ValueError("device-pointer Fock builds need a standard (single-centre) basis")
        with self.ctx.lock:
            _lib.check(_lib.load().qbx_jk_device(self.ctx.h, self.h, int(n_set), ctypes.c_void_p(dDJ_ptr),
                                                 ctypes.c_void_p(dDK_ptr), ctypes.c_void_p(dG_ptr), float(thr),
                                                 int(precision), ctypes.c_void_p(stream or 0)))

    def shells(self):
        """The kernels' shell table (qbx_basis_shells): per shell L, radial
        primitive offset and count, component functions and scales, centre;
        per radial primitive exponent and coefficient."""
        L = _lib.load()
        ns, nr = ctypes.c_int32(), ctypes.c_int32()
        _lib.check(L.qbx_basis_shells(self.h, ctypes.byref(ns), ctypes.byref(nr), None, None, None, None, None, None,
                                      None, None))
        ns, nr = ns.value, nr.value
        Ls, npr, poff = (np.zeros(ns, dtype=np.int32) for _ in range(3))
        mc = _lib.QBX_SHELL_MAXC
        bf = np.zeros(mc * ns, dtype=np.int32)
        sc = np.zeros(mc * ns)
        cen = np.zeros(3 * ns)
        xpn, coef = np.zeros(nr), np.zeros(nr)
        _lib.check(L.qbx_basis_shells(self.h, None, None, _lib.ip(Ls), _lib.ip(npr), _lib.ip(poff), _lib.ip(bf),
                                      _lib.dp(sc), _lib.dp(cen), _lib.dp(xpn), _lib.dp(coef)))
        return dict(L=Ls, nprim=npr, poff=poff, bf=bf.reshape(ns, mc), sc=sc.reshape(ns, mc), center=cen.reshape(ns, 3),
                    xpn=xpn, coef=coef)

    def rhf_gradient(self, Dt, W, Z, R, thr=1e-12):
        """qbx_rhf_gradient: per radial primitive dE/d(exponent) and c dE/dc,
        per shell dE/d(centre), every function's weights held fixed."""
        if self.T is not None:
            raise ValueError("the gradient needs a standard (single-centre) basis")
        Dt, W = _lib.f64(Dt), _lib.f64(W)
        Z, R = _lib.f64(Z), _lib.f64(R).reshape(-1)
        sh = self.shells()
        nr, ns = len(sh["xpn"]), len(sh["L"])
        gx, gl, gc = np.zeros(nr), np.zeros(nr), np.zeros(3 * ns)
        with self.ctx.lock:
            _lib.check(_lib.load().qbx_rhf_gradient(self.ctx.h, self.h, _lib.dp(Dt), _lib.dp(W), len(Z), _lib.dp(Z),
                                                    _lib.dp(R), float(thr), _lib.dp(gx), _lib.dp(gl), _lib.dp(gc)))
        return gx, gl, gc.reshape(ns, 3), sh

    def uhf_gradient(self, Dt, Da, Db, W, Z, R, thr=1e-12):
        """qbx_uhf_gradient: as rhf_gradient for a UHF state (Dt = Da + Db,
        the spin densities Da, Db, W summed over both spins)."""
        if self.T is not None:
            raise ValueError("the gradient needs a standard (single-centre) basis")
        Dt, Da, Db, W = _lib.f64(Dt), _lib.f64(Da), _lib.f64(Db), _lib.f64(W)
        Z, R = _lib.f64(Z), _lib.f64(R).reshape(-1)
        sh = self.shells()
        nr, ns = len(sh["xpn"]), len(sh["L"])
        gx, gl, gc = np.zeros(nr), np.zeros(nr), np.zeros(3 * ns)
        with self.ctx.lock:
            _lib.check(_lib.load().qbx_uhf_gradient(self.ctx.h, self.h, _lib.dp(Dt), _lib.dp(Da), _lib.dp(Db),
                                                    _lib.dp(W), len(Z), _lib.dp(Z), _lib.dp(R), float(thr),
                                                    _lib.dp(gx), _lib.dp(gl), _lib.dp(gc)))
        return gx, gl, gc.reshape(ns, 3), sh

    def jk_stats(self, thr=1e-12, cut=0.0):
        n = _lib.load().qbx_n_classes()
        arrs = [np.zeros(n, dtype=np.int64) for _ in range(4)]
        _lib.check(_lib.load().qbx_jk_stats(self.ctx.h, self.h, float(thr), float(cut),
                                            *[a.ctypes.data_as(_lib._pl) for a in arrs]))
        return dict(shell_quartets=arrs[0], cart_quartets=arrs[1], prim_quartets=arrs[2], fp64_shell_quartets=arrs[3])
