"""ctypes binding of the CPU checkers -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg; never by the
flint_amd package.  Two libraries are bound:

* libflint_oracle.so  the C restatement of the reference path (flint_oracle.c, oracle_bdf.c,
                      oracle_cea.c), built by ``make -C oracle``;
* _ref/libflint_ref.so the reference Fortran modules themselves, compiled by ``make -C oracle ref``
                      where the reference checkout exists (never on the GPU host).
"""
import ctypes as C
import os
import subprocess

import numpy as np

from flint_amd.mechanism import flatten

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libflint_oracle.so")
BENCH_LIB = os.path.join(HERE, "libflint_oracle_bench.so")  # release flags, bench.py's CPU baseline
REF_LIB = os.path.join(HERE, "_ref", "libflint_ref.so")

_P = C.c_void_p
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class _OrcMech(C.Structure):
    _fields_ = [("ns", C.c_int), ("nr", C.c_int), ("n_arrh", C.c_int), ("n_troe", C.c_int),
                ("n_lind", C.c_int), ("nms", C.c_int)] + \
               [(f, _P) for f in ("rx_kind", "rx_table", "rx_mset", "rx_fwd_ptr", "rx_fwd_sp", "rx_fwd_pow",
                                  "rx_rev_ptr", "rx_rev_sp", "rx_rev_pow", "ms_ptr", "ms_sp", "ms_is_sum",
                                  "ms_eff", "wd_ptr", "wd_rx", "wd_coef", "W")] + \
               [("Tk_min", C.c_int), ("Tk_max", C.c_int), ("kf", _P), ("kb", _P), ("troe", _P), ("lind", _P),
                ("Th_min", C.c_int), ("Th_max", C.c_int), ("cp", _P), ("h", _P), ("s", _P), ("is_wd", C.c_int)]


class _OrcCea(C.Structure):
    _fields_ = [("ne", C.c_int), ("A", _P), ("elw", _P), ("tol", C.c_double)]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None
_bench_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
    return _lib


def bench_lib():
    global _bench_lib
    if _bench_lib is None:
        if not os.path.exists(BENCH_LIB):
            build()
        _bench_lib = C.CDLL(BENCH_LIB)
    return _bench_lib


def _ptr(a):
    return a.ctypes.data_as(_P)


class Oracle:
    """The C restatement bound to one mechanism's program and tables (a MechData)."""

    def __init__(self, md, release=False):
        """release: the baseline build at release flags (libflint_oracle_bench.so) instead of the
        checker build."""
        self.md = md
        self.ns = md.ns
        self._keep = flatten(md.prog)
        self._keep.update(kf=md.kf, kb=md.kb, troe=md.troe, lind=md.lind, cp=md.cp, h=md.h, s=md.s)
        m = _OrcMech()
        m.ns, m.nr = md.ns, len(md.prog.reactions)
        m.n_arrh, m.n_troe, m.n_lind, m.nms = md.n_arrh, md.n_troe, md.n_lind, len(md.prog.m_sets)
        for k, v in self._keep.items():
            if hasattr(m, k) and isinstance(v, np.ndarray):
                setattr(m, k, _ptr(v))
        m.Tk_min, m.Tk_max, m.Th_min, m.Th_max = md.Tk_min, md.Tk_max, md.Th_min, md.Th_max
        m.is_wd = int(md.is_wd)
        self.m = m
        self.L = bench_lib() if release else lib()

    def wdot(self, rho, T):
        """rho [ns, n] SoA, T [n] -> (wdot [ns, n], scale [ns, n])."""
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        T = np.ascontiguousarray(T, dtype=np.float64)
        n = T.size
        w = np.empty_like(rho)
        sc = np.empty_like(rho)
        self.L.orc_wdot(C.byref(self.m), C.c_long(n), C.c_long(n), _ptr(rho), _ptr(T), _ptr(w), _ptr(sc))
        return w, sc

    def wd_hrr(self, rho, T):
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        T = np.ascontiguousarray(T, dtype=np.float64)
        n = T.size
        w = np.empty_like(rho)
        q = np.empty(n)
        self.L.orc_wd_hrr(C.byref(self.m), C.c_long(n), C.c_long(n), _ptr(rho), _ptr(T), _ptr(w), _ptr(q))
        return w, q

    def general(self, rho, T):
        """The data-driven general routine with the mechanism's tables: rho [ns, n], T [n] ->
        (wdot [ns, n], scale [ns, n])."""
        na, nt, nl, ni1, ni2, eps = self.md.general
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        T = np.ascontiguousarray(T, dtype=np.float64)
        n = T.size
        w, sc = np.empty_like(rho), np.empty_like(rho)
        self.L.orc_general(C.byref(self.m), C.c_int(na), C.c_int(nt), C.c_int(nl), _ptr(ni1), _ptr(ni2), _ptr(eps),
                           C.c_long(n), C.c_long(n), _ptr(rho), _ptr(T), _ptr(w), _ptr(sc))
        return w, sc

    def cons2prim(self, cons, T):
        """cons [ns+nd+1, n] -> (prim [ns+nd+1, n], T_out [n], iters [n], status [n]) (cons2prim)."""
        cons = np.ascontiguousarray(cons, dtype=np.float64)
        T = np.ascontiguousarray(T, dtype=np.float64)
        npr, n = cons.shape
        prim, Tout = np.empty_like(cons), np.empty(n)
        it, st = np.empty(n, dtype=np.int32), np.empty(n, dtype=np.int32)
        self.L.orc_cons2prim(C.byref(self.m), C.c_long(n), C.c_long(n), C.c_int(npr), _ptr(cons), _ptr(T), _ptr(prim),
                             _ptr(Tout), _ptr(it), _ptr(st))
        return prim, Tout, it, st

    def wilke(self, tables, rhoi, rho, T):
        """co_k_mi_lam_Wilke with species tables (Tt_min, Tt_max, mu_tab [ns, nT], k_tab [ns, nT]):
        rhoi [ns, n], rho [n], T [n] -> (mu [n], lam [n], status [n])."""
        Tt_min, Tt_max, mu_tab, k_tab = tables
        rhoi = np.ascontiguousarray(rhoi, dtype=np.float64)
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        T = np.ascontiguousarray(T, dtype=np.float64)
        n = T.size
        mu, lam, st = np.empty(n), np.empty(n), np.empty(n, dtype=np.int32)
        self.L.orc_wilke(C.c_int(self.md.ns), _ptr(np.ascontiguousarray(self.md.W, dtype=np.float64)),
                         C.c_int(Tt_min), C.c_int(Tt_max), _ptr(mu_tab), _ptr(k_tab), C.c_long(n), C.c_long(n),
                         _ptr(rhoi), _ptr(rho), _ptr(T), _ptr(mu), _ptr(lam), _ptr(st))
        return mu, lam, st

    def laminar_viscosity(self, tables, rhoi, p, rho, Rtot):
        """f_laminarViscosity: the Wilke viscosity at T = p / (Rtot rho)."""
        Tt_min, Tt_max, mu_tab, _ = tables
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rhoi, p, rho, Rtot)]
        n = arrs[1].size
        mu = np.empty(n)
        self.L.orc_laminar_viscosity(C.c_int(self.md.ns), _ptr(np.ascontiguousarray(self.md.W, dtype=np.float64)),
                                     C.c_int(Tt_min), C.c_int(Tt_max), _ptr(mu_tab), C.c_long(n), C.c_long(n),
                                     *[_ptr(a) for a in arrs], _ptr(mu))
        return mu

    def prim2cons(self, prim):
        """prim [ns+nd+1, n] -> (cons, status) (prim2cons)."""
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        npr, n = prim.shape
        cons = np.empty_like(prim)
        st = np.empty(n, dtype=np.int32)
        self.L.orc_prim2cons(C.byref(self.m), C.c_long(n), C.c_long(n), C.c_int(npr), _ptr(prim), _ptr(cons), _ptr(st))
        return cons, st

    def rhs(self, Z):
        Z = np.ascontiguousarray(Z, dtype=np.float64)
        n = Z.shape[1]
        F = np.empty_like(Z)
        self.L.orc_rhs(C.byref(self.m), C.c_long(n), C.c_long(n), _ptr(Z), _ptr(F))
        return F

    def jac(self, Z, fscale=False):
        """Z [nz, n] -> (F [nz, n], J [nz*nz, n] row-major per cell, |J| scale) and, with fscale,
        the scale of F (sum of |terms| of each component) as a fourth result."""
        Z = np.ascontiguousarray(Z, dtype=np.float64)
        nz, n = Z.shape
        F = np.empty_like(Z)
        J = np.empty((nz * nz, n))
        Ja = np.empty((nz * nz, n))
        Fa = np.empty_like(Z) if fscale else None
        self.L.orc_jac(C.byref(self.m), C.c_long(n), C.c_long(n), _ptr(Z), _ptr(F), _ptr(J), _ptr(Ja),
                       None if Fa is None else _ptr(Fa))
        return (F, J, Ja, Fa) if fscale else (F, J, Ja)

    def jac_cstep(self, Z, h=1e-30):
        Z = np.ascontiguousarray(Z, dtype=np.float64)
        nz, n = Z.shape
        J = np.empty((nz * nz, n))
        self.L.orc_jac_cstep(C.byref(self.m), C.c_long(n), C.c_long(n), _ptr(Z), C.c_double(h), _ptr(J))
        return J

    def rhs_cp(self, X, p):
        X = np.ascontiguousarray(X, dtype=np.float64)
        p = np.ascontiguousarray(p, dtype=np.float64)
        n = X.shape[1]
        F = np.empty_like(X)
        self.L.orc_rhs_cp(C.byref(self.m), C.c_long(n), C.c_long(n), _ptr(X), _ptr(p), _ptr(F))
        return F

    def jac_cp(self, X, p, scales=False):
        """Constant-pressure right-hand side and Jacobian: X [nz, n] = (Y_k, T), p [n] ->
        (F, J) and, with scales, (F, J, |F| scale, |J| scale)."""
        X = np.ascontiguousarray(X, dtype=np.float64)
        p = np.ascontiguousarray(p, dtype=np.float64)
        nz, n = X.shape
        F = np.empty_like(X)
        J = np.empty((nz * nz, n))
        Fa = np.empty_like(X) if scales else None
        Ja = np.empty((nz * nz, n)) if scales else None
        self.L.orc_jac_cp(C.byref(self.m), C.c_long(n), C.c_long(n), _ptr(X), _ptr(p), _ptr(F), _ptr(J),
                          None if Fa is None else _ptr(Fa), None if Ja is None else _ptr(Ja))
        return (F, J, Fa, Ja) if scales else (F, J)

    def jac_cp_cstep(self, X, p, h=1e-30):
        X = np.ascontiguousarray(X, dtype=np.float64)
        p = np.ascontiguousarray(p, dtype=np.float64)
        nz, n = X.shape
        J = np.empty((nz * nz, n))
        self.L.orc_jac_cp_cstep(C.byref(self.m), C.c_long(n), C.c_long(n), _ptr(X), _ptr(p), C.c_double(h), _ptr(J))
        return J

    def cea(self, p, T, rho, hp=True, guess=None, tol=0.0):
        """Equilibrium of n states (orc_cea): p [n] (bar for HP), T [n] (the input mixture's
        temperature, also the first iterate), rho [ns, n] amounts; guess [ns+1, n] = (T, ln n_k) or
        None; tol = 0 the reference's convergence tests, > 0 the tightened mode.  Returns
        (T_eq [n], Y_eq [ns, n], iterations [n], status [n])."""
        from flint_amd.elements import element_weights
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        ns, n = rho.shape
        if not hasattr(self, "_cea"):
            A = np.ascontiguousarray(self.md.composition, dtype=np.float64)
            elw = element_weights(self.md.elements)
            c = _OrcCea()
            c.ne, c.A, c.elw = A.shape[0], _ptr(A), _ptr(elw)
            self._cea = (c, A, elw)
        p = np.ascontiguousarray(p, dtype=np.float64)
        T = np.ascontiguousarray(T, dtype=np.float64)
        g = None if guess is None else np.ascontiguousarray(guess, dtype=np.float64)
        self._cea[0].tol = float(tol)
        To = np.empty(n)
        Yo = np.empty((ns, n))
        it = np.empty(n, dtype=np.int32)
        st = np.empty(n, dtype=np.int32)
        self.L.orc_cea(C.byref(self.m), C.byref(self._cea[0]), C.c_long(n), C.c_long(n), C.c_int(int(hp)), _ptr(p),
                       _ptr(T), _ptr(rho), None if g is None else _ptr(g), _ptr(To), _ptr(Yo), _ptr(it), _ptr(st))
        return To, Yo, it, st

    def bdf(self, X, p=None, mode=0, rtol=1e-6, atol=1e-12, t_end=1e-4, max_steps=20000, newton=0):
        """Advance reactors (X [nz, n], mode 0 = (rho, T) constant volume, 1 = (Y, T) constant p);
        newton = 0 scipy's Newton policy, 1 CVODE's (oracle_bdf.c)."""
        X = np.array(X, dtype=np.float64, order="C", copy=True)
        nz, n = X.shape
        pp = np.ascontiguousarray(p if p is not None else np.zeros(n), dtype=np.float64)
        stats = np.zeros((n, 8), dtype=np.int32)
        status = np.zeros(n, dtype=np.int32)
        self.L.orc_bdf_method(C.byref(self.m), C.c_long(n), C.c_long(n), C.c_int(mode), _ptr(X), _ptr(pp),
                              C.c_double(rtol), C.c_double(atol), C.c_double(t_end), C.c_int(max_steps), _ptr(stats),
                              _ptr(status), C.c_int(newton))
        return X, stats, status

    def num_threads(self):
        return self.L.orc_num_threads()

    def set_threads(self, n):
        self.L.orc_set_threads(C.c_int(n))


class Reference:
    """The reference Fortran modules (oracle/_ref), loaded with one mechanism's tables."""

    def __init__(self, md, lib_path=REF_LIB):
        if not os.path.exists(lib_path):
            raise FileNotFoundError(lib_path)
        self.L = C.CDLL(lib_path)
        self.md = md
        ns = md.ns
        nTk = md.kf.shape[0]
        nTh = md.cp.shape[0]
        names = "".join(s.ljust(20) for s in md.species).encode()
        els = "".join(e.ljust(20) for e in md.elements).encode() or b" " * 20
        comp = np.ascontiguousarray(md.composition.T)  # Fortran (ne, ns) column-major
        troe = md.troe if md.n_troe else np.zeros((nTk, 1, 4))
        lind = md.lind if md.n_lind else np.zeros((nTk, 1, 3))
        self._keep = [md.W, md.kf, md.kb, troe, lind, md.cp, md.h, md.s, comp]
        mech = md.desc["name"].encode()
        self.L.ref_setup(C.c_int(ns), _ptr(md.W), C.c_int(nTk), C.c_int(md.Tk_min), C.c_int(md.n_arrh), _ptr(md.kf),
                         _ptr(md.kb), C.c_int(md.n_troe), _ptr(np.ascontiguousarray(troe)), C.c_int(md.n_lind),
                         _ptr(np.ascontiguousarray(lind)), C.c_int(nTh), C.c_int(md.Th_min), _ptr(md.cp),
                         _ptr(md.h), _ptr(md.s), C.c_int(len(md.elements)), _ptr(comp), C.c_char_p(names),
                         C.c_char_p(els), C.c_char_p(mech), C.c_int(len(mech)))

    def wdot(self, rho, T):
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        T = np.ascontiguousarray(T, dtype=np.float64)
        n = T.size
        w = np.empty_like(rho)
        self.L.ref_wdot(C.c_long(n), C.c_long(n), _ptr(rho), _ptr(T), _ptr(w))
        return w

    def general(self, rho, T):
        na, nt, nl, ni1, ni2, eps = self.md.general
        self._gkeep = (ni1, ni2, eps)
        self.L.ref_general_setup(C.c_int(na), C.c_int(nt), C.c_int(nl), _ptr(ni1), _ptr(ni2), _ptr(eps))
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        T = np.ascontiguousarray(T, dtype=np.float64)
        n = T.size
        w = np.empty_like(rho)
        self.L.ref_general(C.c_long(n), C.c_long(n), _ptr(rho), _ptr(T), _ptr(w))
        return w

    def cons2prim(self, cons, T):
        cons = np.ascontiguousarray(cons, dtype=np.float64)
        npr, n = cons.shape
        prim = np.empty_like(cons)
        self.L.ref_cons2prim(C.c_long(n), C.c_long(n), C.c_int(npr), _ptr(cons),
                             _ptr(np.ascontiguousarray(T, dtype=np.float64)), _ptr(prim))
        return prim

    def wilke(self, tables, rhoi, rho, T, p, Rtot):
        """co_k_mi_lam_Wilke and f_laminarViscosity of the reference modules per cell, with the
        species tables on the thermo rows -> (mu, lam, mu_laminar)."""
        Tt_min, Tt_max, mu_tab, k_tab = tables
        assert Tt_min == self.md.Th_min and mu_tab.shape[1] == self.md.cp.shape[0]
        self._tkeep = (mu_tab, k_tab)
        self.L.ref_transport_setup(C.c_int(mu_tab.shape[1]), _ptr(mu_tab), _ptr(k_tab))
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rhoi, rho, T, p, Rtot)]
        n = arrs[2].size
        mu, lam, mul = np.empty(n), np.empty(n), np.empty(n)
        self.L.ref_wilke(C.c_long(n), C.c_long(n), *[_ptr(a) for a in arrs], _ptr(mu), _ptr(lam), _ptr(mul))
        return mu, lam, mul

    def prim2cons(self, prim):
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        npr, n = prim.shape
        cons = np.empty_like(prim)
        self.L.ref_prim2cons(C.c_long(n), C.c_long(n), C.c_int(npr), _ptr(prim), _ptr(cons))
        return cons

    def cea(self, p, T, rho, hp=False):
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        n = T.size
        Tout = np.empty(n)
        rout = np.empty_like(rho)
        self.L.ref_cea(C.c_long(n), C.c_long(n), _ptr(np.ascontiguousarray(p, dtype=np.float64)),
                       _ptr(np.ascontiguousarray(T, dtype=np.float64)), _ptr(rho), C.c_int(int(hp)), _ptr(Tout),
                       _ptr(rout))
        return Tout, rout
