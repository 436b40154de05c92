"""ctypes binding of libflint_amd.so (the C ABI declared in include/flint_amd.h).

Device arrays are torch tensors on a ROCm device (torch is used for HBM allocation and streams
only); their data pointers go straight to the C ABI.  The library must have been built with
``python -m flint_amd.build``: nothing here falls back to a CPU path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libflint_amd.so")

_lib = None

FLINT_OK = 0
ERRORS = {1: "invalid argument", 2: "HIP error", 3: "input file error", 4: "unknown mechanism", 5: "unsupported"}


class FlintError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FlintError(f"{LIB_PATH} is missing: build it with `python -m flint_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp, i, l, d = C.c_void_p, C.c_int, C.c_long, C.c_double
        L.flint_last_error.restype = C.c_char_p
        L.flint_mech_create.argtypes = [C.c_char_p, i, vp, i, i, i, vp, vp, i, vp, i, vp, i, i, vp, vp, i, vp,
                                        C.POINTER(vp)]
        L.flint_mech_load.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(vp)]
        L.flint_mech_destroy.argtypes = [vp]
        L.flint_mech_query.argtypes = [vp] + [C.POINTER(i)] * 5
        L.flint_source.argtypes = [vp, l, l, vp, vp, vp, vp, vp, vp]
        L.flint_rhs.argtypes = [vp, l, l, vp, vp, vp, vp, vp]
        L.flint_reactor_advance.argtypes = [vp, l, l, i, vp, vp, d, d, d, i, vp, vp, vp]
        L.flint_reactor_advance_graph.argtypes = [vp, l, l, i, vp, vp, d, d, d, i, vp, vp, vp]
        L.flint_reactor_advance_noalloc.argtypes = [vp, l, l, i, vp, vp, d, d, d, i, vp, vp, vp]
        L.flint_reactor_jacobian.argtypes = [vp, l, l, i, vp, vp, vp, vp, vp]
        L.flint_reactor_check_step.argtypes = [vp, l, l, vp, vp, vp, d, d, d, vp, vp, vp]
        L.flint_reactor_autotune.argtypes = [vp, l, l, i, vp, vp, d, d, d, C.POINTER(d)]
        L.flint_reactor_set_launch.argtypes = [vp, i, i]
        L.flint_reactor_get_launch.argtypes = [vp, C.POINTER(i), C.POINTER(i)]
        L.flint_reactor_set_newton.argtypes = [vp, i]
        L.flint_reactor_set_order.argtypes = [vp, i]
        L.flint_cea_setup.argtypes = [vp, vp, vp]
        L.flint_cea_solve.argtypes = [vp, l, l, i, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.flint_source_host.argtypes = [vp, l, vp, vp, vp]
        L.flint_rhs_host.argtypes = [vp, l, vp, vp, vp]
        L.flint_general_setup.argtypes = [vp, i, i, i, vp, vp, vp]
        L.flint_source_general.argtypes = [vp, l, l, vp, vp, vp, vp, vp]
        L.flint_source_general_host.argtypes = [vp, l, vp, vp, vp]
        L.flint_cons2prim.argtypes = [vp, l, l, i] + [vp] * 7
        L.flint_prim2cons.argtypes = [vp, l, l, i, vp, vp, vp, vp]
        L.flint_cons2prim_host.argtypes = [vp, l, i] + [vp] * 5
        L.flint_transport_setup.argtypes = [vp, i, i, vp, vp]
        L.flint_transport.argtypes = [vp, l, l] + [vp] * 7
        L.flint_transport_host.argtypes = [vp, l] + [vp] * 6
        L.flint_prim2cons_host.argtypes = [vp, l, i, vp, vp]
        L.flint_reactor_advance_host.argtypes = [vp, l, i, vp, vp, d, d, d, i, vp]
        L.flint_cea_solve_host.argtypes = [vp, l, i, vp, vp, vp, vp, vp, vp]
        L.flint_cea_solve_warm.argtypes = [vp, vp, l, l, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.flint_cea_warm_options.argtypes = [vp, i, i]
        L.flint_cea_set_tolerance.argtypes = [vp, d]
        _lib = L
    return _lib


def check(rc):
    if rc != FLINT_OK:
        msg = lib().flint_last_error().decode(errors="replace")
        raise FlintError(f"flint error {rc} ({ERRORS.get(rc, '?')}): {msg}")


def _p(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(C.c_void_p)
    return C.c_void_p(a.data_ptr())


def _stream(stream):
    if stream is None:
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return C.c_void_p(stream)
    return C.c_void_p(stream.cuda_stream)


def _chk(t, name, ld, rows=None, dtype="float64", optional=False):
    """A device array the kernels read or write: [rows, ld] (structure of arrays with leading
    dimension ld), or, with rows None, a vector of at least ld entries (callers pass the cell
    count n: the kernels index vectors by cell); a contiguous torch tensor of the given dtype on
    the GPU.  Anything else would make the kernel read or write past the allocation."""
    if t is None:
        if optional:
            return
        raise FlintError(f"{name}: array required")
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise FlintError(f"{name}: must be a torch tensor on the GPU")
    if t.dtype != getattr(torch, dtype):
        raise FlintError(f"{name}: dtype {t.dtype}, need {dtype}")
    if not t.is_contiguous():
        raise FlintError(f"{name}: must be contiguous")
    if rows is None:
        if t.dim() != 1 or t.shape[0] < ld:
            raise FlintError(f"{name}: shape {tuple(t.shape)}, need [>= {ld}]")
    elif t.dim() != 2 or t.shape[0] != rows or t.shape[1] != ld:
        raise FlintError(f"{name}: shape {tuple(t.shape)}, need [{rows}, {ld}]")


def _fortran(a):
    """numpy [nT, ncol] (row = T) -> FLINT column-major tab(T, col) as a flat C buffer."""
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)


class Mechanism:
    """Opaque mechanism handle: tables on the device plus the compiled kernels for its name."""

    def __init__(self, handle, ns, name):
        self._h = handle
        self.ns = ns
        self.nz = ns + 1
        self.name = name

    @classmethod
    def from_data(cls, md):
        """From a flint_amd.mechdata.MechData (generated tables)."""
        L = lib()
        nTk = md.kf.shape[0]
        kf, kb = _fortran(md.kf), _fortran(md.kb)
        troe = np.ascontiguousarray(np.transpose(md.troe, (2, 1, 0))) if md.n_troe else None
        lind = np.ascontiguousarray(np.transpose(md.lind, (2, 1, 0))) if md.n_lind else None
        cp, h = _fortran(md.cp), _fortran(md.h)
        comp = np.ascontiguousarray(md.composition.T) if md.composition.size else None
        W = np.ascontiguousarray(md.W)
        out = C.c_void_p()
        check(L.flint_mech_create(md.desc["name"].encode(), md.ns, _p(W), md.Tk_min, md.Tk_min + nTk - 1, md.n_arrh,
                                  _p(kf), _p(kb), md.n_troe, _p(troe), md.n_lind, _p(lind), md.Th_min, md.Th_max,
                                  _p(cp), _p(h), len(md.elements), _p(comp), C.byref(out)))
        mech = cls(out, md.ns, md.desc["name"])
        if md.composition.size:
            from .elements import element_weights
            rc = L.flint_cea_setup(out, _p(_fortran(md.s)), _p(element_weights(md.elements)))
            if rc != 5:  # 5: no compiled equilibrium kernel (the hand-written WD routine)
                check(rc)
        return mech

    @classmethod
    def load(cls, folder, prefix=""):
        """From FLINT's native input folder (read_idealgas_thermo + read_chemistry + Assign_Mechanism)."""
        L = lib()
        out = C.c_void_p()
        check(L.flint_mech_load(str(folder).encode(), prefix.encode(), C.byref(out)))
        ns = C.c_int()
        check(L.flint_mech_query(out, C.byref(ns), None, None, None, None))
        return cls(out, ns.value, os.path.basename(str(folder)))

    def query(self):
        v = [C.c_int() for _ in range(5)]
        check(lib().flint_mech_query(self._h, *[C.byref(x) for x in v]))
        return dict(zip(("ns", "Tk_min", "Tk_max", "Th_min", "Th_max"), (x.value for x in v)))

    def close(self):
        if self._h:
            lib().flint_mech_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- batched device calls (torch tensors, SoA [rows, ld]) ----
    def source(self, rho, T, wdot, hrr=None, status=None, n=None, stream=None):
        ld = rho.shape[-1]
        n = ld if n is None else n
        _chk(rho, "rho", ld, self.ns), _chk(T, "T", n), _chk(wdot, "wdot", ld, self.ns)
        _chk(hrr, "hrr", n, optional=True), _chk(status, "status", n, dtype="int32", optional=True)
        check(lib().flint_source(self._h, n, ld, _p(rho), _p(T), _p(wdot), _p(hrr), _p(status), _stream(stream)))

    def rhs(self, Z, F, J=None, status=None, n=None, stream=None):
        ld = Z.shape[-1]
        n = ld if n is None else n
        _chk(Z, "Z", ld, self.nz), _chk(F, "F", ld, self.nz), _chk(J, "J", ld, self.nz * self.nz, optional=True)
        _chk(status, "status", n, dtype="int32", optional=True)
        check(lib().flint_rhs(self._h, n, ld, _p(Z), _p(F), _p(J), _p(status), _stream(stream)))

    # ---- the data-driven general routine ----
    def general_setup(self, tables):
        """tables = (na, nt, nl, ni1, ni2, eps) as flint_amd.mechanism.general_tables returns them
        (flint_general_setup)."""
        na, nt, nl, ni1, ni2, eps = tables
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (ni1, ni2, eps)]
        check(lib().flint_general_setup(self._h, int(na), int(nt), int(nl), *(_p(x) for x in a)))

    def source_general(self, rho, T, wdot, status=None, n=None, stream=None):
        """rho [ns, ld], T [ld] -> wdot [ns, ld] with the general routine (flint_source_general)."""
        ld = rho.shape[-1]
        n = ld if n is None else n
        _chk(rho, "rho", ld, self.ns), _chk(T, "T", n), _chk(wdot, "wdot", ld, self.ns)
        _chk(status, "status", n, dtype="int32", optional=True)
        check(lib().flint_source_general(self._h, n, ld, _p(rho), _p(T), _p(wdot), _p(status), _stream(stream)))

    # ---- conservative <-> primitive variables ----
    def cons2prim(self, cons, T, prim, T_out=None, iters=None, status=None, n=None, stream=None):
        """cons [ns+nd+1, ld] (rho_k, rho u, rho E) -> prim (rho_k, u, p), Newton on T from T
        (flint_cons2prim)."""
        ld = cons.shape[-1]
        n = ld if n is None else n
        nd = cons.shape[0] - self.ns - 1
        _chk(cons, "cons", ld, cons.shape[0]), _chk(T, "T", n), _chk(prim, "prim", ld, cons.shape[0])
        _chk(T_out, "T_out", n, optional=True), _chk(iters, "iters", n, dtype="int32", optional=True)
        _chk(status, "status", n, dtype="int32", optional=True)
        check(lib().flint_cons2prim(self._h, n, ld, nd, _p(cons), _p(T), _p(prim), _p(T_out), _p(iters), _p(status),
                                    _stream(stream)))

    def prim2cons(self, prim, cons, status=None, n=None, stream=None):
        """prim [ns+nd+1, ld] (rho_k, u, p) -> cons (rho_k, rho u, rho E) (flint_prim2cons)."""
        ld = prim.shape[-1]
        n = ld if n is None else n
        nd = prim.shape[0] - self.ns - 1
        _chk(prim, "prim", ld, prim.shape[0]), _chk(cons, "cons", ld, prim.shape[0])
        _chk(status, "status", n, dtype="int32", optional=True)
        check(lib().flint_prim2cons(self._h, n, ld, nd, _p(prim), _p(cons), _p(status), _stream(stream)))

    # ---- mixture transport (Wilke) ----
    def transport_setup(self, tables):
        """Species tables (Tt_min, Tt_max, mu_tab [ns, nT], k_tab [ns, nT]), temperature fastest
        (flint_transport_setup)."""
        Tt_min, Tt_max, mu_tab, k_tab = tables
        mu_tab = np.ascontiguousarray(mu_tab, dtype=np.float64)
        k_tab = np.ascontiguousarray(k_tab, dtype=np.float64)
        if mu_tab.shape != (self.ns, Tt_max - Tt_min + 1) or k_tab.shape != mu_tab.shape:
            raise FlintError("transport tables must be [ns, Tt_max - Tt_min + 1]")
        check(lib().flint_transport_setup(self._h, int(Tt_min), int(Tt_max), _p(mu_tab), _p(k_tab)))

    def transport(self, rhoi, rho, T, mu, lam, status=None, n=None, stream=None):
        """rhoi [ns, ld], rho [ld], T [ld] (device) -> mu, lam [ld] by Wilke's rule (flint_transport)."""
        ld = rhoi.shape[-1]
        n = ld if n is None else n
        _chk(rhoi, "rhoi", ld, self.ns), _chk(rho, "rho", n), _chk(T, "T", n), _chk(mu, "mu", n)
        _chk(lam, "lam", n), _chk(status, "status", n, dtype="int32", optional=True)
        check(lib().flint_transport(self._h, n, ld, _p(rhoi), _p(rho), _p(T), _p(mu), _p(lam), _p(status),
                                    _stream(stream)))

    def transport_host(self, rhoi, rho, T):
        """rhoi [ns, n], rho [n], T [n] (numpy) -> (mu, lam, status)."""
        rhoi = np.ascontiguousarray(rhoi, dtype=np.float64)
        n = rhoi.shape[1]
        mu, lam, st = np.empty(n), np.empty(n), np.empty(n, dtype=np.int32)
        check(lib().flint_transport_host(self._h, n, _p(rhoi), _p(np.ascontiguousarray(rho, dtype=np.float64)),
                                         _p(np.ascontiguousarray(T, dtype=np.float64)), _p(mu), _p(lam), _p(st)))
        return mu, lam, st

    # ---- batched reactors ----
    CV, CP = 0, 1

    def advance(self, X, p=None, mode=0, t_end=1e-4, rtol=1e-6, atol=1e-12, max_steps=0, stats=None, status=None,
                n=None, graph=False, stream=None):
        """Advance X [ns+1, ld] (torch, device) in place from 0 to t_end."""
        ld = X.shape[-1]
        n = ld if n is None else n
        _chk(X, "X", ld, self.nz), _chk(p, "p", n, optional=mode == 0)
        _chk(stats, "stats", ld, 8, dtype="int32", optional=True)
        _chk(status, "status", n, dtype="int32", optional=True)
        fn = lib().flint_reactor_advance_graph if graph else lib().flint_reactor_advance
        check(fn(self._h, n, ld, mode, _p(X), _p(p), t_end, rtol, atol, max_steps, _p(stats), _p(status),
                 _stream(stream)))

    def reactor_jacobian(self, X, F, J, p=None, mode=0, n=None, stream=None):
        """F [ns+1, ld] and J [(ns+1)^2, ld] of the reactor model at X [ns+1, ld] (torch, device)."""
        ld = X.shape[-1]
        n = ld if n is None else n
        _chk(X, "X", ld, self.nz), _chk(p, "p", n, optional=mode == 0), _chk(F, "F", ld, self.nz)
        _chk(J, "J", ld, self.nz * self.nz, optional=True)
        check(lib().flint_reactor_jacobian(self._h, n, ld, mode, _p(X), _p(p), _p(F), _p(J), _stream(stream)))

    def check_step(self, X, Xp, p, dt, flag, tol_y=5e-5, tol_t=0.5, check_out=None, n=None, stream=None):
        """The FP64 judge of a candidate constant-pressure step X -> Xp over dt ([ns+1, ld] each):
        flag [>= n] int32 gets 0 where the backward-Euler and trapezoidal Newton corrections from Xp
        stay within (tol_y, tol_t), else 1; check_out [2, ld] the larger correction's max |dY_k|
        and |dT|."""
        ld = X.shape[-1]
        n = ld if n is None else n
        _chk(X, "X", ld, self.nz), _chk(Xp, "Xp", ld, self.nz), _chk(p, "p", n)
        _chk(flag, "flag", n, dtype="int32"), _chk(check_out, "check_out", ld, 2, optional=True)
        check(lib().flint_reactor_check_step(self._h, n, ld, _p(X), _p(Xp), _p(p), dt, tol_y, tol_t, _p(flag),
                                             _p(check_out), _stream(stream)))

    def autotune(self, X, p=None, mode=0, t_end=1e-4, rtol=1e-6, atol=1e-12):
        ld = X.shape[-1]
        _chk(X, "X", ld, self.nz), _chk(p, "p", ld, optional=mode == 0)
        best = C.c_double()
        check(lib().flint_reactor_autotune(self._h, ld, ld, mode, _p(X), _p(p), t_end, rtol, atol, C.byref(best)))
        return best.value, self.get_launch()

    def set_launch(self, teams_per_block, blocks):
        check(lib().flint_reactor_set_launch(self._h, teams_per_block, blocks))

    def set_newton(self, policy):
        """Newton policy of the reactor's BDF: 1 as SUNDIALS CVODE (default), 0 as scipy's BDF."""
        check(lib().flint_reactor_set_newton(self._h, policy))

    def set_order(self, policy):
        """Work order of advance: 0 the caller's order, 1 by descending |dT/dt| at the initial
        state (flint_reactor_set_order; results are identical)."""
        check(lib().flint_reactor_set_order(self._h, policy))

    def get_launch(self):
        a, b = C.c_int(), C.c_int()
        check(lib().flint_reactor_get_launch(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    # ---- batched equilibrium ----
    def equilibrium(self, T, rho, T_out, Y_out, p=None, hp=True, guess=None, iters=None, status=None, n=None,
                    stream=None):
        """CEA_solve over a batch (torch, device): T [ld], rho [ns, ld] -> T_out [ld], Y_out [ns, ld]."""
        ld = T.shape[-1]
        n = ld if n is None else n
        _chk(T, "T", n), _chk(rho, "rho", ld, self.ns), _chk(T_out, "T_out", n), _chk(Y_out, "Y_out", ld, self.ns)
        _chk(p, "p", n, optional=not hp), _chk(guess, "guess", ld, self.nz, optional=True)
        _chk(iters, "iters", n, dtype="int32", optional=True), _chk(status, "status", n, dtype="int32", optional=True)
        check(lib().flint_cea_solve(self._h, n, ld, int(hp), _p(p), _p(T), _p(rho), _p(guess), _p(T_out), _p(Y_out),
                                    _p(iters), _p(status), _stream(stream)))

    def equilibrium_warm(self, net, scaling, T, rho, T_out, Y_out, p, iters=None, status=None, n=None, stream=None):
        """HP equilibria warm-started by the network net (flint_amd.mlp.MLP, ns+2 -> ns+1) with the
        standardisation scaling = (in_mean, in_scale, out_mean, out_scale) (flint_cea_solve_warm)."""
        ld = T.shape[-1]
        n = ld if n is None else n
        sc = np.ascontiguousarray(np.concatenate([np.asarray(a, dtype=np.float64) for a in scaling]))
        if sc.size != 4 * self.ns + 6:
            raise FlintError("scaling must hold 4 ns + 6 values")
        check(lib().flint_cea_solve_warm(self._h, net._h, n, ld, _p(p), _p(T), _p(rho), _p(sc), _p(T_out), _p(Y_out),
                                         _p(iters), _p(status), _stream(stream)))

    def cea_tolerance(self, tol):
        """Convergence tests of the equilibrium solves: 0 the reference's, > 0 all at tol
        (flint_cea_set_tolerance)."""
        check(lib().flint_cea_set_tolerance(self._h, float(tol)))

    def warm_options(self, max_iter=6, fused=True):
        """Iteration cap of the warm-started pass of equilibrium_warm, and whether its features,
        network and Newton start run as one kernel (flint_cea_warm_options)."""
        check(lib().flint_cea_warm_options(self._h, int(max_iter), int(bool(fused))))

    # ---- host-memory convenience (numpy, SoA [rows, n]) ----
    def source_host(self, rho, T):
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        T = np.ascontiguousarray(T, dtype=np.float64)
        w = np.empty_like(rho)
        check(lib().flint_source_host(self._h, T.size, _p(rho), _p(T), _p(w)))
        return w

    def source_general_host(self, rho, T):
        """The general routine over host arrays rho [ns, n], T [n] (flint_source_general_host)."""
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        T = np.ascontiguousarray(T, dtype=np.float64)
        w = np.empty_like(rho)
        check(lib().flint_source_general_host(self._h, T.size, _p(rho), _p(T), _p(w)))
        return w

    def rhs_host(self, Z, jac=False):
        Z = np.ascontiguousarray(Z, dtype=np.float64)
        n = Z.shape[1]
        F = np.empty_like(Z)
        J = np.empty((self.nz * self.nz, n)) if jac else None
        check(lib().flint_rhs_host(self._h, n, _p(Z), _p(F), _p(J)))
        return (F, J) if jac else F

    def cons2prim_host(self, cons, T):
        """cons [ns+nd+1, n] (numpy) -> (prim, T_out, status)."""
        cons = np.ascontiguousarray(cons, dtype=np.float64)
        npr, n = cons.shape
        prim, T_out = np.empty_like(cons), np.empty(n)
        st = np.empty(n, dtype=np.int32)
        check(lib().flint_cons2prim_host(self._h, n, npr - self.ns - 1, _p(cons),
                                         _p(np.ascontiguousarray(T, dtype=np.float64)), _p(prim), _p(T_out), _p(st)))
        return prim, T_out, st

    def prim2cons_host(self, prim):
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        npr, n = prim.shape
        cons = np.empty_like(prim)
        check(lib().flint_prim2cons_host(self._h, n, npr - self.ns - 1, _p(prim), _p(cons)))
        return cons

    def advance_host(self, X, p=None, mode=0, t_end=1e-4, rtol=1e-6, atol=1e-12, max_steps=0):
        """X [ns+1, n] (numpy) advanced from 0 to t_end; returns (X_new, status)."""
        X = np.array(X, dtype=np.float64, order="C")
        n = X.shape[1]
        p = None if p is None else np.ascontiguousarray(p, dtype=np.float64)
        st = np.empty(n, dtype=np.int32)
        check(lib().flint_reactor_advance_host(self._h, n, mode, _p(X), _p(p), t_end, rtol, atol, max_steps, _p(st)))
        return X, st

    def equilibrium_host(self, T, rho, p=None, hp=True):
        """CEA_solve over host arrays T [n], rho [ns, n]; returns (T_out, Y_out, status)."""
        T = np.ascontiguousarray(T, dtype=np.float64)
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        p = None if p is None else np.ascontiguousarray(p, dtype=np.float64)
        To, Yo, st = np.empty_like(T), np.empty_like(rho), np.empty(T.size, dtype=np.int32)
        check(lib().flint_cea_solve_host(self._h, T.size, int(hp), _p(p), _p(T), _p(rho), _p(To), _p(Yo), _p(st)))
        return To, Yo, st
