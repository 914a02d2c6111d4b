"""ctypes front end of the CPU oracle (oracle/gp_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker or the CPU baseline,
never as part of the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libgp_oracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    lib.orc_create.restype = C.c_void_p
    lib.orc_create.argtypes = [C.c_void_p]
    lib.orc_destroy.argtypes = [C.c_void_p]
    lib.orc_nnz.argtypes = [C.c_void_p]
    lib.orc_pattern.argtypes = [C.c_void_p, _ip, _ip]
    lib.orc_residual.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, C.c_double, _dp]
    lib.orc_jacobian.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_double, _dp]
    lib.orc_linsolve.argtypes = [C.c_void_p, _dp, _dp, _dp]
    lib.orc_integrate.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_int, _ip, _dp, _ip, _ip,
                                  C.c_int, C.c_double, C.c_int, _dp, _dp, _dp]
    lib.orc_schur_step.argtypes = [C.c_void_p, _dp, _dp, _dp]
    lib.orc_set_trips.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _ip, _ip, _dp]
    lib.orc_set_drops.argtypes = [C.c_void_p, C.c_int, _ip, _ip]
    lib.orc_set_stop_on_fail.argtypes = [C.c_void_p, C.c_int]
    lib.orc_linsolve_t.argtypes = [C.c_void_p, _dp, _dp, _dp]
    lib.orc_adjoint.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_int, _ip, _dp, _ip, _ip, C.c_int,
                                C.c_double, C.c_int, C.c_int, _ip, C.c_double, C.c_double, C.c_int, _ip,
                                C.c_int, _dp, _dp, _dp]
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _load()
    return _LIB


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


class Oracle:
    """CPU restatement of the reference integrator for one system."""

    def __init__(self, desc):
        self.desc = desc
        self._cs, self._keep = desc.as_ctypes()
        self.h = lib().orc_create(C.byref(self._cs))
        self.n = desc.n
        self.nnz = lib().orc_nnz(self.h)
        self.cp = np.zeros(self.n + 1, np.int32)
        self.ri = np.zeros(self.nnz, np.int32)
        lib().orc_pattern(self.h, _i(self.cp), _i(self.ri))
        self.p = np.ascontiguousarray(desc.p, dtype=np.float64)

    def __del__(self):
        h = getattr(self, "h", None)
        if h and _LIB is not None:
            _LIB.orc_destroy(h)
            self.h = None

    def _fault_g(self, fault_g):
        g = np.zeros(self.desc.nbus)
        if fault_g is not None:
            g[:] = fault_g
        return g

    def residual(self, z, zold, dt, fault_g=None, p=None):
        z = np.ascontiguousarray(z, np.float64)
        zold = np.ascontiguousarray(zold, np.float64)
        p = self.p if p is None else np.ascontiguousarray(p, np.float64)
        f = np.zeros(self.n)
        g = self._fault_g(fault_g)
        lib().orc_residual(self.h, _d(z), _d(zold), _d(p), _d(g), dt, _d(f))
        return f

    def jacobian(self, z, dt, fault_g=None, p=None):
        """Backward-Euler Jacobian values in the oracle's CSC pattern."""
        z = np.ascontiguousarray(z, np.float64)
        p = self.p if p is None else np.ascontiguousarray(p, np.float64)
        J = np.zeros(self.nnz)
        g = self._fault_g(fault_g)
        lib().orc_jacobian(self.h, _d(z), _d(p), _d(g), dt, _d(J))
        return J

    def jacobian_csc(self, z, dt, fault_g=None, p=None):
        import scipy.sparse as sp
        return sp.csc_matrix((self.jacobian(z, dt, fault_g, p=p), self.ri, self.cp), shape=(self.n, self.n))

    def linsolve(self, Jvals, b):
        x = np.zeros(self.n)
        rc = lib().orc_linsolve(self.h, _d(np.ascontiguousarray(Jvals)), _d(np.ascontiguousarray(b)), _d(x))
        if rc:
            raise RuntimeError("singular matrix")
        return x

    def schur_step(self, Jvals, f):
        dz = np.zeros(self.n)
        lib().orc_schur_step(self.h, _d(np.ascontiguousarray(Jvals)), _d(np.ascontiguousarray(f)), _d(dz))
        return dz

    def _set_trips(self, trips):
        """trips: iterable of (from_bus, to_bus, step_trip, step_reclose (0:
        never), y[8] = yff, yft, ytf, ytt as re, im)."""
        tr = list(trips)
        fr = np.array([t[0] for t in tr] or [0], np.int32)
        to = np.array([t[1] for t in tr] or [0], np.int32)
        on = np.array([t[2] for t in tr] or [0], np.int32)
        off = np.array([t[3] for t in tr] or [0], np.int32)
        y = np.ascontiguousarray(np.array([list(t[4]) for t in tr], np.float64).reshape(-1) if tr else np.zeros(8))
        lib().orc_set_trips(self.h, len(tr), _i(fr), _i(to), _i(on), _i(off), _d(y))

    def _set_drops(self, drops):
        """drops: iterable of (ZIP table row, step) or (kind, row, step) with
        kind "zip", "sexs" or "esdc1a" (row of that device table)."""
        from gradpower_hip.desc import GP_ESDC1A, GP_SEXS, GP_ZIPLOAD
        nd = [len(self.desc.dev[t]) for t in (GP_ZIPLOAD, GP_SEXS, GP_ESDC1A)]
        base = {"zip": 0, "sexs": nd[0], "esdc1a": nd[0] + nd[1]}
        d = [(base[x[0]] + int(x[1]), int(x[2])) if len(x) == 3 else (int(x[0]), int(x[1])) for x in drops]
        k = np.array([x[0] for x in d] or [0], np.int32)
        st = np.array([x[1] for x in d] or [0], np.int32)
        lib().orc_set_drops(self.h, len(d), _i(k), _i(st))

    def integrate(self, dt, nsteps, events=(), z0=None, tol=1e-10, itermax=50, traj=True, p=None, trips=(),
                  drops=(), stop_on_fail=False):
        """events: iterable of (bus, r_fault, step_on, step_off); trips: line
        trips, see _set_trips; drops: ZIP load disconnects, see _set_drops.
        stop_on_fail: stop at the first Newton solve that does not converge
        (the reference keeps integrating).  Returns (traj [(nsteps+1) x n]
        or None, zfinal, stats dict); stats["first_fail_step"] is the number
        of steps completed before the first such solve (-1: none)."""
        self._set_trips(trips)
        self._set_drops(drops)
        lib().orc_set_stop_on_fail(self.h, 1 if stop_on_fail else 0)
        z0 = np.ascontiguousarray(self.desc.z0 if z0 is None else z0, np.float64)
        p = self.p if p is None else np.ascontiguousarray(p, np.float64)
        ev = list(events)
        eb = np.array([e[0] for e in ev] or [0], np.int32)
        er = np.array([e[1] for e in ev] or [1.0], np.float64)
        eon = np.array([e[2] for e in ev] or [0], np.int32)
        eoff = np.array([e[3] for e in ev] or [0], np.int32)
        tr = np.zeros((nsteps + 1, self.n)) if traj else None
        zf = np.zeros(self.n)
        st = np.zeros(4)
        rc = lib().orc_integrate(self.h, _d(z0), _d(p), dt, nsteps, _i(eb), _d(er), _i(eon), _i(eoff),
                                 len(ev), tol, itermax, _d(tr) if traj else None, _d(zf), _d(st))
        return tr, zf, {"rc": rc, "newton_iters": st[0], "not_converged": st[1], "max_residual": st[2],
                        "first_fail_step": int(st[3])}

    def linsolve_t(self, Jvals, c):
        """y with J^T y = c."""
        y = np.zeros(self.n)
        rc = lib().orc_linsolve_t(self.h, _d(np.ascontiguousarray(Jvals)), _d(np.ascontiguousarray(c)), _d(y))
        if rc:
            raise RuntimeError("singular matrix")
        return y

    def adjoint(self, dt, nsteps, params, events=(), rows=None, scale=1.0, offset=0.0, z0=None, tol=1e-10,
                itermax=50, p=None, fault_shift=0, trips=(), drops=()):
        """Objective scale * sum_k dt sum_r (z_k[r] - offset)^2 and its gradient
        w.r.t. p[params] by the oracle's discrete adjoint.  rows=None: the
        GENROU speed deviations.  fault_shift != 0 plants a misaligned fault
        state (tests only).  Returns (value, grad, dz0)."""
        from gradpower_hip.desc import gen_speed_indices
        self._set_trips(trips)
        self._set_drops(drops)
        z0 = np.ascontiguousarray(self.desc.z0 if z0 is None else z0, np.float64)
        p = self.p if p is None else np.ascontiguousarray(p, np.float64)
        rows = np.ascontiguousarray(gen_speed_indices(self.desc) if rows is None else rows, np.int32)
        par = np.ascontiguousarray(params, np.int32)
        ev = list(events)
        eb = np.array([e[0] for e in ev] or [0], np.int32)
        er = np.array([e[1] for e in ev] or [1.0], np.float64)
        eon = np.array([e[2] for e in ev] or [0], np.int32)
        eoff = np.array([e[3] for e in ev] or [0], np.int32)
        val = C.c_double()
        grad = np.zeros(max(1, len(par)))
        dz0 = np.zeros(self.n)
        rc = lib().orc_adjoint(self.h, _d(z0), _d(p), dt, nsteps, _i(eb), _d(er), _i(eon), _i(eoff), len(ev),
                               tol, itermax, len(rows), _i(rows), scale, offset, len(par), _i(par),
                               int(fault_shift), C.byref(val), _d(grad), _d(dz0))
        if rc:
            raise RuntimeError(f"oracle forward sweep failed ({rc})")
        return val.value, grad[:len(par)], dz0

    def objective(self, dt, nsteps, events=(), rows=None, scale=1.0, offset=0.0, z0=None, tol=1e-10,
                  itermax=50, p=None, trips=(), drops=()):
        """scale * sum_{k>=1} dt sum_r (z_k[r] - offset)^2 along the oracle's trajectory."""
        from gradpower_hip.desc import gen_speed_indices
        rows = gen_speed_indices(self.desc) if rows is None else np.asarray(rows)
        tr, _, st = self.integrate(dt, nsteps, events, z0=z0, tol=tol, itermax=itermax, p=p, trips=trips,
                                   drops=drops)
        if st["rc"]:
            raise RuntimeError(f"oracle forward sweep failed ({st['rc']})")
        e = tr[1:, rows] - offset
        return scale * dt * float((e * e).sum())
