"""ctypes binding of libgradpower_hip.so (the C ABI in include/gradpower_hip.h).

The library is required: importing this module on a machine without the
built library, or creating a context without a usable GPU, raises -- there
is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .desc import SystemDesc, CEvent, CLineTrip, CLoadDrop, CDeviceDrop

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libgradpower_hip.so")

GP_OK = 0
GP_ERR_BADARG, GP_ERR_HIP, GP_ERR_OOM, GP_ERR_SINGULAR_PIVOT = -1, -2, -3, -4
GP_ERR_NEWTON_NOT_CONVERGED, GP_ERR_NAN, GP_ERR_UNSUPPORTED = -5, -6, -7
METHOD_MONO_FP64, METHOD_SCHUR_FP64, METHOD_SCHUR_FP32, METHOD_SCHUR_SHARED = 0, 1, 2, 3
REFACTOR_EVERY_ITER, REFACTOR_FIRST_ITER, REFACTOR_ADAPTIVE, REFACTOR_ASYNC = 0, 1, 2, 3
_NAMES = {-1: "BADARG", -2: "HIP", -3: "OOM", -4: "SINGULAR_PIVOT", -5: "NEWTON_NOT_CONVERGED",
          -6: "NAN", -7: "UNSUPPORTED"}


class GpError(RuntimeError):
    def __init__(self, code, msg, stats=None):
        super().__init__(f"gp error {code} ({_NAMES.get(code, '?')}): {msg}")
        self.code = code
        self.stats = stats


class COptions(C.Structure):
    _fields_ = [("record_all", C.c_int32), ("n_record", C.c_int32),
                ("record_idx", C.POINTER(C.c_int32)), ("refactor_policy", C.c_int32),
                ("stall_window", C.c_int32), ("profile", C.c_int32), ("lock_step", C.c_int32),
                ("dense_block", C.c_int32), ("retry_window", C.c_int32), ("global_x", C.c_int32),
                ("refactor_theta", C.c_double),
                ("async_below", C.c_double), ("predictor", C.c_int32), ("pivot_rule", C.c_int32),
                ("solve_levels", C.c_int32)]


class CStats(C.Structure):
    _fields_ = [("newton_iters", C.c_int64), ("scenario_steps", C.c_int64),
                ("event_resolves", C.c_int64), ("refactorizations", C.c_int64),
                ("n_not_converged", C.c_int32), ("n_nan", C.c_int32),
                ("max_final_residual", C.c_double), ("elapsed_ms", C.c_double),
                ("solve_ms", C.c_double), ("newton_rounds", C.c_int64), ("factor_launches", C.c_int64),
                ("solve_kernel_ms", C.c_double), ("solve_kernel_launches", C.c_int64),
                ("solve_kernel_scenarios", C.c_int64), ("factor_kernel_ms", C.c_double),
                ("factor_kernel_launches", C.c_int64), ("factor_kernel_scenarios", C.c_int64),
                ("solve_bytes_per_scenario", C.c_double), ("factor_bytes_per_scenario", C.c_double),
                ("exact_retries", C.c_int64), ("repivots", C.c_int64)]


class CSymInfo(C.Structure):
    _fields_ = [("n", C.c_int32), ("nnz_j", C.c_int32), ("nnz_lu_j", C.c_int32),
                ("levels_j", C.c_int32), ("n_red", C.c_int32), ("nnz_s", C.c_int32),
                ("nnz_lu_s", C.c_int32), ("levels_s", C.c_int32), ("n_clusters_nt", C.c_int32),
                ("w_total_nt", C.c_int32), ("nd_j", C.c_int32), ("nd_s", C.c_int32), ("nd_jt", C.c_int32),
                ("res_j", C.c_int32), ("res_s", C.c_int32), ("ld_scen", C.c_int32),
                ("solve_levels_j", C.c_int32), ("solve_levels_s", C.c_int32)]


class CObjective(C.Structure):
    _fields_ = [("n_rows", C.c_int32), ("rows", C.POINTER(C.c_int32)), ("scale", C.c_double),
                ("offset", C.c_double)]


class CAdjStats(C.Structure):
    _fields_ = [("elapsed_ms", C.c_double), ("forward_ms", C.c_double), ("backward_ms", C.c_double),
                ("scenario_steps", C.c_int64), ("adjoint_steps", C.c_int64), ("n_segments", C.c_int32),
                ("n_excluded", C.c_int32), ("newton_iters", C.c_int64), ("reserved", C.c_int64 * 4)]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint8)
_LIB = None

EXPORTS = ["gp_create", "gp_destroy", "gp_set_state", "gp_set_events", "gp_set_line_trips", "gp_set_load_drops",
           "gp_set_device_drops", "gp_integrate",
           "gp_get_trajectory", "gp_get_state", "gp_eval_residual", "gp_jacobian_pattern",
           "gp_eval_jacobian", "gp_newton_step", "gp_get_symbolic_info", "gp_get_scenario_status",
           "gp_copy_kernel", "gp_adjoint", "gp_analyse_pivots", "gp_eval_residual_noalloc", "gp_last_error",
           "gp_version"]


def load():
    """Load the library (raises FileNotFoundError/OSError if it is not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBPATH):
        raise FileNotFoundError(f"{LIBPATH} is not built; run __graft_entry__.build() "
                                f"or `make -C gradpower.jl_amd`")
    lib = C.CDLL(LIBPATH)
    lib.gp_create.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.POINTER(C.c_void_p)]
    lib.gp_destroy.argtypes = [C.c_void_p]
    lib.gp_destroy.restype = None
    lib.gp_set_state.argtypes = [C.c_void_p, _dp, C.c_int32, _dp]
    lib.gp_set_events.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, _ip]
    lib.gp_set_line_trips.argtypes = [C.c_void_p, C.c_void_p, C.c_int32]
    lib.gp_set_load_drops.argtypes = [C.c_void_p, C.c_void_p, C.c_int32]
    lib.gp_set_device_drops.argtypes = [C.c_void_p, C.c_void_p, C.c_int32]
    lib.gp_integrate.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_double, C.c_int32, C.c_void_p]
    lib.gp_get_trajectory.argtypes = [C.c_void_p, C.c_int32, _dp, C.c_int64]
    lib.gp_get_state.argtypes = [C.c_void_p, _dp]
    lib.gp_eval_residual.argtypes = [C.c_void_p, _dp, _dp, C.c_double, _up, _dp]
    lib.gp_jacobian_pattern.argtypes = [C.c_void_p, _ip, _ip, _ip]
    lib.gp_eval_jacobian.argtypes = [C.c_void_p, _dp, C.c_double, _up, _dp]
    lib.gp_get_symbolic_info.argtypes = [C.c_void_p, C.c_void_p]
    lib.gp_get_scenario_status.argtypes = [C.c_void_p, _ip, _ip]
    lib.gp_copy_kernel.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
    lib.gp_newton_step.argtypes = [C.c_void_p, C.c_int32, _dp, _dp, C.c_double, _up, _dp]
    lib.gp_adjoint.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_double, C.c_int32, C.c_int32,
                               C.c_void_p, C.c_int32, _ip, _dp, _dp, _dp, _dp, C.c_void_p]
    lib.gp_analyse_pivots.argtypes = [C.c_int32, _ip, _ip, _dp, C.c_int32, C.c_void_p, C.c_double, _ip, _ip]
    lib.gp_eval_residual_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p,
                                             C.c_void_p, C.c_void_p]
    lib.gp_last_error.restype = C.c_char_p
    lib.gp_version.restype = C.c_char_p
    _LIB = lib
    return lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _stats(s):
    return {f: getattr(s, f) for f, _ in CStats._fields_}


class HipContext:
    """One library context: a system, M scenarios, device state on the current GPU."""

    def __init__(self, desc: SystemDesc, M: int, record_idx=None,
                 refactor_policy=REFACTOR_EVERY_ITER, stall_window=0, refactor_theta=0.0, async_below=0.0,
                 profile=False, lock_step=False, dense_block=0, retry_window=0, global_x=False,
                 predictor=0, pivot_rule=0, solve_levels=0):
        self.lib = load()
        self.desc = desc
        self.M = int(M)
        self.n = desc.n
        self._cs, self._keep = desc.as_ctypes()
        opt = COptions()
        if record_idx is None:
            opt.record_all = 1
            self.nrec = desc.n
        else:
            self._rec = np.ascontiguousarray(record_idx, dtype=np.int32)
            opt.record_all = 0
            opt.n_record = len(self._rec)
            opt.record_idx = _i(self._rec)
            self.nrec = len(self._rec)
        opt.refactor_policy = refactor_policy
        opt.stall_window = int(stall_window)
        opt.refactor_theta = float(refactor_theta)
        opt.async_below = float(async_below)
        opt.profile = int(profile) if profile else 0   # k: every k-th solve launch timed
        opt.lock_step = 1 if lock_step else 0
        opt.dense_block = int(dense_block)
        opt.retry_window = int(retry_window)
        opt.global_x = 1 if global_x else 0
        opt.predictor = int(predictor)
        opt.pivot_rule = int(pivot_rule)
        opt.solve_levels = int(solve_levels)
        self._opt = opt
        h = C.c_void_p()
        self._check(self.lib.gp_create(C.byref(self._cs), self.M, C.byref(opt), C.byref(h)))
        self.h = h
        self.nsteps = np.zeros(self.M, np.int32)

    def _check(self, rc, stats=None):
        if rc != GP_OK:
            raise GpError(rc, self.lib.gp_last_error().decode(), stats)

    def close(self):
        if getattr(self, "h", None):
            self.lib.gp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, z0, p, per_scenario=False):
        z0 = np.ascontiguousarray(z0, np.float64)
        p = np.ascontiguousarray(p, np.float64)
        self._check(self.lib.gp_set_state(self.h, _d(z0), int(per_scenario), _d(p)))

    def set_events(self, events, nsteps):
        """events: iterable of (scenario, bus, r_fault, step_on, step_off)."""
        ev = list(events)
        arr = (CEvent * max(1, len(ev)))()
        for i, (m, b, r, on, off) in enumerate(ev):
            arr[i].scenario, arr[i].bus, arr[i].r_fault = int(m), int(b), float(r)
            arr[i].step_on, arr[i].step_off = int(on), int(off)
        self.nsteps = np.ascontiguousarray(nsteps, np.int32)
        self._check(self.lib.gp_set_events(self.h, C.byref(arr), len(ev), _i(self.nsteps)))

    def set_line_trips(self, trips):
        """trips: iterable of (scenario, from_bus, to_bus, step_trip,
        step_reclose (0: never), y) with y the line's stamp (yff, yft, ytf,
        ytt as 4 complex or 8 floats, see system.TripLineEvent.stamp).
        Call after set_events, which clears them."""
        tr = list(trips)
        arr = (CLineTrip * max(1, len(tr)))()
        for i, (m, a, b, on, off, y) in enumerate(tr):
            yv = np.asarray(y)
            if np.iscomplexobj(yv):
                yv = np.column_stack([yv.real, yv.imag]).reshape(-1)
            yv = np.asarray(yv, np.float64).reshape(-1)
            if yv.size != 8:
                raise ValueError("a line trip's stamp has 4 complex (8 real) entries")
            arr[i].scenario, arr[i].from_bus, arr[i].to_bus = int(m), int(a), int(b)
            arr[i].step_trip, arr[i].step_reclose = int(on), int(off)
            for k in range(8):
                arr[i].y[k] = float(yv[k])
        self._check(self.lib.gp_set_line_trips(self.h, C.byref(arr), len(tr)))

    def set_load_drops(self, drops):
        """drops: iterable of (scenario, ZIP table row, step): the load goes
        offline for good at the end of that step (system.DisconnectDeviceEvent).
        Call after set_events, which clears them."""
        d = list(drops)
        arr = (CLoadDrop * max(1, len(d)))()
        for i, (m, k, step) in enumerate(d):
            arr[i].scenario, arr[i].load, arr[i].step = int(m), int(k), int(step)
        self._check(self.lib.gp_set_load_drops(self.h, C.byref(arr), len(d)))

    def set_device_drops(self, drops):
        """drops: iterable of (scenario, device type, table row, step): a ZIP
        load (GP_ZIPLOAD) or an SEXS / ESDC1A exciter goes offline for good
        at the end of that step (system.DisconnectDeviceEvent).  Call after
        set_events, which clears them; replaces earlier drops."""
        d = list(drops)
        arr = (CDeviceDrop * max(1, len(d)))()
        for i, (m, t, k, step) in enumerate(d):
            arr[i].scenario, arr[i].type, arr[i].row, arr[i].step = int(m), int(t), int(k), int(step)
        self._check(self.lib.gp_set_device_drops(self.h, C.byref(arr), len(d)))

    def integrate(self, method, dt, tol=1e-10, itermax=50, raise_on_fail=True):
        st = CStats()
        rc = self.lib.gp_integrate(self.h, int(method), float(dt), float(tol), int(itermax), C.byref(st))
        stats = _stats(st)
        stats["status"] = rc
        if rc != GP_OK and (raise_on_fail or rc in (GP_ERR_BADARG, GP_ERR_HIP, GP_ERR_OOM, GP_ERR_UNSUPPORTED)):
            raise GpError(rc, self.lib.gp_last_error().decode(), stats)
        return stats

    def trajectory(self, m):
        ln = int(self.nsteps[m]) + 1
        out = np.zeros((self.nrec, ln))
        self._check(self.lib.gp_get_trajectory(self.h, int(m), _d(out), ln))
        return out

    def scenario_status(self):
        """(status[M], steps_done[M]) of the last integrate call."""
        st = np.zeros(self.M, np.int32)
        sd = np.zeros(self.M, np.int32)
        self._check(self.lib.gp_get_scenario_status(self.h, _i(st), _i(sd)))
        return st, sd

    def state(self):
        out = np.zeros((self.M, self.n))
        self._check(self.lib.gp_get_state(self.h, _d(out)))
        return out

    def eval_residual(self, z, zold, dt, fault_on=None):
        z = np.ascontiguousarray(np.broadcast_to(z, (self.M, self.n)), np.float64)
        zold = np.ascontiguousarray(np.broadcast_to(zold, (self.M, self.n)), np.float64)
        fo = np.ascontiguousarray(np.zeros(self.M) if fault_on is None else fault_on, np.uint8)
        f = np.zeros((self.M, self.n))
        self._check(self.lib.gp_eval_residual(self.h, _d(z), _d(zold), float(dt),
                                              fo.ctypes.data_as(_up), _d(f)))
        return f

    def jacobian_pattern(self):
        nnz = C.c_int32()
        self._check(self.lib.gp_jacobian_pattern(self.h, C.byref(nnz), None, None))
        rp = np.zeros(self.n + 1, np.int32)
        ci = np.zeros(nnz.value, np.int32)
        self._check(self.lib.gp_jacobian_pattern(self.h, C.byref(nnz), _i(rp), _i(ci)))
        return rp, ci

    def eval_jacobian(self, z, dt, fault_on=None):
        rp, ci = self.jacobian_pattern()
        z = np.ascontiguousarray(np.broadcast_to(z, (self.M, self.n)), np.float64)
        fo = np.ascontiguousarray(np.zeros(self.M) if fault_on is None else fault_on, np.uint8)
        vals = np.zeros((self.M, len(ci)))
        self._check(self.lib.gp_eval_jacobian(self.h, _d(z), float(dt), fo.ctypes.data_as(_up), _d(vals)))
        return rp, ci, vals

    def newton_step(self, method, z, zold, dt, fault_on=None):
        """One Newton correction dz per scenario (z_new = z + dz)."""
        z = np.ascontiguousarray(np.broadcast_to(z, (self.M, self.n)), np.float64)
        zold = np.ascontiguousarray(np.broadcast_to(zold, (self.M, self.n)), np.float64)
        fo = np.ascontiguousarray(np.zeros(self.M) if fault_on is None else fault_on, np.uint8)
        dz = np.zeros((self.M, self.n))
        self._check(self.lib.gp_newton_step(self.h, int(method), _d(z), _d(zold), float(dt),
                                            fo.ctypes.data_as(_up), _d(dz)))
        return dz

    def symbolic_info(self):
        info = CSymInfo()
        self._check(self.lib.gp_get_symbolic_info(self.h, C.byref(info)))
        return {f: getattr(info, f) for f, _ in CSymInfo._fields_ if f != "reserved"}

    def adjoint(self, method, dt, params, tol=1e-10, itermax=50, segment=30, rows=None, scale=1.0,
                offset=0.0, want_dz0=False, raise_on_fail=True):
        """Objective and gradient by the discrete adjoint (gp_adjoint).

        Returns a dict: value (total), value_scen [M], grad [len(params)],
        grad_scen [M x len(params)], dz0 [M x n] (when want_dz0), stats, status."""
        par = np.ascontiguousarray(params, np.int32)
        obj = CObjective()
        if rows is not None:
            self._objrows = np.ascontiguousarray(rows, np.int32)
            obj.n_rows = len(self._objrows)
            obj.rows = _i(self._objrows)
        obj.scale = float(scale)
        obj.offset = float(offset)
        val = np.zeros(self.M)
        grad = np.zeros(max(1, len(par)))
        gs = np.zeros((self.M, max(1, len(par))))
        dz0 = np.zeros((self.M, self.n)) if want_dz0 else None
        st = CAdjStats()
        rc = self.lib.gp_adjoint(self.h, int(method), float(dt), float(tol), int(itermax), int(segment),
                                 C.byref(obj), len(par), _i(par) if len(par) else None, _d(val), _d(grad),
                                 _d(gs), _d(dz0) if want_dz0 else None, C.byref(st))
        stats = {f: getattr(st, f) for f, _ in CAdjStats._fields_ if f != "reserved"}
        if rc != GP_OK and (raise_on_fail or rc in (GP_ERR_BADARG, GP_ERR_HIP, GP_ERR_OOM, GP_ERR_UNSUPPORTED)):
            raise GpError(rc, self.lib.gp_last_error().decode(), stats)
        return {"value": float(val.sum()), "value_scen": val, "grad": grad[:len(par)],
                "grad_scen": gs[:, :len(par)], "dz0": dz0, "stats": stats, "status": rc}


def analyse_pivots(cp, ri, val, extra=(), threshold=1e-3):
    """Column order and static pivot rows of the library's LU analysis
    (gp_analyse_pivots; host only, no GPU): returns (q, prow) with column
    q[k] and row prow[k] at pivot k."""
    lib = load()
    cp = np.ascontiguousarray(cp, np.int32)
    ri = np.ascontiguousarray(ri, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    n = len(cp) - 1
    ex = [np.ascontiguousarray(e, np.float64) for e in extra]
    arr = (C.c_void_p * max(1, len(ex)))(*[e.ctypes.data for e in ex])
    q = np.zeros(n, np.int32)
    prow = np.zeros(n, np.int32)
    rc = lib.gp_analyse_pivots(n, _i(cp), _i(ri), _d(val), len(ex), arr, float(threshold), _i(q), _i(prow))
    if rc != GP_OK:
        raise GpError(rc, lib.gp_last_error().decode())
    return q, prow
