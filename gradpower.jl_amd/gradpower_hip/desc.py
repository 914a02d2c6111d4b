"""Flatten a set-up PowerSystem into the C-ABI system descriptor.

The descriptor is the data format that crosses into libgradpower_hip
(`gp_system_desc` in include/gradpower_hip.h): SoA device tables of absolute
0-based z indices after cluster reordering, the realified admittance in CSR,
the cluster table, and the constant controls used by unwired machines.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from .system import PowerSystem, DynamicProblem, net_ptr, system_size
from .init import genrou_u0

DEVTYPES = ["genrou", "ieesgo", "tgov1", "sexs", "esdc1a", "ieeest", "zipload", "static_gen"]
DEVCOLS = [6, 4, 4, 4, 4, 4, 2, 4]
GP_GENROU, GP_IEESGO, GP_TGOV1, GP_SEXS, GP_ESDC1A, GP_IEEEST, GP_ZIPLOAD, GP_STATICGEN = range(8)
CLUSTER_COLS = 11
(CL_WSTART, CL_WSIZE, CL_BUS, CL_TRIVIAL, CL_GENROU, CL_GOVTYPE, CL_GOV, CL_EXCTYPE,
 CL_EXC, CL_PSS, CL_SG) = range(11)


class CSystemDesc(C.Structure):
    _fields_ = [("n", C.c_int32), ("nbus", C.c_int32), ("net_ptr", C.c_int32),
                ("n_par", C.c_int32), ("is_diff", C.POINTER(C.c_uint8)),
                ("y_nnz", C.c_int32), ("y_rowptr", C.POINTER(C.c_int32)),
                ("y_colind", C.POINTER(C.c_int32)), ("y_val", C.POINTER(C.c_double)),
                ("n_dev", C.c_int32 * 8), ("dev", C.POINTER(C.c_int32) * 8),
                ("genrou_u0", C.POINTER(C.c_double)), ("n_clusters", C.c_int32),
                ("clusters", C.POINTER(C.c_int32))]


class CEvent(C.Structure):
    _fields_ = [("scenario", C.c_int32), ("bus", C.c_int32), ("r_fault", C.c_double),
                ("step_on", C.c_int32), ("step_off", C.c_int32)]


class CLineTrip(C.Structure):
    _fields_ = [("scenario", C.c_int32), ("from_bus", C.c_int32), ("to_bus", C.c_int32),
                ("step_trip", C.c_int32), ("step_reclose", C.c_int32), ("reserved", C.c_int32),
                ("y", C.c_double * 8)]


class CLoadDrop(C.Structure):
    _fields_ = [("scenario", C.c_int32), ("load", C.c_int32), ("step", C.c_int32), ("reserved", C.c_int32)]


class CDeviceDrop(C.Structure):
    _fields_ = [("scenario", C.c_int32), ("type", C.c_int32), ("row", C.c_int32), ("step", C.c_int32)]


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


@dataclass
class SystemDesc:
    n: int
    nbus: int
    net_ptr: int
    n_par: int
    is_diff: np.ndarray
    y_rowptr: np.ndarray
    y_colind: np.ndarray
    y_val: np.ndarray
    dev: list                     # 8 int32 arrays [n_t, cols_t]
    genrou_u0: np.ndarray
    clusters: np.ndarray          # int32 [n_clusters, 11]
    z0: np.ndarray = None
    p: np.ndarray = None
    meta: dict = field(default_factory=dict)

    def as_ctypes(self):
        """Return (struct, keepalive) -- keep `keepalive` referenced while the
        struct is in use."""
        keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a
        s = CSystemDesc()
        s.n, s.nbus, s.net_ptr, s.n_par = self.n, self.nbus, self.net_ptr, self.n_par
        s.is_diff = _p(arr(self.is_diff, np.uint8), C.c_uint8)
        s.y_nnz = len(self.y_colind)
        s.y_rowptr = _p(arr(self.y_rowptr, np.int32), C.c_int32)
        s.y_colind = _p(arr(self.y_colind, np.int32), C.c_int32)
        s.y_val = _p(arr(self.y_val, np.float64), C.c_double)
        for t in range(8):
            d = arr(self.dev[t].reshape(-1), np.int32)
            s.n_dev[t] = self.dev[t].shape[0]
            s.dev[t] = _p(d if d.size else arr(np.zeros(1), np.int32), C.c_int32)
        s.genrou_u0 = _p(arr(self.genrou_u0.reshape(-1) if self.genrou_u0.size else np.zeros(2),
                             np.float64), C.c_double)
        s.n_clusters = self.clusters.shape[0]
        s.clusters = _p(arr(self.clusters.reshape(-1), np.int32), C.c_int32)
        return s, keep

    @property
    def n_dev(self):
        return [d.shape[0] for d in self.dev]


def build_desc(ps: PowerSystem, dp: DynamicProblem) -> SystemDesc:
    psd = ps.dynamic
    npt = net_ptr(ps)
    n = system_size(ps)
    is_diff = np.zeros(n, dtype=np.uint8)
    is_diff[:npt] = psd.is_diff
    yr = ps.network.ybus_real.tocsr()
    yr.sort_indices()

    rows = {k: [] for k in DEVTYPES}
    local = {}                    # device index -> row in its table
    u = psd.uvec_idx
    for i, dev in enumerate(psd.devices):
        d = dev.dtype
        k = d.kind
        local[i] = len(rows[k])
        b = psd.dev_bus[i]
        vr = npt + 2 * b
        if k == "genrou":
            rows[k].append([dev.diff_ptr, dev.alg_ptr, dev.par_ptr, b,
                            int(u[dev.ctrl_ptr + 1]), int(u[dev.ctrl_ptr])])
        elif k in ("ieesgo", "tgov1", "ieeest"):
            rows[k].append([dev.diff_ptr, dev.alg_ptr, dev.par_ptr, int(u[dev.ctrl_ptr])])
        elif k in ("sexs", "esdc1a"):
            rows[k].append([dev.diff_ptr, dev.par_ptr, vr, -1])
        elif k == "zipload":
            rows[k].append([b, dev.par_ptr])
        elif k == "static_gen":
            rows[k].append([b, dev.alg_ptr, dev.par_ptr, d.bus_type])
    # PSS output wiring into its exciter's vs_idx (src/tables/ieeest.jl:95-130)
    for cl in psd.clusters:
        if cl.pss_idx >= 0 and cl.exc_idx >= 0:
            exc = psd.devices[cl.exc_idx]
            pss = psd.devices[cl.pss_idx]
            rows[exc.dtype.kind][local[cl.exc_idx]][3] = pss.alg_ptr
    dev_arrays = []
    for k, c in zip(DEVTYPES, DEVCOLS):
        a = np.array(rows[k], dtype=np.int32).reshape(-1, c)
        dev_arrays.append(a)

    cl_rows = []
    for cl in psd.clusters:
        gk = psd.devices[cl.gen_idx].dtype.kind
        r = [cl.w_start, cl.w_size, cl.bus, int(cl.trivial), -1, 0, -1, 0, -1, -1, -1]
        if gk == "genrou":
            r[CL_GENROU] = local[cl.gen_idx]
        else:
            r[CL_SG] = local[cl.gen_idx]
        if cl.gov_idx >= 0:
            r[CL_GOVTYPE] = GP_IEESGO if psd.devices[cl.gov_idx].dtype.kind == "ieesgo" else GP_TGOV1
            r[CL_GOV] = local[cl.gov_idx]
        if cl.exc_idx >= 0:
            r[CL_EXCTYPE] = GP_SEXS if psd.devices[cl.exc_idx].dtype.kind == "sexs" else GP_ESDC1A
            r[CL_EXC] = local[cl.exc_idx]
        if cl.pss_idx >= 0:
            r[CL_PSS] = local[cl.pss_idx]
        cl_rows.append(r)
    clusters = np.array(cl_rows, dtype=np.int32).reshape(-1, CLUSTER_COLS)

    return SystemDesc(n=n, nbus=ps.nbus, net_ptr=npt, n_par=len(dp.pvec), is_diff=is_diff,
                      y_rowptr=yr.indptr.astype(np.int32), y_colind=yr.indices.astype(np.int32),
                      y_val=yr.data.astype(np.float64), dev=dev_arrays,
                      genrou_u0=genrou_u0(ps, dp), clusters=clusters,
                      z0=dp.zvec.copy(), p=dp.pvec.copy())


def gen_speed_indices(desc: SystemDesc):
    """z indices of the GENROU speed deviations w (src/utils.jl:46-59)."""
    g = desc.dev[GP_GENROU]
    return (g[:, 0] + 4).astype(np.int64) if g.size else np.zeros(0, np.int64)


def make_events(events_per_scenario, dt):
    """events_per_scenario: list (len M) of lists of ContingencyEvent-like
    objects with bus, rfault, ton, toff.  Steps are round(t/dt) with ties to
    even, like Julia's round in src/dynamics.jl:528-542."""
    out = []
    for m, evs in enumerate(events_per_scenario):
        for e in evs:
            on = int(round(e.ton / dt))
            off = int(round(e.toff / dt))
            out.append((m, e.bus, float(e.rfault), on, off))
    arr = (CEvent * max(1, len(out)))()
    for i, (m, b, r, on, off) in enumerate(out):
        arr[i].scenario, arr[i].bus, arr[i].r_fault = m, b, r
        arr[i].step_on, arr[i].step_off = on, off
    return arr, len(out), out
