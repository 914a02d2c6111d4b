"""Reference-API mirror of the batched GPU integrator.

    gbl = HipBatchedLayout(dp, ps, M)          # GpuBatchedLayout(dp, ps, M), ext:466
    tvec, trajs = integrate_hip(gbl, ps, tf)   # integrate_gpu_cudss!, ext:2924-2996
    tvec, trajs = integrate_hip_schur(gbl, ps, tf)        # integrate_gpu_schur_cudss!, ext:3008-3080
    tvec, trajs = integrate_hip_multi(gbl, ps, tf, fault_buses, rfaults, ton, toff)
                                               # integrate_gpu_shared_multi!, ext:2640-2708

Semantics follow the reference: every scenario starts from dp.zvec; the
events registered on ps.dynamic.events apply to every scenario; steps are
round(tf/dt); after the Newton solve of step k the state is stored, then the
events of step k toggle and a dt = 0 re-solve follows (src/dynamics.jl:599-645);
the returned trajs[m] is n_sys x (nsteps+1) in the cluster-reordered z order
and all events are deactivated at the end.  Indices in the fault lists are
0-based bus indices.
"""
from __future__ import annotations

import numpy as np

from .desc import build_desc
from .lib import (HipContext, METHOD_MONO_FP64, METHOD_SCHUR_FP64, METHOD_SCHUR_FP32, METHOD_SCHUR_SHARED,
                  REFACTOR_ASYNC, REFACTOR_EVERY_ITER)


class HipBatchedLayout:
    """Device layout of M scenarios of one system (all start from dp.zvec)."""

    def __init__(self, dp, ps, M, record_idx=None, refactor_policy=REFACTOR_ASYNC):
        self.desc = build_desc(ps, dp)
        self.M = int(M)
        self.ctx = HipContext(self.desc, self.M, record_idx=record_idx,
                              refactor_policy=refactor_policy)
        self.ctx.set_state(dp.zvec, dp.pvec)
        self.zvec0 = np.array(dp.zvec, copy=True)
        self.pvec = np.array(dp.pvec, copy=True)
        self.last_stats = None

    def reset(self):
        self.ctx.set_state(self.zvec0, self.pvec)


def _step(t, dt):
    return int(round(t / dt))


def _run(gbl, method, events, nsteps, dt, newton_tol, itermax, raise_on_fail, ps=None):
    gbl.ctx.set_events(events, nsteps)
    if ps is not None:
        # line trips and load disconnects of ps.dynamic, in every scenario (the
        # reference runs them in its CPU integrate!, src/dynamics.jl:533-540)
        trips = [(m, t.from_bus, t.to_bus) + t.steps(dt) + (t.stamp(),)
                 for m in range(gbl.M) for t in ps.dynamic.trip_events]
        if trips:
            gbl.ctx.set_line_trips(trips)
        drops = [(m,) + d.target(ps) + (_step(d.ton, dt),) for m in range(gbl.M) for d in ps.dynamic.disconnect_events]
        if drops:
            gbl.ctx.set_device_drops(drops)
    gbl.last_stats = gbl.ctx.integrate(method, dt, newton_tol, itermax, raise_on_fail=raise_on_fail)
    trajs = [gbl.ctx.trajectory(m) for m in range(gbl.M)]
    return trajs


def _shared_events(ps, M, dt):
    ev = []
    for m in range(M):
        for e in ps.dynamic.events:
            ev.append((m, e.bus, e.rfault, _step(e.ton, dt), _step(e.toff, dt)))
    return ev


def integrate_hip(gbl, ps, tf, dt=1.0 / 120.0, newton_tol=1e-10, itermax=30, raise_on_fail=True):
    """Plain FP64 monolithic Newton: sparse LU of J refactorised every iteration."""
    nsteps = _step(tf, dt)
    tvec = np.arange(nsteps + 1) * dt
    trajs = _run(gbl, METHOD_MONO_FP64, _shared_events(ps, gbl.M, dt), [nsteps] * gbl.M, dt,
                 newton_tol, itermax, raise_on_fail, ps)
    return tvec, trajs


def integrate_hip_schur(gbl, ps, tf, dt=1.0 / 120.0, newton_tol=1e-10, itermax=30, fp32=False,
                        raise_on_fail=True):
    """Cluster Schur reduction; fp32=True factors the blocks and the reduced
    network in FP32 with FP64 residuals and updates (inexact Newton)."""
    nsteps = _step(tf, dt)
    tvec = np.arange(nsteps + 1) * dt
    trajs = _run(gbl, METHOD_SCHUR_FP32 if fp32 else METHOD_SCHUR_FP64, _shared_events(ps, gbl.M, dt),
                 [nsteps] * gbl.M, dt, newton_tol, itermax, raise_on_fail, ps)
    return tvec, trajs


def integrate_hip_multi(gbl, ps, tf, fault_buses, rfaults, ton, toff, dt=1.0 / 120.0,
                        newton_tol=1e-10, itermax=30, method=METHOD_SCHUR_FP64, raise_on_fail=True):
    """One fault per scenario (bus, r_fault, ton, toff); ton/toff/tf may be
    scalars or per-scenario arrays (uneven horizons are allowed)."""
    M = gbl.M
    tfv = np.broadcast_to(np.asarray(tf, float), (M,))
    tonv = np.broadcast_to(np.asarray(ton, float), (M,))
    toffv = np.broadcast_to(np.asarray(toff, float), (M,))
    ev = [(m, int(fault_buses[m]), float(rfaults[m]), _step(tonv[m], dt), _step(toffv[m], dt))
          for m in range(M)]
    nsteps = [_step(t, dt) for t in tfv]
    trajs = _run(gbl, method, ev, nsteps, dt, newton_tol, itermax, raise_on_fail, ps)
    tvecs = [np.arange(ns + 1) * dt for ns in nsteps]
    return tvecs, trajs


def integrate_hip_shared(gbl, ps, tf, dt=1.0 / 120.0, newton_tol=1e-10, itermax=30, raise_on_fail=True):
    """Shared-factor Schur Newton with Woodbury fault corrections
    (integrate_gpu_shared!, ext:2296-2375): one FP64 factorisation of the
    reduced network at the initial state serves every scenario and
    iteration; the events of ps.dynamic in every scenario."""
    nsteps = _step(tf, dt)
    tvec = np.arange(nsteps + 1) * dt
    trajs = _run(gbl, METHOD_SCHUR_SHARED, _shared_events(ps, gbl.M, dt), [nsteps] * gbl.M, dt,
                 newton_tol, itermax, raise_on_fail, ps)
    return tvec, trajs


def integrate_hip_shared_multi(gbl, ps, tf, fault_buses, rfaults, ton, toff, dt=1.0 / 120.0,
                               newton_tol=1e-10, itermax=30, raise_on_fail=True):
    """integrate_gpu_shared_multi! (ext:2640-2708): one fault per scenario,
    the shared factorisation with a rank-two Woodbury correction each."""
    return integrate_hip_multi(gbl, ps, tf, fault_buses, rfaults, ton, toff, dt=dt, newton_tol=newton_tol,
                               itermax=itermax, method=METHOD_SCHUR_SHARED, raise_on_fail=raise_on_fail)
