"""GPU parity on the benchmark workload itself: the 2,000-bus grid with 256
uneven fault scenarios (synth.case2_grid(), n = 11,872 states, every device
type: GENROU, SEXS, ESDC1A, TGOV1, IEESGO, IEEEST, ZIP loads, PV static
generator stubs).

1. Trajectories of 16 scenarios over their first 60 steps (four of them fail
   there in the oracle: bolted faults whose faulted algebraic system has no
   Newton-reachable solution) against the oracle, both sides iterating to
   |F|inf < 1e-12 with at most 30 corrections: within 1e-9 pu per step up
   to the step before the first Newton solve that does not converge, and
   that step itself the same.
2. The per-scenario outcome of the whole benchmark run (full horizons,
   |F|inf < 1e-10, at most 30 corrections, the reference's Newton rule:
   stall_window = 0) against tests/golden/c2_oracle_failures.json (written
   by tests/golden/make_c2_oracle_failures.py from the oracle).  Exact Newton
   (a factorisation at every iteration, as the reference's newton_step! and
   its GPU Schur path, ext:2721-2813) stops the same scenarios after the
   same steps.  The policies that reuse factorisations (ADAPTIVE, ASYNC)
   run a failed solve again as exact Newton before they stop a scenario, so
   they never stop a scenario the oracle integrates, nor earlier than the
   oracle does.
"""
import json
import os

import numpy as np
import pytest

from conftest import get_case

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c2_oracle_failures.json")
METHODS = {"mono": 0, "schur64": 1, "schur32": 2}
TRAJ_SCEN = [0, 1, 2, 4, 5, 6, 7, 8, 9, 11, 12, 27, 47, 54, 184, 218]
CAP = 60


def _events(case, m):
    return [(e.bus, e.rfault, int(round(e.ton / case.dt)), int(round(e.toff / case.dt))) for e in case.events[m]]


@pytest.mark.parametrize("method,policy", [("mono", 0), ("schur64", 0), ("schur64", 3), ("schur32", 3)])
def test_c2_trajectories_match_oracle(method, policy):
    from gradpower_hip.lib import HipContext
    from oracle.oracle import Oracle
    case = get_case("case2_grid")
    scen = TRAJ_SCEN
    M = len(scen)
    tol, itermax = 1e-12, 30
    ns = [min(CAP, int(case.nsteps[m])) for m in scen]
    ctx = HipContext(case.desc, M, refactor_policy=policy, stall_window=0)
    ctx.set_state(case.desc.z0, case.desc.p)
    ctx.set_events([(j,) + e for j, m in enumerate(scen) for e in _events(case, m)], ns)
    ctx.integrate(METHODS[method], case.dt, tol, itermax, raise_on_fail=False)
    status, done = ctx.scenario_status()
    o = Oracle(case.desc)
    n_failed = 0
    for j, m in enumerate(scen):
        ref, _, st = o.integrate(case.dt, ns[j], _events(case, m), tol=tol, itermax=itermax)
        ok_steps = ns[j] if st["first_fail_step"] < 0 else st["first_fail_step"]
        n_failed += st["first_fail_step"] >= 0
        if policy == 0 and method != "schur32":
            assert done[j] == ok_steps, (method, m, done[j], ok_steps)
            assert (status[j] == 0) == (st["first_fail_step"] < 0), (method, m)
        else:
            assert done[j] >= ok_steps, (method, m, done[j], ok_steps)
        k = min(done[j], ok_steps)
        dev = np.max(np.abs(ctx.trajectory(m=j)[:, :k + 1].T - ref[:k + 1]))
        assert dev < 1e-9, (method, m, dev)
    assert n_failed >= 4   # the sample keeps scenarios that stop early
    ctx.close()


@pytest.mark.parametrize("method,policy", [("mono", 0), ("schur64", 0), ("schur32", 0), ("schur64", 3),
                                           ("schur32", 3)])
def test_c2_failed_scenarios_match_oracle(method, policy):
    """Exact Newton in FP64 (MONO, SCHUR64 with a factorisation at every
    iteration) stops exactly the oracle's scenarios after the same steps.
    FP32 factors, or factors kept across iterations, are an inexact Newton
    method: a solve that fails is run again as exact Newton with FP64
    factors, so these stop only scenarios the oracle stops, never earlier
    (where the inexact iteration converges and exact Newton does not, they
    go further)."""
    from gradpower_hip.desc import gen_speed_indices
    from gradpower_hip.lib import HipContext
    with open(FIX) as fh:
        fx = json.load(fh)
    case = get_case("case2_grid")
    assert [int(x) for x in case.nsteps] == fx["nsteps"]
    ctx = HipContext(case.desc, case.M, record_idx=gen_speed_indices(case.desc).astype(np.int32),
                     refactor_policy=policy, stall_window=0)
    ctx.set_state(case.desc.z0, case.desc.p)
    ctx.set_events([(m,) + e for m in range(case.M) for e in _events(case, m)], case.nsteps)
    st = ctx.integrate(METHODS[method], case.dt, fx["newton_tol"], fx["itermax"], raise_on_fail=False)
    status, done = ctx.scenario_status()
    ctx.close()
    gpu_failed = {int(m) for m in np.flatnonzero(status != 0)}
    ref_failed = set(fx["failed"])
    ref_done = {m: (fx["first_fail_step"][str(m)] if m in ref_failed else int(case.nsteps[m]))
                for m in range(case.M)}
    if policy == 0 and method != "schur32":
        assert gpu_failed == ref_failed, (method, sorted(gpu_failed ^ ref_failed))
        bad = {m: (int(done[m]), ref_done[m]) for m in range(case.M) if done[m] != ref_done[m]}
        assert not bad, (method, bad)
    else:
        assert gpu_failed <= ref_failed, (method, sorted(gpu_failed - ref_failed))
        early = {m: (int(done[m]), ref_done[m]) for m in range(case.M) if done[m] < ref_done[m]}
        assert not early, (method, early)
    assert st["scenario_steps"] == int(np.sum(done))
