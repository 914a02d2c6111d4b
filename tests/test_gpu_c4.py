"""The large-system path on the GPU against the oracle.

The 70,000-bus workload (synth.case4_grid: 35 tiles of the 2,000-bus grid) is
too large for the oracle to follow within a test, so its building block is
checked here at 4 tiles: 8,000 buses, 43k states, a reduced network of about
18k rows -- beyond what the triangular solves can keep in the LDS of one CU,
so they keep x in global memory by themselves, as at 70,000 buses (and the
monolithic J, 43k rows, too).  Four fault scenarios over 24 steps (the fault
goes on at step 2): trajectories of MONO, SCHUR64 and SCHUR32/ASYNC within
1e-9 pu of the oracle, both sides iterating to |F|inf < 1e-11."""
import numpy as np
import pytest

from conftest import get_case

pytestmark = pytest.mark.gpu
METHODS = {"mono": 0, "schur64": 1, "schur32": 2}
LDS = 160 * 1024


@pytest.fixture(scope="module")
def oracle_trajs():
    from oracle.oracle import Oracle
    case = get_case("case4_grid", tiles=(2, 2), M=4, tf=0.2)
    o = Oracle(case.desc)
    out = []
    for m in range(case.M):
        ev = [(e.bus, e.rfault, int(round(e.ton / case.dt)), int(round(e.toff / case.dt))) for e in case.events[m]]
        tr, _, st = o.integrate(case.dt, int(case.nsteps[m]), ev, tol=1e-11, itermax=50)
        assert st["not_converged"] == 0
        out.append(tr)
    return out


@pytest.mark.parametrize("method,policy", [("mono", 0), ("schur64", 0), ("schur32", 3)])
def test_large_system_trajectories_match_oracle(method, policy, oracle_trajs):
    from gradpower_hip.lib import HipContext
    case = get_case("case4_grid", tiles=(2, 2), M=4, tf=0.2)
    ctx = HipContext(case.desc, case.M, refactor_policy=policy)
    ctx.set_state(case.desc.z0, case.desc.p)
    ctx.set_events([(m, e.bus, e.rfault, int(round(e.ton / case.dt)), int(round(e.toff / case.dt)))
                    for m in range(case.M) for e in case.events[m]], case.nsteps)
    st = ctx.integrate(METHODS[method], case.dt, 1e-11, 50)
    info = ctx.symbolic_info()
    rows = info["n"] if method == "mono" else info["n_red"]
    assert 8 * rows + 64 * 1024 > LDS   # the solves keep x in global memory
    for m in range(case.M):
        dev = np.max(np.abs(ctx.trajectory(m).T - oracle_trajs[m]))
        assert dev < 1e-9, (method, m, dev, st)
    ctx.close()
