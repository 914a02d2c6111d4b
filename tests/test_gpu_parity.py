"""GPU parity: the HIP library (through its C ABI) against the CPU oracle.

Tolerances: residuals and Jacobian values are the same FP64 expressions, so
they agree to rounding (<= 1e-12 relative); trajectories compare converged
Newton solutions (both sides iterate to |F|inf < 1e-12), so the per-step
state deviation bound is 1e-9 pu.
"""
import numpy as np
import pytest

from conftest import get_case

pytestmark = pytest.mark.gpu


def _ctx(case, M=None, **kw):
    from gradpower_hip.lib import HipContext
    c = HipContext(case.desc, M or case.M, **kw)
    c.set_state(case.desc.z0, case.desc.p)
    return c


def _oracle(case):
    from oracle.oracle import Oracle
    return Oracle(case.desc)


def _perturbed(case, M, seed, scale=0.02):
    rng = np.random.default_rng(seed)
    z0 = case.desc.z0
    return z0[None, :] * (1 + scale * rng.standard_normal((M, len(z0)))) + scale * rng.standard_normal((M, len(z0)))


@pytest.mark.parametrize("name", ["case_2bus_testdata", "case0_2bus", "case1_9bus", "case3_200bus", "case2_60bus"])
def test_residual_matches_oracle(name):
    case = get_case(name)
    M = min(case.M, 8)
    ctx = _ctx(case, M)
    o = _oracle(case)
    z = _perturbed(case, M, 1)
    zold = _perturbed(case, M, 2, 0.01)
    ev = [(m, case.events[m][0].bus, case.events[m][0].rfault, 1, 2) for m in range(M) if case.events[m]]
    ctx.set_events(ev, case.nsteps[:M])
    fault_on = np.array([1 if m % 2 else 0 for m in range(M)], np.uint8)
    for dt in (case.dt, 0.0):
        f = ctx.eval_residual(z, zold, dt, fault_on)
        for m in range(M):
            g = np.zeros(case.desc.nbus)
            if fault_on[m] and case.events[m]:
                g[case.events[m][0].bus] += 1.0 / case.events[m][0].rfault
            fo = o.residual(z[m], zold[m], dt, g)
            scale = 1.0 + np.abs(fo)
            assert np.max(np.abs(f[m] - fo) / scale) < 1e-12, (name, m, dt)


@pytest.mark.parametrize("name", ["case0_2bus", "case1_9bus", "case3_200bus", "case2_60bus"])
def test_jacobian_matches_oracle(name):
    import scipy.sparse as sp
    case = get_case(name)
    M = min(case.M, 4)
    ctx = _ctx(case, M)
    o = _oracle(case)
    z = _perturbed(case, M, 3)
    ev = [(m, case.events[m][0].bus, case.events[m][0].rfault, 1, 2) for m in range(M) if case.events[m]]
    ctx.set_events(ev, case.nsteps[:M])
    fault_on = np.ones(M, np.uint8)
    rp, ci, vals = ctx.eval_jacobian(z, case.dt, fault_on)
    n = case.desc.n
    for m in range(M):
        g = np.zeros(case.desc.nbus)
        if case.events[m]:
            g[case.events[m][0].bus] += 1.0 / case.events[m][0].rfault
        Jo = o.jacobian_csc(z[m], case.dt, g).toarray()
        Jg = sp.csr_matrix((vals[m], ci, rp), shape=(n, n)).toarray()
        assert np.max(np.abs(Jg - Jo) / (1 + np.abs(Jo))) < 1e-12, (name, m)


def _oracle_trajs(case, scen, tol, nsteps=None):
    o = _oracle(case)
    out = {}
    for m in scen:
        ev = [(e.bus, e.rfault, int(round(e.ton / case.dt)), int(round(e.toff / case.dt))) for e in case.events[m]]
        ns = int(case.nsteps[m]) if nsteps is None else nsteps
        tr, zf, st = o.integrate(case.dt, ns, ev, tol=tol, itermax=50)
        assert st["not_converged"] == 0
        out[m] = tr
    return out


METHODS = {"mono": 0, "schur64": 1, "schur32": 2}


@pytest.mark.parametrize("method", ["mono", "schur64", "schur32"])
@pytest.mark.parametrize("name", ["case0_2bus", "case1_9bus", "case3_200bus", "case2_60bus"])
def test_newton_step_matches_oracle(name, method):
    """One Newton correction of every method equals the oracle's -J^{-1}F
    (the Schur elimination is exact algebra; FP32 factors give the step to
    ~1e-6 relative)."""
    case = get_case(name)
    M = min(case.M, 4)
    ctx = _ctx(case, M)
    o = _oracle(case)
    z = _perturbed(case, M, 5, 0.005)
    zold = _perturbed(case, M, 6, 0.002)
    ev = [(m, case.events[m][0].bus, case.events[m][0].rfault, 1, 2) for m in range(M) if case.events[m]]
    ctx.set_events(ev, case.nsteps[:M])
    fault_on = np.ones(M, np.uint8)
    for dt in (case.dt, 0.0):
        dz = ctx.newton_step(METHODS[method], z, zold, dt, fault_on)
        for m in range(M):
            g = np.zeros(case.desc.nbus)
            if case.events[m]:
                g[case.events[m][0].bus] += 1.0 / case.events[m][0].rfault
            f = o.residual(z[m], zold[m], dt, g)
            ref = -o.linsolve(o.jacobian(z[m], dt, g), f)
            tol = 1e-4 if method == "schur32" else 1e-9
            assert np.max(np.abs(dz[m] - ref)) <= tol * (1 + np.max(np.abs(ref))), (name, method, m, dt)


@pytest.mark.parametrize("method,policy", [("mono", 0), ("schur64", 0), ("schur64", 2), ("schur32", 2),
                                           ("schur64", 3), ("schur32", 3)])
@pytest.mark.parametrize("name,nsteps", [("case_2bus_testdata", 120), ("case0_2bus", 240), ("case1_9bus", None),
                                         ("case3_200bus", 60), ("case2_60bus", 120)])
def test_trajectory_matches_oracle(name, nsteps, method, policy):
    case = get_case(name)
    M = min(case.M, 8)
    tol = 1e-12
    ctx = _ctx(case, M, refactor_policy=policy)
    ns = [int(case.nsteps[m]) if nsteps is None else nsteps for m in range(M)]
    ev = []
    for m in range(M):
        for e in case.events[m]:
            ev.append((m, e.bus, e.rfault, int(round(e.ton / case.dt)), int(round(e.toff / case.dt))))
    ctx.set_events(ev, ns)
    st = ctx.integrate(METHODS[method], case.dt, tol, 50)
    ref = _oracle_trajs(case, range(M), tol, nsteps)
    for m in range(M):
        tr = ctx.trajectory(m)
        dev = np.max(np.abs(tr.T - ref[m]))
        assert dev < 1e-9, (name, method, m, dev, st)


@pytest.mark.parametrize("method,policy", [("mono", 0), ("schur64", 0), ("schur64", 1), ("schur32", 1)])
@pytest.mark.parametrize("name", ["case1_9bus", "case2_grid_small"])
def test_scenarios_advancing_alone_match_lock_step(name, method, policy):
    """Scenarios advancing on their own (the default time loop) give the same
    iterates as the lock-step loop, bit for bit, when every Newton solve
    starts with its own factorisation (EVERY_ITER, FIRST_ITER): each
    scenario's arithmetic does not depend on the others.  Uneven horizons and
    per-scenario fault times included."""
    case = get_case("case2_60bus" if name == "case2_grid_small" else name)
    M = min(case.M, 8)
    ns = [int(case.nsteps[m]) for m in range(M)]
    ev = [(m, e.bus, e.rfault, int(round(e.ton / case.dt)), int(round(e.toff / case.dt)))
          for m in range(M) for e in case.events[m]]
    out = []
    for lock in (False, True):
        ctx = _ctx(case, M, refactor_policy=policy, lock_step=lock, dense_block=-1)   # same plan in both
        ctx.set_events(ev, ns)
        st = ctx.integrate(METHODS[method], case.dt, 1e-12, 50, raise_on_fail=False)
        status, steps = ctx.scenario_status()
        out.append(([ctx.trajectory(m) for m in range(M)], status.copy(), steps.copy(), st))
        ctx.close()
    (ta, sa, ka, st_free), (tb, sb, kb, st_lock) = out
    assert np.array_equal(sa, sb) and np.array_equal(ka, kb)
    for m in range(M):
        if sa[m] == 0:
            assert np.array_equal(ta[m], tb[m]), (name, method, m)
    assert st_free["newton_iters"] == st_lock["newton_iters"]
    assert st_free["scenario_steps"] == st_lock["scenario_steps"]


@pytest.mark.parametrize("method,policy", [("mono", 0), ("schur32", 3)])
def test_gensal_trajectory_matches_oracle(method, policy):
    """GENSAL enters as a GENROU with x_qp = x_dp and T_q0p = T_d0p, as the
    reference's parser builds it (src/generators.jl:41-66): the 9-bus case
    with its bus-4 machine a GENSAL, trajectories within 1e-9 of the oracle."""
    case = get_case("case1_9bus", gensal=True)
    M = 8
    ctx = _ctx(case, M, refactor_policy=policy)
    ev = [(m, e.bus, e.rfault, int(round(e.ton / case.dt)), int(round(e.toff / case.dt)))
          for m in range(M) for e in case.events[m]]
    ctx.set_events(ev, case.nsteps[:M])
    st = ctx.integrate(METHODS[method], case.dt, 1e-12, 50)
    ref = _oracle_trajs(case, range(M), 1e-12)
    for m in range(M):
        dev = np.max(np.abs(ctx.trajectory(m).T - ref[m]))
        assert dev < 1e-9, (method, m, dev, st)


def _edge_events(case):
    """Ragged horizons (0, 1, 37, 120, 60 steps) and awkward fault tables:
    on and off in the same step, two overlapping faults, a fault past the
    horizon, a fault that never clears."""
    b = case.desc.nbus
    nsteps = [0, 1, 37, 120, 60]
    ev = [(0, 1, 0.05, 1, 2),
          (1, 2, 0.05, 1, 1),
          (2, 3, 0.08, 3, 20), (2, 5, 0.1, 10, 30),
          (3, 4, 0.05, 200, 210), (3, 6, 0.07, 2, 6),
          (4, b - 1, 0.05, 5, 500)]
    return nsteps, ev


@pytest.mark.parametrize("method,policy,lock", [("mono", 0, False), ("mono", 0, True), ("schur64", 2, False),
                                                ("schur32", 3, False), ("schur32", 3, True)])
def test_ragged_horizons_and_event_tables_match_oracle(method, policy, lock):
    """Five scenarios (M not a multiple of 64) with ragged horizons, including
    0 and 1 step, and the awkward event tables above, against the oracle."""
    case = get_case("case1_9bus")
    nsteps, ev = _edge_events(case)
    M = len(nsteps)
    ctx = _ctx(case, M, refactor_policy=policy, lock_step=lock)
    ctx.set_events(ev, nsteps)
    st = ctx.integrate(METHODS[method], case.dt, 1e-12, 50)
    assert st["scenario_steps"] == sum(nsteps)
    o = _oracle(case)
    for m in range(M):
        tr = ctx.trajectory(m)
        assert tr.shape == (case.desc.n, nsteps[m] + 1)
        ref, _, rs = o.integrate(case.dt, nsteps[m], [e[1:] for e in ev if e[0] == m], tol=1e-12, itermax=50)
        assert rs["not_converged"] == 0
        dev = np.max(np.abs(tr.T - ref))
        assert dev < 1e-9, (method, lock, m, dev)


@pytest.mark.parametrize("method,policy", [("mono", 0), ("schur64", 0), ("schur32", 3)])
@pytest.mark.parametrize("name,nsteps", [("case3_200bus", 60), ("case2_60bus", 120)])
def test_global_memory_solve_matches_oracle(name, nsteps, method, policy):
    """The triangular solves that keep x in global memory (the form they take
    above about 12k rows, where x does not fit in LDS; forced here with
    gp_options.global_x) give trajectories within 1e-9 pu of the oracle."""
    case = get_case(name)
    M = min(case.M, 8)
    tol = 1e-12
    ctx = _ctx(case, M, refactor_policy=policy, global_x=True)
    ev = [(m, e.bus, e.rfault, int(round(e.ton / case.dt)), int(round(e.toff / case.dt)))
          for m in range(M) for e in case.events[m]]
    ctx.set_events(ev, [nsteps] * M)
    st = ctx.integrate(METHODS[method], case.dt, tol, 50)
    ref = _oracle_trajs(case, range(M), tol, nsteps)
    for m in range(M):
        dev = np.max(np.abs(ctx.trajectory(m).T - ref[m]))
        assert dev < 1e-9, (name, method, m, dev, st)
    ctx.close()
