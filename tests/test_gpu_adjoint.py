"""GPU discrete adjoint (gp_adjoint through the C ABI) against the oracle's
adjoint and against central finite differences of the GPU's own forward
sweep.

Bars, all FP64:
  - per-scenario objective, gradient and dJ/dz0 within 1e-7 of the oracle's
    adjoint (relative to the largest entry; the oracle takes its parameter
    partials by central differences of the residual, the library by its
    hand-written pullbacks, ~1e-9 apart);
  - the 2-bus event-offset test of test/test_adjoint_event_offset.jl:26-84
    within 1e-6 of the centred difference;
  - 20 random H / D / K of the 200-bus, 128-scenario, 3 s workload within
    1e-5 of centred differences with h = max(1e-6 |p|, 1e-5): a relative
    step, floored where the quotient meets the FP64 noise of the objective
    (a damping D of 0.2 at h = 2e-7 is off by up to 3e-5 while the adjoint
    value does not move; a gain K = 50 at h = 5e-4 crosses kinks of the
    constant-power load clamp and is off by 3e-5 the other way);
  - checkpoint segments of any length give bit-identical gradients.
"""
import numpy as np
import pytest

from conftest import get_case

pytestmark = pytest.mark.gpu

MONO = 0


def _ctx(case, M, **kw):
    from gradpower_hip.lib import HipContext
    c = HipContext(case.desc, M, **kw)
    c.set_state(case.desc.z0, case.desc.p)
    return c


def _oracle(case):
    from oracle.oracle import Oracle
    return Oracle(case.desc)


def _events(case, m):
    return [(e.bus, e.rfault, int(round(e.ton / case.dt)), int(round(e.toff / case.dt))) for e in case.events[m]]


def _all_events(case, M):
    return [(m,) + e for m in range(M) for e in _events(case, m)]


def _speed_rows(case):
    return [int(r[0] + 4) for r in case.desc.dev[0]]


def _gains(case):
    """Every parameter with a pullback: GENROU H, D; SEXS K; ESDC1A KA;
    TGOV1 R; IEESGO K1; IEEEST KS."""
    d = case.desc.dev
    out = [int(r[2] + 6) for r in d[0]] + [int(r[2] + 7) for r in d[0]]
    out += [int(r[1] + 2) for r in d[3]] + [int(r[1]) for r in d[4]]
    out += [int(r[2]) for r in d[2]] + [int(r[2] + 6) for r in d[1]] + [int(r[2] + 12) for r in d[5]]
    return out


def _objective_fd(case, M, nsteps, ev, p, rows, scale, tol, scen=None):
    """scale * sum_m sum_{k>=1} dt sum_r z_k[r]^2 from gp_integrate's
    trajectories, over the scenarios in `scen` (all by default), each of
    which must have converged."""
    ctx = _ctx(case, M, record_idx=rows)
    ctx.set_state(case.desc.z0, p)
    ctx.set_events(ev, nsteps)
    ctx.integrate(MONO, case.dt, tol=tol, itermax=50, raise_on_fail=scen is None)
    scen = range(M) if scen is None else scen
    st, _ = ctx.scenario_status()
    assert all(st[m] == 0 for m in scen)
    tot = 0.0
    for m in scen:
        tr = ctx.trajectory(m)
        tot += scale * case.dt * float((tr[:, 1:] ** 2).sum())
    ctx.close()
    return tot


def _check_vs_oracle(case, M, nsteps, params, scale, segment=30, method=MONO):
    ctx = _ctx(case, M)
    ctx.set_events(_all_events(case, M), nsteps)
    r = ctx.adjoint(method, case.dt, params, tol=1e-12, segment=segment, scale=scale, want_dz0=True)
    o = _oracle(case)
    for m in range(M):
        val, g, dz0 = o.adjoint(case.dt, int(nsteps[m]), params, _events(case, m), tol=1e-12, scale=scale)
        assert r["value_scen"][m] == pytest.approx(val, rel=1e-9), m
        gs = r["grad_scen"][m]
        assert np.max(np.abs(gs - g)) <= 1e-7 * np.max(np.abs(g)), (m, gs, g)
        assert np.max(np.abs(r["dz0"][m] - dz0)) <= 1e-7 * np.max(np.abs(dz0)), m
    assert np.allclose(r["grad"], r["grad_scen"].sum(axis=0), rtol=1e-14, atol=0)
    assert r["stats"]["n_excluded"] == 0
    return r


def test_adjoint_event_offset_2bus():
    """test/test_adjoint_event_offset.jl:26-84 on the library: dJ/dH within
    1e-6 of the centred difference with h = 1e-5 max(|H|, 1)."""
    case = get_case("case_2bus_testdata", fault=False, tf=4.8)
    dt = 1.0 / 120.0
    nsteps = np.array([576], np.int32)
    ev = [(0, 1, 0.02, 24, 36)]
    par = int(case.desc.dev[0][0][2] + 6)
    ctx = _ctx(case, 1)
    ctx.set_events(ev, nsteps)
    r = ctx.adjoint(MONO, dt, [par], tol=1e-12)
    p0 = case.desc.p.copy()
    h = 1e-5 * max(abs(p0[par]), 1.0)
    pp, pm = p0.copy(), p0.copy()
    pp[par] += h
    pm[par] -= h
    rows = _speed_rows(case)
    fd = (_objective_fd(case, 1, nsteps, ev, pp, rows, 1.0, 1e-12) -
          _objective_fd(case, 1, nsteps, ev, pm, rows, 1.0, 1e-12)) / (2 * h)
    assert abs(r["grad"][0] - fd) / abs(fd) < 1e-6
    assert r["value"] == pytest.approx(_objective_fd(case, 1, nsteps, ev, p0, rows, 1.0, 1e-12), rel=1e-12)


@pytest.mark.parametrize("method", [0, 1, 2])
def test_adjoint_matches_oracle_9bus(method):
    """Any forward method: the Schur ones (FP64 and FP32 factors) converge to
    the same trajectory, so the gradient is the same to the bar."""
    case = get_case("case1_9bus")
    M = 8
    nsteps = case.nsteps[:M]
    d = case.desc.dev
    params = [int(r[2] + 6) for r in d[0]] + [int(r[2] + 7) for r in d[0]] + \
             [int(r[1] + 2) for r in d[3]] + [int(r[2]) for r in d[2]]
    _check_vs_oracle(case, M, nsteps, params, 1.0 / 3.0, method=method)


@pytest.mark.parametrize("method", [0, 2])
def test_adjoint_all_gains_uneven_horizons_match_oracle(method):
    """All seven parameter kinds on a 60-bus grid with every controller type,
    four scenarios with horizons of 152 to 408 steps, segments of 30."""
    case = get_case("case2_grid", M=4, nbus=60, n_genrou=12, n_pss=4, n_static=6, seed=77)
    _check_vs_oracle(case, 4, case.nsteps, _gains(case), 1.0 / 12.0, method=method)


@pytest.mark.parametrize("method", [0, 2])
def test_adjoint_segments_bit_identical(method):
    """Segments of 7, 30 and 1000 steps give identical gradients: every
    recomputed segment repeats the forward sweep's iterates bit for bit, for
    the monolithic method and for the Schur method with FP32 factors (the
    bench's forward sweep)."""
    case = get_case("case1_9bus")
    M = 8
    ctx = _ctx(case, M)
    ctx.set_events(_all_events(case, M), case.nsteps[:M])
    params = _gains(case)
    res = [ctx.adjoint(method, case.dt, params, tol=1e-12, segment=s, want_dz0=True) for s in (7, 30, 1000)]
    for r in res[1:]:
        assert np.array_equal(r["grad_scen"], res[0]["grad_scen"])
        assert np.array_equal(r["value_scen"], res[0]["value_scen"])
        assert np.array_equal(r["dz0"], res[0]["dz0"])
    assert res[0]["stats"]["n_segments"] == 18 and res[2]["stats"]["n_segments"] == 1
    # the forward sweep's trajectory stays available
    o = _oracle(case)
    tr, _, _ = o.integrate(case.dt, int(case.nsteps[0]), _events(case, 0), tol=1e-12)
    assert np.max(np.abs(ctx.trajectory(0) - tr.T)) < 1e-9


def test_adjoint_unsupported_parameter():
    from gradpower_hip.lib import GpError, GP_ERR_UNSUPPORTED
    case = get_case("case1_9bus")
    ctx = _ctx(case, 2)
    ctx.set_events([], case.nsteps[:2])
    zip_par = int(case.desc.dev[6][0][1])
    with pytest.raises(GpError) as e:
        ctx.adjoint(MONO, case.dt, [zip_par])
    assert e.value.code == GP_ERR_UNSUPPORTED


def test_adjoint_200bus_vs_central_differences():
    """The 200-bus workload: 40 GENROU+SEXS+TGOV1, 128 fault scenarios, 3 s at
    dt = 1/120, objective sum_m sum_k dt mean_g w_g^2, 30-step segments;
    20 random parameters among all H, D and SEXS K against centred
    differences with h = max(1e-6 |p|, 1e-5), relative error <= 1e-5, for
    the adjoint whose forward sweep runs the monolithic method and for the
    bench's configuration (forward sweep and recomputation by the Schur
    method with FP32 factors).  Scenarios whose forward sweep fails (bolted
    faults on this weak grid) are left out of both sides."""
    case = get_case("case3_200bus")
    M = case.M
    nsteps = case.nsteps
    ev = _all_events(case, M)
    d = case.desc.dev
    allp = [int(r[2] + 6) for r in d[0]] + [int(r[2] + 7) for r in d[0]] + [int(r[1] + 2) for r in d[3]]
    ctx = _ctx(case, M)
    ctx.set_events(ev, nsteps)
    scale = 1.0 / len(d[0])
    tol = 1e-12
    r = ctx.adjoint(MONO, case.dt, allp, tol=tol, segment=30, scale=scale, raise_on_fail=False)
    st, _ = ctx.scenario_status()
    ok = [m for m in range(M) if st[m] == 0]
    ctx.set_state(case.desc.z0, case.desc.p)
    ctx.set_events(ev, nsteps)
    r32 = ctx.adjoint(2, case.dt, allp, tol=tol, segment=30, scale=scale, raise_on_fail=False)
    st32, _ = ctx.scenario_status()
    assert np.array_equal(st32 != 0, st != 0)
    print(f"{M - len(ok)} of {M} scenarios left out (forward Newton failure)", flush=True)
    assert r["stats"]["n_excluded"] == M - len(ok) and len(ok) >= M - 16
    assert r["stats"]["n_segments"] == 12
    assert np.all(r["grad_scen"][[m for m in range(M) if st[m] != 0]] == 0.0)
    ctx.close()
    rng = np.random.default_rng(2024)
    pick = rng.choice(len(allp), size=20, replace=False)
    rows = _speed_rows(case)
    p0 = case.desc.p.copy()
    worst = 0.0
    for j in pick:
        q = allp[j]
        h = max(1e-6 * abs(p0[q]), 1e-5)
        pp, pm = p0.copy(), p0.copy()
        pp[q] += h
        pm[q] -= h
        fd = (_objective_fd(case, M, nsteps, ev, pp, rows, scale, tol, ok) -
              _objective_fd(case, M, nsteps, ev, pm, rows, scale, tol, ok)) / (2 * h)
        rel = max(abs(r["grad"][j] - fd), abs(r32["grad"][j] - fd)) / abs(fd)
        print(f"param {q}: adjoint {r['grad'][j]:.12e} (FP32-factor forward {r32['grad'][j]:.12e}) "
              f"fd {fd:.12e} rel {rel:.2e}", flush=True)
        worst = max(worst, rel)
    assert worst <= 1e-5


def test_adjoint_ragged_horizons_and_event_tables():
    """The adjoint with ragged horizons (0 and 1 step included) and awkward
    event tables (on and off in one step, overlapping faults, a fault past
    the horizon or never cleared), short segments, against the oracle."""
    from test_gpu_parity import _edge_events
    case = get_case("case1_9bus")
    nsteps, ev = _edge_events(case)
    M = len(nsteps)
    ctx = _ctx(case, M)
    ctx.set_events(ev, nsteps)
    params = _gains(case)
    r = ctx.adjoint(MONO, case.dt, params, tol=1e-12, segment=7, scale=0.5, want_dz0=True)
    o = _oracle(case)
    for m in range(M):
        val, g, dz0 = o.adjoint(case.dt, nsteps[m], params, [e[1:] for e in ev if e[0] == m], tol=1e-12, scale=0.5)
        if nsteps[m] == 0:
            assert r["value_scen"][m] == 0.0 and np.all(r["grad_scen"][m] == 0.0) and np.all(r["dz0"][m] == 0.0)
            continue
        assert r["value_scen"][m] == pytest.approx(val, rel=1e-9, abs=1e-18), m
        assert np.max(np.abs(r["grad_scen"][m] - g)) <= 1e-7 * np.max(np.abs(g)) + 1e-15, m
        assert np.max(np.abs(r["dz0"][m] - dz0)) <= 1e-7 * np.max(np.abs(dz0)) + 1e-15, m


def test_adjoint_with_global_memory_solves_matches_oracle():
    """The discrete adjoint with the triangular solves (forward J, backward
    J^T) keeping x in global memory: within 1e-7 of the oracle's adjoint."""
    case = get_case("case1_9bus")
    M = 4
    from gradpower_hip.lib import HipContext
    ctx = HipContext(case.desc, M, global_x=True)
    ctx.set_state(case.desc.z0, case.desc.p)
    ctx.set_events(_all_events(case, M), case.nsteps[:M])
    params = _gains(case)
    r = ctx.adjoint(MONO, case.dt, params, tol=1e-12, segment=30)
    o = _oracle(case)
    for m in range(M):
        _, g, _ = o.adjoint(case.dt, int(case.nsteps[m]), params, _events(case, m), tol=1e-12)
        assert np.max(np.abs(r["grad_scen"][m] - g)) <= 1e-7 * np.max(np.abs(g)), m
    ctx.close()
