"""GPU parity of exciters taken offline (gp_set_device_drops;
DisconnectDeviceEvent, src/GradPower.jl:130-143,681-708) against the oracle.
An offline exciter's rows of F are zero (the reference's kernels skip it into
a zeroed F, ext/GradPowerCUDAExt/GradPowerCUDAExt.jl:1340,1072-1199), so
backward Euler holds its states; the oracle's restatement is pinned on the
CPU (tests/test_oracle_exciter_drops.py: the same as the system read without
that exciter when dropped at the operating point).

  - residual and Jacobian: an offline exciter's rows are z - zold and
    identity rows, bit for bit, the other rows those of the online system;
  - trajectories with faults, exciter drops (SEXS, ESDC1A) and ZIP-load drops
    mixed per scenario, every method, free-running and lock step: within
    1e-9 pu of the oracle;
  - the discrete adjoint with exciter drops (the dropped exciters' gains
    among the parameters): within 1e-7 relative of the oracle's;
  - devices with algebraic states are refused.
"""
import os

import numpy as np
import pytest

from conftest import get_case

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "psse")
METHODS = {"mono": 0, "schur64": 1, "schur32": 2}
TOL = 1e-12
SEXS, ESDC1A, ZIP = 3, 4, 6


def _case():
    return get_case("case_psse", raw=os.path.join(G, "ieee9_v33.raw"), dyr=os.path.join(G, "ieee9bus_SEXS.dyr"),
                    M=8)


def _ctx(case, M, **kw):
    from gradpower_hip.lib import HipContext
    c = HipContext(case.desc, M, **kw)
    c.set_state(case.desc.z0, case.desc.p)
    return c


def _tables(case, M):
    """Per scenario: its fault, SEXS row m % 3 offline at step 8 + 7m (none
    for m % 4 == 3), and in scenario 1 also ZIP load 2 at step 30."""
    ev, dr, per = [], [], []
    for m in range(M):
        e = case.events[m][0]
        f = [(e.bus, e.rfault, int(round(e.ton / case.dt)), int(round(e.toff / case.dt)))]
        ev += [(m,) + f[0]]
        d = []
        if m % 4 != 3:
            dr.append((m, SEXS, m % 3, 8 + 7 * m))
            d.append(("sexs", m % 3, 8 + 7 * m))
        if m == 1:
            dr.append((m, ZIP, 2, 30))
            d.append(("zip", 2, 30))
        per.append((f, d))
    return ev, dr, per


def test_offline_exciter_rows_of_residual_and_jacobian():
    import scipy.sparse as sp
    case = _case()
    M = 2
    ctx = _ctx(case, M)
    ctx.set_events([], case.nsteps[:M])
    ctx.set_device_drops([(m, SEXS, 1, 1) for m in range(M)])
    rng = np.random.default_rng(9)
    z = case.desc.z0[None, :] * (1 + 1e-3 * rng.standard_normal((M, case.desc.n)))
    zold = np.tile(case.desc.z0, (M, 1))
    on = np.array([1, 0], np.uint8)   # the drop is in force in scenario 0 only
    dp = int(case.desc.dev[SEXS][1][0])
    rows = [dp, dp + 1]
    n = case.desc.n
    f = ctx.eval_residual(z, zold, case.dt, on)
    rp, ci, vals = ctx.eval_jacobian(z, case.dt, on)
    assert np.array_equal(f[0, rows], z[0, rows] - zold[0, rows])
    others = np.setdiff1d(np.arange(n), rows)
    f_on = ctx.eval_residual(z, zold, case.dt, np.zeros(M, np.uint8))   # nothing in force
    assert np.array_equal(f[0, others], f_on[0, others])
    assert np.max(np.abs(f_on[0, rows] - f[0, rows])) > 1e-6
    J0 = sp.csr_matrix((vals[0], ci, rp), shape=(n, n)).toarray()
    J1 = sp.csr_matrix((vals[1], ci, rp), shape=(n, n)).toarray()
    assert np.array_equal(J0[rows], np.eye(n)[rows])
    assert np.max(np.abs(J1[rows] - np.eye(n)[rows])) > 1e-3   # online: the exciter's own rows
    ctx.close()


@pytest.mark.parametrize("method,policy,lock", [("mono", 0, False), ("schur64", 0, False), ("schur64", 2, True),
                                                ("schur32", 3, False)])
def test_trajectories_with_exciter_drops_match_oracle(method, policy, lock):
    from oracle.oracle import Oracle
    case = _case()
    M = 8
    nsteps = np.full(M, 120, np.int32)
    ev, dr, per = _tables(case, M)
    ctx = _ctx(case, M, refactor_policy=policy, lock_step=lock)
    ctx.set_events(ev, nsteps)
    ctx.set_device_drops(dr)
    st = ctx.integrate(METHODS[method], case.dt, TOL, 50)
    o = Oracle(case.desc)
    for m in range(M):
        faults, drops = per[m]
        ref, _, ost = o.integrate(case.dt, int(nsteps[m]), faults, tol=TOL, drops=drops)
        assert ost["not_converged"] == 0
        dev = np.max(np.abs(ctx.trajectory(m).T - ref))
        assert dev < 1e-9, (method, policy, lock, m, dev, st)
    ctx.close()


@pytest.mark.parametrize("method", ["mono", "schur32"])
def test_esdc1a_drop_matches_oracle(method):
    from oracle.oracle import Oracle
    case = get_case("case_psse", raw=os.path.join(G, "2bus.raw"), dyr=os.path.join(G, "2bus_ESDC1A.dyr"), M=4)
    M = 4
    nsteps = np.full(M, 120, np.int32)
    ev = [(m, 1, 0.03 + 0.01 * m, 6, 12 + m) for m in range(M)]
    ctx = _ctx(case, M, refactor_policy=3 if method == "schur32" else 0)
    ctx.set_events(ev, nsteps)
    ctx.set_device_drops([(m, ESDC1A, 0, 5 + 9 * m) for m in range(M)])
    ctx.integrate(METHODS[method], case.dt, TOL, 50)
    o = Oracle(case.desc)
    for m in range(M):
        ref, _, ost = o.integrate(case.dt, 120, [ev[m][1:]], tol=TOL, drops=[("esdc1a", 0, 5 + 9 * m)])
        assert ost["not_converged"] == 0
        dev = np.max(np.abs(ctx.trajectory(m).T - ref))
        assert dev < 1e-9, (method, m, dev)
    ctx.close()


def test_adjoint_with_exciter_drops_matches_oracle():
    from oracle.oracle import Oracle
    case = _case()
    M = 8
    nsteps = np.array([90, 120, 100, 60, 80, 120, 110, 100], np.int32)
    ev, dr, per = _tables(case, M)
    d = case.desc.dev
    params = [int(r[2] + 6) for r in d[0]] + [int(r[1] + 2) for r in d[SEXS]]   # GENROU H, SEXS K
    ctx = _ctx(case, M)
    ctx.set_events(ev, nsteps)
    ctx.set_device_drops(dr)
    r = ctx.adjoint(0, case.dt, params, tol=TOL, segment=25, scale=0.5, want_dz0=True)
    o = Oracle(case.desc)
    for m in range(M):
        faults, drops = per[m]
        val, g, dz0 = o.adjoint(case.dt, int(nsteps[m]), params, faults, tol=TOL, scale=0.5, drops=drops)
        assert r["value_scen"][m] == pytest.approx(val, rel=1e-9, abs=1e-18), m
        assert np.max(np.abs(r["grad_scen"][m] - g)) <= 1e-7 * np.max(np.abs(g)) + 1e-15, m
        assert np.max(np.abs(r["dz0"][m] - dz0)) <= 1e-7 * np.max(np.abs(dz0)) + 1e-15, m
    ctx.close()


def test_device_drop_refusals():
    from gradpower_hip.lib import GP_ERR_BADARG, GP_ERR_UNSUPPORTED, GpError
    case = _case()
    ctx = _ctx(case, 2)
    ctx.set_events([], case.nsteps[:2])
    with pytest.raises(GpError) as e:
        ctx.set_device_drops([(0, 0, 0, 5)])       # a GENROU
    assert e.value.code == GP_ERR_UNSUPPORTED
    with pytest.raises(GpError) as e:
        ctx.set_device_drops([(0, SEXS, 3, 5)])    # three SEXS rows
    assert e.value.code == GP_ERR_BADARG
    ctx.close()
