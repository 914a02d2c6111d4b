"""GPU parity of the shared-factor method (GP_METHOD_SCHUR_SHARED: the
reference's integrate_gpu_shared! and integrate_gpu_shared_multi!,
ext/GradPowerCUDAExt/GradPowerCUDAExt.jl:2296-2375, 2640-2708) against the
oracle, through the C ABI.

Newton step: when every scenario sits at scenario 0's state, a scenario's
reduced matrix with faults in force is P plus the faults' diagonal terms
exactly, so the Woodbury-corrected step equals the exact step -J^{-1} F to
rounding (1e-9 relative) for any set of faults: one bus (rank two), two
buses (rank four), one bus faulted twice (the admittances add).  A wrong
sign or a missing term of the correction fails it by O(1).
Trajectories: the shared method iterates each scenario to the same
|F|inf < 1e-12 as the oracle (the inexact Newton needs more iterations,
and a solve that stalls on the shared factor runs again as exact Newton),
so converged states agree to 1e-9 pu.  The reference's own tests (test/test_gpu_backend.jl:281-357)
compare against its exact Schur path at 1e-4.
"""
import numpy as np
import pytest

from conftest import get_case

pytestmark = pytest.mark.gpu
SHARED = 3


def _ctx(case, M, **kw):
    from gradpower_hip.lib import HipContext
    c = HipContext(case.desc, M, **kw)
    c.set_state(case.desc.z0, case.desc.p)
    return c


def _oracle(case):
    from oracle.oracle import Oracle
    return Oracle(case.desc)


@pytest.mark.parametrize("name", ["case0_2bus", "case1_9bus", "case3_200bus", "case2_60bus"])
def test_shared_newton_step_equals_exact_step(name):
    case = get_case(name)
    M = 4
    ctx = _ctx(case, M)
    o = _oracle(case)
    rng = np.random.default_rng(7)
    z0 = case.desc.z0
    z1 = z0 * (1 + 0.005 * rng.standard_normal(len(z0))) + 0.005 * rng.standard_normal(len(z0))
    z = np.tile(z1, (M, 1))   # every scenario at scenario 0's state
    zold = z0[None, :] * (1 + 0.002 * rng.standard_normal((M, len(z0))))
    b0, b1 = (int(x) for x in rng.choice(case.desc.nbus, 2, replace=False))
    # scenario 0: no fault; 1: one bus; 2: two buses; 3: one bus faulted twice
    ev = [(1, b0, 0.05, 1, 2), (2, b0, 0.05, 1, 2), (2, b1, 0.1, 1, 2), (3, b1, 0.05, 1, 2), (3, b1, 0.2, 1, 2)]
    ctx.set_events(ev, [10] * M)
    for dt in (case.dt, 0.0):
        dz = ctx.newton_step(SHARED, z, zold, dt, np.ones(M, np.uint8))
        for m in range(M):
            g = np.zeros(case.desc.nbus)
            for (mm, bus, r, _, _) in ev:
                if mm == m:
                    g[bus] += 1.0 / r
            f = o.residual(z[m], zold[m], dt, g)
            ref = -o.linsolve(o.jacobian(z[m], dt, g), f)
            err = np.max(np.abs(dz[m] - ref))
            assert err <= 1e-9 * (1 + np.max(np.abs(ref))), (name, m, dt, err)
    ctx.close()


@pytest.mark.parametrize("name,nsteps", [("case0_2bus", 240), ("case1_9bus", None), ("case3_200bus", 60),
                                         ("case2_60bus", 120)])
def test_shared_trajectory_matches_oracle(name, nsteps):
    case = get_case(name)
    M = min(case.M, 8)
    tol = 1e-12
    ctx = _ctx(case, M)
    ns = [int(case.nsteps[m]) if nsteps is None else nsteps for m in range(M)]
    ev = [(m, e.bus, e.rfault, int(round(e.ton / case.dt)), int(round(e.toff / case.dt)))
          for m in range(M) for e in case.events[m]]
    ctx.set_events(ev, ns)
    st = ctx.integrate(SHARED, case.dt, tol, 50)
    # factorisations only for the solves run again as exact Newton
    assert st["exact_retries"] > 0 or (st["refactorizations"] == 0 and st["factor_launches"] == 0)
    o = _oracle(case)
    for m in range(M):
        evm = [(e.bus, e.rfault, int(round(e.ton / case.dt)), int(round(e.toff / case.dt))) for e in case.events[m]]
        ref, _, so = o.integrate(case.dt, ns[m], evm, tol=tol, itermax=50)
        assert so["not_converged"] == 0
        dev = np.max(np.abs(ctx.trajectory(m).T - ref))
        assert dev < 1e-9, (name, m, dev, st)
    ctx.close()


def test_shared_multi_front_end_matches_single_scenarios():
    """integrate_hip_shared_multi (integrate_gpu_shared_multi!): four
    scenarios faulted at four buses in one call give the trajectories of
    four one-scenario calls (the reference's test_gpu_backend.jl:309-357 at
    1e-6; here each scenario's arithmetic is its own, so bit for bit)."""
    from gradpower_hip.batched import integrate_hip_shared, integrate_hip_shared_multi
    from gradpower_hip.lib import HipContext
    case = get_case("case1_9bus")
    buses = [0, 1, 2, 3]

    class G:   # the layout object the front end takes
        pass

    def layout(M):
        g = G()
        g.M = M
        g.ctx = HipContext(case.desc, M, dense_block=-1)   # the same plan in every context
        g.ctx.set_state(case.desc.z0, case.desc.p)
        return g

    gb = layout(4)
    _, tr = integrate_hip_shared_multi(gb, None, 1.0, buses, [0.02] * 4, 0.1, 0.2)
    gb.ctx.close()
    for m, bus in enumerate(buses):
        g1 = layout(1)
        tvec, one = integrate_hip_shared_multi(g1, None, 1.0, [bus], [0.02], 0.1, 0.2)
        g1.ctx.close()
        assert np.array_equal(tr[m], one[0]), (m, bus)


def test_shared_refusals():
    from gradpower_hip.lib import GpError
    case = get_case("case1_9bus")
    M = 2
    ev = [(m, case.events[m][0].bus, case.events[m][0].rfault, 2, 4) for m in range(M)]
    ctx = _ctx(case, M, lock_step=True)
    ctx.set_events(ev, [20] * M)
    with pytest.raises(GpError) as e:
        ctx.integrate(SHARED, case.dt, 1e-10, 30)
    assert e.value.code == -7
    ctx.close()
    ctx = _ctx(case, M)
    ctx.set_events(ev, [20] * M)
    with pytest.raises(GpError) as e:
        ctx.adjoint(SHARED, case.dt, [0])
    assert e.value.code == -7
    ctx.close()
