"""CPU checks of the oracle's discrete adjoint (no GPU needed).

The reference snapshot calls GradPower.adjoint (examples/adjoint.jl:34,106)
but no longer contains it, so the oracle's adjoint is pinned the way the
reference's own regression test pins its gradient: against central finite
differences of the objective along the forward sweep
(test/test_adjoint_event_offset.jl:26-84: 2-bus system, fault on bus 2 from
0.2 s to 0.3 s with r = 0.02, 4.8 s at dt = 1/120, dJ/dH relative error
< 1e-6).  A planted misalignment of the fault state (the "+1" of that file's
comment) must break the agreement, and the transposed sparse solve is checked
against scipy.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import get_case


def _oracle(case):
    from oracle.oracle import Oracle
    return Oracle(case.desc)


def _events(case, m=0):
    return [(e.bus, e.rfault, int(round(e.ton / case.dt)), int(round(e.toff / case.dt))) for e in case.events[m]]


def _fd(o, dt, nsteps, ev, q, h, **kw):
    p0 = o.p.copy()
    pp, pm = p0.copy(), p0.copy()
    pp[q] += h
    pm[q] -= h
    return (o.objective(dt, nsteps, ev, p=pp, **kw) - o.objective(dt, nsteps, ev, p=pm, **kw)) / (2 * h)


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def test_transposed_solve_vs_scipy():
    case = get_case("case1_9bus")
    o = _oracle(case)
    rng = np.random.default_rng(3)
    z = case.desc.z0 * (1 + 0.01 * rng.standard_normal(case.desc.n))
    J = o.jacobian_csc(z, case.dt)
    c = rng.standard_normal(case.desc.n)
    y = o.linsolve_t(o.jacobian(z, case.dt), c)
    y_ref = spla.spsolve(sp.csc_matrix(J.T), c)
    assert np.max(np.abs(J.T @ y - c)) <= 1e-10 * max(1.0, np.max(np.abs(c)))
    assert np.max(np.abs(y - y_ref)) <= 1e-9 * max(1.0, np.max(np.abs(y_ref)))


def _two_bus():
    case = get_case("case_2bus_testdata", fault=False, tf=4.8)
    dt = 1.0 / 120.0
    nsteps = int(round(4.8 / dt))
    ev = [(1, 0.02, int(round(0.2 / dt)), int(round(0.3 / dt)))]   # bus 2, r 0.02, 0.2 s - 0.3 s
    par = int(case.desc.dev[0][0][2] + 6)                          # GENROU H: par_ptr + 6
    return case, dt, nsteps, ev, par


def test_adjoint_event_offset_2bus():
    """test/test_adjoint_event_offset.jl:26-84: dJ/dH by the adjoint within
    1e-6 of the centred difference with h = 1e-5 max(|H|, 1)."""
    case, dt, nsteps, ev, par = _two_bus()
    o = _oracle(case)
    val, grad, _ = o.adjoint(dt, nsteps, [par], ev, tol=1e-12)
    assert val == pytest.approx(o.objective(dt, nsteps, ev, tol=1e-12), rel=1e-13)
    fd = _fd(o, dt, nsteps, ev, par, 1e-5 * max(abs(o.p[par]), 1.0), tol=1e-12)
    assert _rel(grad[0], fd) < 1e-6


@pytest.mark.parametrize("shift", [-1, 1])
def test_misaligned_fault_state_is_detected(shift):
    """The planted "+1"/"-1" alignment error moves dJ/dH far outside the
    1e-6 bar (the reference comment reports ~0.15 for its variant)."""
    case, dt, nsteps, ev, par = _two_bus()
    o = _oracle(case)
    _, grad, _ = o.adjoint(dt, nsteps, [par], ev, tol=1e-12, fault_shift=shift)
    fd = _fd(o, dt, nsteps, ev, par, 1e-5 * max(abs(o.p[par]), 1.0), tol=1e-12)
    assert _rel(grad[0], fd) > 1e-3


def test_adjoint_all_gains_vs_fd():
    """Every parameter kind with a pullback (GENROU H, D; SEXS K; ESDC1A KA;
    TGOV1 R; IEESGO K1; IEEEST KS) on a 60-bus grid with every controller
    type, against centred differences with h = 1e-6 max(|p|, 1)."""
    case = get_case("case2_grid", M=4, nbus=60, n_genrou=12, n_pss=4, n_static=6, seed=77)
    o = _oracle(case)
    d = case.desc.dev
    pars = [int(d[0][0][2] + 6), int(d[0][1][2] + 7), int(d[3][0][1] + 2), int(d[4][0][1]),
            int(d[2][0][2]), int(d[1][0][2] + 6), int(d[5][0][2] + 12)]
    ev = _events(case, 1)
    nsteps = 120
    val, grad, _ = o.adjoint(case.dt, nsteps, pars, ev, tol=1e-12)
    for j, q in enumerate(pars):
        fd = _fd(o, case.dt, nsteps, ev, q, 1e-6 * max(abs(o.p[q]), 1.0), tol=1e-12)
        assert _rel(grad[j], fd) < 1e-5, (q, grad[j], fd)


def test_adjoint_initial_state_gradient():
    """dJ/dz0: centred differences on differential entries of z0; zero on the
    algebraic entries (they only seed the first Newton solve)."""
    case = get_case("case1_9bus")
    o = _oracle(case)
    ev = _events(case, 0)
    nsteps = 60
    scale = 1.0 / 3.0
    _, _, dz0 = o.adjoint(case.dt, nsteps, [], ev, tol=1e-12, scale=scale)
    isd = case.desc.is_diff.astype(bool)
    assert np.all(dz0[~isd] == 0.0)
    z0 = case.desc.z0
    w_rows = [int(r[0] + 4) for r in case.desc.dev[0]]
    for i in w_rows[:2] + [int(np.flatnonzero(isd)[0])]:
        h = 1e-6 * max(abs(z0[i]), 1.0)
        zp, zm = z0.copy(), z0.copy()
        zp[i] += h
        zm[i] -= h
        fd = (o.objective(case.dt, nsteps, ev, z0=zp, tol=1e-12, scale=scale) -
              o.objective(case.dt, nsteps, ev, z0=zm, tol=1e-12, scale=scale)) / (2 * h)
        assert abs(dz0[i] - fd) <= 1e-5 * max(abs(fd), 1e-8), (i, dz0[i], fd)
