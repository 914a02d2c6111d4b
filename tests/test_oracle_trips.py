"""CPU checks of line trips and load disconnects in the oracle (TripLineEvent,
src/GradPower.jl:140-167,509-623; integrate!'s trip/reclose handling,
src/dynamics.jl:536-540,616-624).  No reference test covers them, so they are
pinned by what the network algebra fixes:

  - the stamp of create_trip_line_event equals Y(all branches) - Y(without
    the branch) from the independent Y builder, for a line with charging and
    for a transformer with off-nominal ratio and phase shift;
  - a sweep that trips a line at step k0 continues exactly as a sweep on the
    network built without that line, started from the state after the
    trip's dt = 0 re-solve (1e-10 pu); a reclosure returns to the sweep on
    the full network the same way;
  - the adjoint with trips (and reclosures) agrees with central differences
    of the objective, and a planted one-step misalignment of the event state
    breaks that agreement;
  - a ZIP load disconnected at step k0 (DisconnectDeviceEvent, src/GradPower.jl:
    130-143,681-698) continues as the sweep whose load has zero power and
    admittance; the adjoint with disconnects agrees with central differences.

The system: the bundled IEEE 9-bus case (ieee9_v33.raw + ieee9bus_SEXS.dyr)
with its first transformer at ratio 1.05 and a 2.5 degree phase shift.
"""
import copy
import os

import numpy as np
import pytest

from conftest import get_case

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "psse")


def _case():
    return get_case("case_psse", raw=os.path.join(G, "ieee9_v33.raw"), dyr=os.path.join(G, "ieee9bus_SEXS.dyr"),
                    M=4, xf_tap=(0, 1.05, 2.5))


def _oracle(desc):
    from oracle.oracle import Oracle
    return Oracle(desc)


def _without(case, fr, to):
    """The case's descriptor with the network rebuilt without branch fr-to."""
    from gradpower_hip.network import build_network
    from gradpower_hip.desc import build_desc
    ps = copy.deepcopy(case.ps)
    k = next(i for i, b in enumerate(ps.branches) if (b.fr, b.to) in ((fr, to), (to, fr)))
    del ps.branches[k]
    build_network(ps)
    return build_desc(ps, case.dp)


def _trip(case, a_ext, b_ext, on, off=0):
    from gradpower_hip.system import create_trip_line_event
    t = create_trip_line_event(case.ps, a_ext, b_ext, 0.0)
    return t, (t.from_bus, t.to_bus, on, off, t.stamp())


@pytest.mark.parametrize("a,b", [(4, 5), (1, 4), (7, 8)])
def test_stamp_is_the_branch_share_of_y(a, b):
    """Line 4-5 has charging; 1-4 is the tapped, phase-shifting transformer."""
    case = _case()
    t, _ = _trip(case, a, b, 1)
    y_full = case.ps.network.ybus.toarray()
    from gradpower_hip.network import build_network
    ps = copy.deepcopy(case.ps)
    k = next(i for i, x in enumerate(ps.branches) if (x.fr, x.to) in ((t.from_bus, t.to_bus), (t.to_bus, t.from_bus)))
    del ps.branches[k]
    y_less = build_network(ps).ybus.toarray()
    d = y_full - y_less
    f, o = t.from_bus, t.to_bus
    for (i, j), y in (((f, f), t.yff), ((f, o), t.yft), ((o, f), t.ytf), ((o, o), t.ytt)):
        assert abs(d[i, j] - y) <= 1e-12 * max(1.0, abs(y)), (i, j)
    d[[f, f, o, o], [f, o, f, o]] = 0.0
    assert np.max(np.abs(d)) <= 1e-12
    if (a, b) == (1, 4):
        assert abs(t.yft - t.ytf) > 1e-3    # the phase shift makes Y unsymmetric


@pytest.mark.parametrize("a,b", [(4, 5), (1, 4)])
def test_trip_continues_as_the_network_without_the_line(a, b):
    case = _case()
    dt, k0, n = case.dt, 30, 90
    ev = [(7, 0.05, 6, 12)]                     # a fault before the trip
    t, row = _trip(case, a, b, k0)
    o = _oracle(case.desc)
    tr, _, st = o.integrate(dt, n, ev, tol=1e-12, trips=[row])
    assert st["not_converged"] == 0
    _, z_k0, _ = o.integrate(dt, k0, ev, tol=1e-12, trips=[row], traj=False)
    ob = _oracle(_without(case, t.from_bus, t.to_bus))
    trb, _, stb = ob.integrate(dt, n - k0, (), z0=z_k0, tol=1e-12)
    assert stb["not_converged"] == 0
    assert np.max(np.abs(tr[k0 + 1:] - trb[1:])) < 1e-10
    # before the trip the sweep is the untripped one
    tr0, _, _ = o.integrate(dt, n, ev, tol=1e-12)
    assert np.array_equal(tr[:k0 + 1], tr0[:k0 + 1])
    assert np.max(np.abs(tr[-1] - tr0[-1])) > 1e-3


def test_reclose_returns_to_the_full_network():
    case = _case()
    dt, k0, k1, n = case.dt, 20, 50, 90
    t, row = _trip(case, 4, 5, k0, k1)
    o = _oracle(case.desc)
    tr, _, st = o.integrate(dt, n, (), tol=1e-12, trips=[row])
    assert st["not_converged"] == 0
    _, z_k1, _ = o.integrate(dt, k1, (), tol=1e-12, trips=[row], traj=False)
    tr1, _, _ = o.integrate(dt, n - k1, (), z0=z_k1, tol=1e-12)
    assert np.max(np.abs(tr[k1 + 1:] - tr1[1:])) < 1e-10
    # and a sweep's Y is restored when it ends
    tr2, _, _ = o.integrate(dt, n, (), tol=1e-12)
    assert np.max(np.abs(tr2 - tr2[0])) < 1e-9


def _speed_rows(case):
    return [int(r[0] + 4) for r in case.desc.dev[0]]


def test_adjoint_with_trips_vs_central_differences():
    case = _case()
    dt, n = case.dt, 80
    ev = [(6, 0.08, 5, 11)]
    _, r1 = _trip(case, 4, 5, 15, 40)
    _, r2 = _trip(case, 1, 4, 25)
    trips = [r1, r2]
    o = _oracle(case.desc)
    d = case.desc.dev
    params = [int(d[0][0][2] + 6), int(d[0][1][2] + 7), int(d[3][2][1] + 2)]   # H, D, SEXS K
    rows = _speed_rows(case)
    val, g, _ = o.adjoint(dt, n, params, ev, rows=rows, tol=1e-12, trips=trips)
    assert val == pytest.approx(o.objective(dt, n, ev, rows=rows, tol=1e-12, trips=trips), rel=1e-12)
    for j, q in enumerate(params):
        h = max(1e-6 * abs(o.p[q]), 1e-6)
        pp, pm = o.p.copy(), o.p.copy()
        pp[q] += h
        pm[q] -= h
        fd = (o.objective(dt, n, ev, rows=rows, tol=1e-12, p=pp, trips=trips) -
              o.objective(dt, n, ev, rows=rows, tol=1e-12, p=pm, trips=trips)) / (2 * h)
        assert abs(g[j] - fd) <= 1e-5 * abs(fd), (q, g[j], fd)
    # the event state one step off (faults and trips alike) breaks it
    _, gs, _ = o.adjoint(dt, n, params, ev, rows=rows, tol=1e-12, trips=trips, fault_shift=1)
    assert np.max(np.abs(gs - g) / np.abs(g)) > 1e-3


def _zip_rows(case):
    """Parameter offset of each ZIP table row."""
    return [int(r[1]) for r in case.desc.dev[6]]


def test_load_drop_continues_as_a_zero_load():
    """A ZIP load disconnected at step k0 (DisconnectDeviceEvent) continues as
    the sweep whose load has zero power and admittance, from the state after
    the dt = 0 re-solve; the helper maps a device index to its ZIP row."""
    from gradpower_hip.system import DisconnectDeviceEvent
    case = _case()
    dt, k0, n = case.dt, 25, 80
    devs = case.ps.dynamic.devices
    di = [i for i, d in enumerate(devs) if d.dtype.kind == "zipload"][1]
    row = DisconnectDeviceEvent(di, k0 * dt).zip_row(case.ps)
    assert row == 1
    with pytest.raises(NotImplementedError):
        DisconnectDeviceEvent(0, 0.1).zip_row(case.ps)
    o = _oracle(case.desc)
    ev = [(7, 0.05, 6, 12)]
    tr, _, st = o.integrate(dt, n, ev, tol=1e-12, drops=[(row, k0)])
    assert st["not_converged"] == 0
    _, z_k0, _ = o.integrate(dt, k0, ev, tol=1e-12, drops=[(row, k0)], traj=False)
    p0 = case.desc.p.copy()
    q = _zip_rows(case)[row]
    p0[[q, q + 1, q + 7, q + 8]] = 0.0          # pinj, qinj, yreal, yimag
    trb, _, _ = o.integrate(dt, n - k0, (), z0=z_k0, tol=1e-12, p=p0)
    assert np.max(np.abs(tr[k0 + 1:] - trb[1:])) < 1e-10
    tr0, _, _ = o.integrate(dt, n, ev, tol=1e-12)
    assert np.array_equal(tr[:k0 + 1], tr0[:k0 + 1])
    assert np.max(np.abs(tr[-1] - tr0[-1])) > 1e-3


def test_adjoint_with_load_drop_vs_central_differences():
    case = _case()
    dt, n = case.dt, 70
    ev = [(6, 0.08, 5, 11)]
    _, r1 = _trip(case, 4, 5, 30)
    drops = [(0, 18), (2, 40)]
    o = _oracle(case.desc)
    d = case.desc.dev
    params = [int(d[0][0][2] + 6), int(d[0][2][2] + 7), int(d[3][1][1] + 2)]
    rows = _speed_rows(case)
    kw = dict(rows=rows, tol=1e-12, trips=[r1], drops=drops)
    val, g, _ = o.adjoint(dt, n, params, ev, **kw)
    for j, q in enumerate(params):
        h = max(1e-6 * abs(o.p[q]), 1e-6)
        pp, pm = o.p.copy(), o.p.copy()
        pp[q] += h
        pm[q] -= h
        fd = (o.objective(dt, n, ev, p=pp, **kw) - o.objective(dt, n, ev, p=pm, **kw)) / (2 * h)
        assert abs(g[j] - fd) <= 1e-5 * abs(fd), (q, g[j], fd)
    _, gs, _ = o.adjoint(dt, n, params, ev, fault_shift=1, **kw)
    assert np.max(np.abs(gs - g) / np.abs(g)) > 1e-3
