"""CPU checks of exciters taken offline in the oracle (DisconnectDeviceEvent,
src/GradPower.jl:130-143,681-708).  The reference's GPU kernels skip an
offline device into a zeroed F (ext/GradPowerCUDAExt/GradPowerCUDAExt.jl:1340,
1072-1199), so an offline exciter's rows of F are zero and backward Euler
holds its states.  No reference test covers it; what the model fixes:

  - an exciter taken offline at the end of step 1 (the system still at its
    operating point) gives the sweep of the same system read without that
    exciter's dyr record, whose generator then keeps the initial field
    voltage: generator speeds and bus voltages agree to 1e-9 pu (SEXS on the
    IEEE 9-bus case, ESDC1A on the 2-bus case);
  - after the drop the exciter's states stay bit for bit constant, and
    before it the sweep is bit for bit the sweep without the drop;
  - the adjoint with an exciter drop agrees with central differences, the
    dropped exciter's own gain included;
  - devices with algebraic states cannot be disconnected.
"""
import os

import numpy as np
import pytest

from conftest import get_case

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "psse")


def _oracle(desc):
    from oracle.oracle import Oracle
    return Oracle(desc)


def _ieee9(dyr="ieee9bus_SEXS.dyr"):
    return get_case("case_psse", raw=os.path.join(G, "ieee9_v33.raw"), dyr=os.path.join(G, dyr) if
                    not os.path.isabs(dyr) else dyr, M=4)


def _observed(case):
    """Generator speeds and every bus voltage (rows common to both systems)."""
    d = case.desc
    return [int(r[0] + 4) for r in d.dev[0]] + list(range(d.net_ptr, d.net_ptr + 2 * d.nbus))


def _target(case, kind, nth):
    from gradpower_hip.system import DisconnectDeviceEvent
    devs = case.ps.dynamic.devices
    di = [i for i, d in enumerate(devs) if d.dtype.kind == kind][nth]
    return DisconnectDeviceEvent(di, 0.0).target(case.ps)


@pytest.mark.parametrize("which", ["sexs", "esdc1a"])
def test_exciter_offline_from_the_start_equals_system_without_it(which, tmp_path):
    if which == "sexs":
        a = _ieee9()
        lines = open(os.path.join(G, "ieee9bus_SEXS.dyr")).read().splitlines()
        keep = [ln for ln in lines if not (ln.split() and ln.split()[0] == "2" and "'SEXS'" in ln)]
        assert len(keep) == len(lines) - 1
        p = tmp_path / "ieee9bus_SEXS_no2.dyr"
        p.write_text("\n".join(keep) + "\n")
        b = _ieee9(str(p))
        t, row = _target(a, "sexs", 1)
        ev = [(4, 0.02, 6, 14)]
    else:
        a = get_case("case_psse", raw=os.path.join(G, "2bus.raw"), dyr=os.path.join(G, "2bus_ESDC1A.dyr"), M=2)
        b = get_case("case_psse", raw=os.path.join(G, "2bus.raw"), dyr=os.path.join(G, "2bus.dyr"), M=2)
        t, row = _target(a, "esdc1a", 0)
        ev = [(1, 0.05, 6, 14)]
    assert (t, row) == ((3, 1) if which == "sexs" else (4, 0))
    n = 120
    ta, _, sa = _oracle(a.desc).integrate(a.dt, n, ev, tol=1e-12, drops=[(which, row, 1)])
    tb, _, sb = _oracle(b.desc).integrate(b.dt, n, ev, tol=1e-12)
    assert sa["not_converged"] == 0 and sb["not_converged"] == 0
    ra, rb = _observed(a), _observed(b)
    dev = np.max(np.abs(ta[:, ra] - tb[:, rb]))
    assert dev < 1e-9, dev
    # and the exciter does matter: online, the sweep differs
    t0, _, _ = _oracle(a.desc).integrate(a.dt, n, ev, tol=1e-12)
    assert np.max(np.abs(t0[:, ra] - tb[:, rb])) > 1e-4


def test_exciter_states_hold_after_the_drop():
    case = _ieee9()
    o = _oracle(case.desc)
    t, row = _target(case, "sexs", 0)
    dp = int(case.desc.dev[t][row][0])
    k0, n = 20, 80
    ev = [(5, 0.03, 6, 12)]
    tr, _, st = o.integrate(case.dt, n, ev, tol=1e-12, drops=[("sexs", row, k0)])
    assert st["not_converged"] == 0
    x = tr[k0:, dp:dp + 2]
    assert np.array_equal(x, np.tile(x[0], (x.shape[0], 1)))
    t0, _, _ = o.integrate(case.dt, n, ev, tol=1e-12)
    assert np.array_equal(tr[:k0 + 1], t0[:k0 + 1])
    assert np.max(np.abs(t0[k0 + 1:, dp + 1] - x[0, 1])) > 1e-4   # online, e_fd moves


def test_adjoint_with_exciter_drop_vs_central_differences():
    case = _ieee9()
    o = _oracle(case.desc)
    dt, n = case.dt, 70
    ev = [(6, 0.08, 5, 11)]
    d = case.desc.dev
    t, row = _target(case, "sexs", 1)
    k_dropped = int(d[3][row][1] + 2)                 # its gain K
    params = [int(d[0][0][2] + 6), int(d[3][0][1] + 2), k_dropped]
    rows = [int(r[0] + 4) for r in d[0]]
    gk = {}
    for k0 in (30, 1):
        kw = dict(rows=rows, tol=1e-12, drops=[("sexs", row, k0)])
        val, g, _ = o.adjoint(dt, n, params, ev, **kw)
        for j, q in enumerate(params):
            h = max(1e-6 * abs(o.p[q]), 1e-6)
            pp, pm = o.p.copy(), o.p.copy()
            pp[q] += h
            pm[q] -= h
            fd = (o.objective(dt, n, ev, p=pp, **kw) - o.objective(dt, n, ev, p=pm, **kw)) / (2 * h)
            assert abs(g[j] - fd) <= 1e-5 * abs(fd) + 1e-14, (k0, q, g[j], fd)
        gk[k0] = g[2]
    # the gain acts only while the exciter is online: dropped after step 1
    # its pullback is that of one step at the operating point
    assert abs(gk[1]) < abs(gk[30]), gk


def test_devices_with_algebraic_states_cannot_be_disconnected():
    case = _ieee9()
    with pytest.raises(NotImplementedError):
        _target(case, "genrou", 0)
    assert _target(case, "zipload", 1) == (6, 1)
