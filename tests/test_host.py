"""Host-side checks (no GPU): setup layer, descriptor, C-ABI exports."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import get_case, ROOT


def test_library_exports_every_declared_symbol():
    from gradpower_hip import lib
    L = lib.load()
    header = open(os.path.join(ROOT, "include", "gradpower_hip.h")).read()
    declared = set(re.findall(r"^\s*(?:int|void|const char \*)\s*(gp_\w+)\s*\(", header, re.M))
    assert declared >= set(lib.EXPORTS)
    for name in declared:
        assert hasattr(L, name), name


@pytest.mark.parametrize("name", ["case_2bus_testdata", "case0_2bus", "case1_9bus", "case3_200bus"])
def test_descriptor_consistency(name):
    case = get_case(name)
    d = case.desc
    assert d.n == d.net_ptr + 2 * d.nbus
    # clusters tile the device states contiguously (src/clusters.jl:225-319)
    cl = d.clusters
    order = np.argsort(cl[:, 0])
    pos = 0
    for c in cl[order]:
        assert c[0] == pos
        pos += c[1]
    assert pos == d.net_ptr
    assert d.is_diff[d.net_ptr:].sum() == 0
    # every wired controller index points at a device state
    g = d.dev[0]
    assert np.all((g[:, 4] == -1) | ((g[:, 4] >= 0) & (g[:, 4] < d.net_ptr)))
    assert np.all((g[:, 5] == -1) | ((g[:, 5] >= 0) & (g[:, 5] < d.net_ptr)))


def test_power_flow_balance():
    from gradpower_hip.network import _power, _injections
    case = get_case("case1_9bus")
    ps = case.ps
    vm = np.array([b.v0m for b in ps.buses])
    va = np.array([b.v0a for b in ps.buses])
    P, Q = _power(ps.network.ybus.tocsr(), vm, va)
    pinj, qinj = _injections(ps)
    for k, b in enumerate(ps.buses):
        if b.type == 1:
            assert abs(P[k] - pinj[k]) < 1e-9 and abs(Q[k] - qinj[k]) < 1e-9
        elif b.type == 2:
            assert abs(P[k] - pinj[k]) < 1e-9


def test_event_steps_round_half_even():
    from gradpower_hip.desc import make_events
    from gradpower_hip.system import ContingencyEvent
    _, n, rows = make_events([[ContingencyEvent(0, 0.1, 2.5 / 120, 3.5 / 120)]], 1.0 / 120)
    assert rows[0][3] == 2 and rows[0][4] == 4      # round(2.5)=2, round(3.5)=4 as in Julia


def test_zip_alpha_and_static_generators_in_grid_case():
    case = get_case("case3_200bus")
    d = case.desc
    zp = d.dev[6]
    assert len(zp) > 0
    alphas = case.desc.p[zp[:, 1] + 2]
    assert np.allclose(alphas, 0.4)


def test_gensal_is_genrou_with_substitutions():
    """src/generators.jl:41-66: GENSAL = GENROU with x_qp = x_dp, T_q0p = T_d0p;
    the case initialises to an equilibrium."""
    from conftest import get_case
    from oracle.oracle import Oracle
    case = get_case("case1_9bus", gensal=True)
    p = case.desc.p
    rows = case.desc.dev[0]
    g = rows[1]                     # the bus-4 machine
    P = p[g[2]:g[2] + 14]
    assert P[3] == P[2] and P[9] == P[8]
    other = p[rows[0][2]:rows[0][2] + 14]
    assert other[3] != other[2] and other[9] != other[8]
    f = Oracle(case.desc).residual(case.desc.z0, case.desc.z0, case.dt)
    assert np.max(np.abs(f)) < 1e-9
