"""Network admittance and power flow.

build_network  -- src/network.jl:1-126 (complex Y, realified 2x2 blocks [g -b; b g])
runpf          -- src/pflow.jl:286-418 (polar Newton power flow + generator write-back)
"""
from __future__ import annotations

import cmath
import math

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from .system import PowerSystem, PQ, PV, SLACK


class Network:
    def __init__(self, adjacency, ybus, ybus_real):
        self.adjacency = adjacency
        self.ybus = ybus              # complex CSR, nbus x nbus
        self.ybus_real = ybus_real    # real CSR, 2nbus x 2nbus


def create_ybus_complex(ps: PowerSystem):
    """Branch pi-models with off-nominal taps and shunts (src/network.jl:29-85)."""
    n = ps.nbus
    acc = {}

    def add(r, c, v):
        acc[(r, c)] = acc.get((r, c), 0.0) + v

    for br in ps.branches:
        tap, shift = br.tap, br.shift
        if tap > 0.0:
            tpsh = tap * cmath.exp(1j * math.pi / 180.0 * shift)
        else:
            tpsh, tap = 1.0 + 0j, 1.0
        y = 1.0 / complex(br.r, br.x)
        fr, to = br.fr, br.to
        add(fr, fr, y / (tap * tap))
        add(to, to, y)
        add(fr, to, -y / tpsh.conjugate())
        add(to, fr, -y / tpsh)
        add(to, to, 1j * 0.5 * br.sh / (tap * tap))
        add(fr, fr, 1j * 0.5 * br.sh)
    for sh in ps.shunts:
        add(sh.bus, sh.bus, complex(sh.gsh, sh.bsh))
    keys = sorted(acc)
    rows = np.array([k[0] for k in keys], dtype=np.int64)
    cols = np.array([k[1] for k in keys], dtype=np.int64)
    vals = np.array([acc[k] for k in keys], dtype=np.complex128)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def realify_ybus(ybus):
    """Each complex entry g+jb becomes [g -b; b g] (src/network.jl:87-126)."""
    c = ybus.tocoo()
    g, b = c.data.real, c.data.imag
    r2, c2 = 2 * c.row, 2 * c.col
    rows = np.concatenate([r2, r2 + 1, r2, r2 + 1])
    cols = np.concatenate([c2, c2 + 1, c2 + 1, c2])
    vals = np.concatenate([g, g, -b, b])
    n = ybus.shape[0]
    m = sp.csr_matrix((vals, (rows, cols)), shape=(2 * n, 2 * n))
    m.sum_duplicates()
    m.sort_indices()
    return m


def adjacency_list(ps: PowerSystem):
    adj = [set() for _ in range(ps.nbus)]
    for br in ps.branches:
        adj[br.fr].add(br.to)
        adj[br.to].add(br.fr)
    return [sorted(a) for a in adj]


def build_network(ps: PowerSystem):
    y = create_ybus_complex(ps)
    ps.network = Network(adjacency_list(ps), y, realify_ybus(y))
    return ps.network


def _injections(ps):
    n = ps.nbus
    pinj = np.zeros(n)
    qinj = np.zeros(n)
    for g in ps.gens:
        pinj[g.bus] += g.psch
        qinj[g.bus] += g.qsch
    for ld in ps.loads:
        pinj[ld.bus] -= ld.pd
        qinj[ld.bus] += ld.qd       # qd already carries the injection sign
    return pinj, qinj


def _power(ybus, vm, va):
    v = vm * np.exp(1j * va)
    s = v * np.conj(ybus @ v)
    return s.real, s.imag


def runpf(ps: PowerSystem, tol=1e-11, maxit=100, dc_start=False, max_step=None):
    """Polar Newton power flow on |V| (PQ buses) and angle (PQ+PV buses),
    the unknown set of src/pflow.jl:286-371.  Returns (vm, va, P, Q) with the
    net injections at the solution.  max_step (angle rad, magnitude pu), when
    given, shortens a Newton step whose largest change exceeds it: on grids
    of tens of thousands of buses with one slack, the angle block of the
    Jacobian is nearly singular and a full first step from a flat start
    moves every angle by hundreds of radians."""
    y = ps.network.ybus.tocsr()
    btype = np.array([b.type for b in ps.buses])
    vm = np.array([b.v0m for b in ps.buses], dtype=float)
    va = np.array([b.v0a for b in ps.buses], dtype=float)
    pinj, qinj = _injections(ps)
    pq = np.where(btype == PQ)[0]
    pvq = np.where((btype == PQ) | (btype == PV))[0]
    npq = len(pq)
    if dc_start and len(pvq):
        # DC power-flow angles as the starting point (B' from -imag(Y)).
        Bp = -y.imag.tocsr()[pvq][:, pvq].tocsc()
        va = va.copy()
        va[pvq] = va[btype == SLACK].mean() + spl.spsolve(Bp, pinj[pvq])
    for it in range(maxit):
        P, Q = _power(y, vm, va)
        f = np.concatenate([Q[pq] - qinj[pq], P[pvq] - pinj[pvq]])
        if np.max(np.abs(f), initial=0.0) < tol:
            break
        v = vm * np.exp(1j * va)
        ibus = y @ v
        diagV = sp.diags(v)
        diagI = sp.diags(ibus)
        diagVn = sp.diags(v / vm)
        dS_dVa = 1j * diagV @ np.conj(diagI - y @ diagV)
        dS_dVm = diagV @ np.conj(y @ diagVn) + np.conj(diagI) @ diagVn
        dS_dVa = sp.csr_matrix(dS_dVa)
        dS_dVm = sp.csr_matrix(dS_dVm)
        J = sp.bmat([[dS_dVm[pq][:, pq].imag, dS_dVa[pq][:, pvq].imag],
                     [dS_dVm[pvq][:, pq].real, dS_dVa[pvq][:, pvq].real]]).tocsc()
        dx = spl.spsolve(J, -f)
        # Backtracking on |f|inf: plain Newton from a flat start overshoots on
        # weak meshed grids; the accepted point is still a Newton solution.
        f0n = np.max(np.abs(f))
        lam = 1.0
        if max_step is not None:
            da = np.max(np.abs(dx[npq:]), initial=0.0) / max_step[0]
            dm = np.max(np.abs(dx[:npq]), initial=0.0) / max_step[1]
            lam = min(1.0, 1.0 / max(da, dm, 1e-300))
        while True:
            vm_t, va_t = vm.copy(), va.copy()
            vm_t[pq] += lam * dx[:npq]
            va_t[pvq] += lam * dx[npq:]
            Pt, Qt = _power(y, vm_t, va_t)
            ft = np.concatenate([Qt[pq] - qinj[pq], Pt[pvq] - pinj[pvq]])
            if np.max(np.abs(ft)) < f0n or lam < 1e-3:
                break
            lam *= 0.5
        vm, va = vm_t, va_t
    else:
        raise RuntimeError(f"power flow did not converge (|f|={np.max(np.abs(f)):.3e})")
    return vm, va, P, Q


def runpf_inplace(ps: PowerSystem, exact_slack=False, **kw):
    """Write the solution back like src/pflow.jl:375-418: bus voltages, PV-bus
    reactive power and slack P/Q split evenly over each bus's generators.

    exact_slack=False repeats the reference's arithmetic (src/pflow.jl:393-397
    subtracts pd and qd from the net injection).  For Q this recovers the
    generator output; for P it is exact only where the slack bus has no load
    (the synthetic cases keep it load-free): a slack-bus load pd leaves the
    slack machines 2 pd short and the initial state off balance on that bus.
    exact_slack=True adds pd instead, so the machines carry the solved
    output."""
    vm, va, P, Q = runpf(ps, **kw)
    for k, b in enumerate(ps.buses):
        b.v0m = float(vm[k])
        b.v0a = float(va[k])
    sP, sQ = P.copy(), Q.copy()
    for ld in ps.loads:
        sP[ld.bus] += ld.pd if exact_slack else -ld.pd
        sQ[ld.bus] -= ld.qd
    bus_gens = [[] for _ in range(ps.nbus)]
    for gi, g in enumerate(ps.gens):
        bus_gens[g.bus].append(gi)
    for k, b in enumerate(ps.buses):
        ng = len(bus_gens[k])
        if b.type == PV:
            for gi in bus_gens[k]:
                ps.gens[gi].qsch = float(sQ[k] / ng)
        elif b.type == SLACK:
            for gi in bus_gens[k]:
                ps.gens[gi].psch = float(sP[k] / ng)
                ps.gens[gi].qsch = float(sQ[k] / ng)
    return vm, va
