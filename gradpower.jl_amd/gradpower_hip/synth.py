"""Seeded synthetic cases for the benchmark and the parity tests.

Each builder returns a `Case` holding the set-up PowerSystem, the initialised
DynamicProblem, the flattened C-ABI descriptor and the per-scenario fault /
horizon tables.  Random draws use numpy's PCG64 with seed 1000 + case index;
the drawn values that the case definitions leave open are recorded in
`case.meta`.

  case_2bus_testdata   -- the reference's test/testdata/2bus.{raw,dyr} system
  case0_2bus           -- 2-bus GENROU+IEESGO, constant-impedance load, 10 s
  case1_9bus           -- 9-bus ring + 3 chords, 3 x GENROU+SEXS+TGOV1, M=64, 1 s
  case2_grid           -- random geometric grid (default 2,000 buses), 400 machines
                          with SEXS/ESDC1A + TGOV1/IEESGO, 100 IEEEST, 300 static
                          generators, M=256 uneven faults and horizons
  case3_200bus         -- 200 buses, 40 x GENROU+SEXS+TGOV1, M=128, 3 s
  case4_grid           -- 70,000 buses (35 tiles of the 2,000-bus grid), 12,005
                          machines (a quarter GENSAL) with exciters, governors
                          and stabilisers, 9,100 static generators, M=8, 2 s
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .system import (Bus, Gen, Load, Branch, Shunt, PowerSystem, PowerSystemDynamics,
                     Genrou, Gensal, IEESGO, TGOV1, SEXS, ESDC1A, IEEEST, ContingencyEvent,
                     esdc1a_sat_coefficients, set_dynamics, DynamicProblem, PQ, PV, SLACK)
from .network import build_network, runpf_inplace
from .init import initialize_dynamics
from .desc import build_desc, SystemDesc

DT = 1.0 / 120.0


@dataclass
class Case:
    name: str
    ps: PowerSystem
    dp: DynamicProblem
    desc: SystemDesc
    M: int
    dt: float
    events: list                  # per scenario: list of ContingencyEvent
    tf: np.ndarray                # per scenario horizon (s)
    meta: dict = field(default_factory=dict)

    @property
    def nsteps(self):
        return np.array([int(round(t / self.dt)) for t in self.tf], dtype=np.int32)


# Machine data of the 2-bus spec (on machine base).
C0_GENROU = dict(x_d=1.8, x_q=1.7, x_dp=0.3, x_qp=0.55, x_ddp=0.25, xl=0.2, H=3.5, D=0.0,
                 T_d0p=8.0, T_q0p=0.4, T_d0dp=0.03, T_q0dp=0.05, S1=0.03, S2=0.4)


def _genrou(bus, gid="1", **over):
    p = dict(C0_GENROU)
    p.update(over)
    return Genrou(bus, gid, p["x_d"], p["x_q"], p["x_dp"], p["x_qp"], p["x_ddp"], p["xl"],
                  p["H"], p["D"], p["T_d0p"], p["T_q0p"], p["T_d0dp"], p["T_q0dp"],
                  p["S1"], p["S2"])


SEXS_K = 50.0       # exciter gains of the synthetic grids
ESDC1A_KA = 46.0


def _sexs(bus, gid="1"):
    return SEXS(bus, gid, 0.1, 10.0, SEXS_K, 0.05, -5.0, 5.0)


def _tgov1(bus, gid="1"):
    return TGOV1(bus, gid, 0.05, 0.5, 1.1, 0.0, 2.1, 7.0, 0.0)


def _esdc1a(bus, gid="1"):
    sa, sb = esdc1a_sat_coefficients(2.8, 0.04, 3.73, 0.33)
    return ESDC1A(bus, gid, ESDC1A_KA, 0.06, 0.1, 1.0, 1.0, 0.46, 0.0, sa, sb)


def _ieesgo(bus, gid, rng, meta):
    T = rng.uniform(0.05, 0.5, size=6)
    meta.setdefault("ieesgo_T", []).append(T.tolist())
    return IEESGO(bus, gid, *T, 20.0, 0.5, 0.5, 1.0, 0.0)


def _ieeest(bus, gid="1"):
    return IEEEST(bus, gid, 1, 0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.15, 0.05, 0.15, 0.05,
                  2.0, 2.0, 5.0, 0.1, -0.1, 0.0, 0.0)


def _solve_pf_scaling(ps, meta, max_step=None, maxit=100, first=1.0, min_step=1.0 / 256):
    """Power flow by continuation in the loading: solve with every load and
    scheduled generation scaled by s, starting at s = 0 and raising s towards
    1 from the previous solution (halving the increment after a failure).  If
    s = 1 cannot be reached the largest solved s is kept and recorded in
    meta["load_scale"]."""
    from .network import runpf
    pd0 = [(ld.pd, ld.qd) for ld in ps.loads]
    pg0 = [g.psch for g in ps.gens]
    vset = [b.v0m for b in ps.buses]

    def apply(sc):
        for ld, (pd, qd) in zip(ps.loads, pd0):
            ld.pd, ld.qd = pd * sc, qd * sc
        for g, p0 in zip(ps.gens, pg0):
            g.psch = p0 * sc

    s_ok, step, sol = None, 1.0, None
    s_try = first
    if first < 1.0:
        step = 0.25
    while True:
        apply(s_try)
        try:
            vm, va, _, _ = runpf(ps, max_step=max_step, maxit=maxit)
            s_ok, sol = s_try, (vm, va)
        except RuntimeError:
            if s_try == 0.0:
                raise
            if s_ok is None:
                s_try, step = 0.0, 0.25
                continue
            step *= 0.5
            if step < min_step:
                break
            s_try = s_ok + step
            continue
        if s_ok >= 1.0:
            break
        for b, m, a, vs in zip(ps.buses, sol[0], sol[1], vset):
            b.v0a = float(a)
            b.v0m = float(m) if b.type == PQ else vs
        s_try = min(1.0, s_ok + step)
    apply(s_ok)
    for b, m, a, vs in zip(ps.buses, sol[0], sol[1], vset):
        b.v0a = float(a)
        b.v0m = float(m) if b.type == PQ else vs
    runpf_inplace(ps, max_step=max_step)
    meta["load_scale"] = s_ok


def _finish(name, ps, psd, M, events, tf, meta, zip_alpha=None, pq_after_pf=None, size_machines=False,
            pf_max_step=None, pf_kw=None):
    pf_done = False
    if pq_after_pf is not None or size_machines:
        # power flow first: optionally freeze the given voltage-controlled
        # buses as constant P/Q injections at the solved operating point, and
        # size each generator's MVA base to its solved loading (at most 80%),
        # which the machine models scale with (the slack machine in particular
        # picks up the losses and the dispatch mismatch)
        build_network(ps)
        _solve_pf_scaling(ps, meta, pf_max_step, **(pf_kw or {}))
        pf_done = True
        for b in pq_after_pf or []:
            ps.buses[int(b)].type = PQ
        if size_machines:
            for g in ps.gens:
                g.mbase = max(g.mbase, 100.0 * math.hypot(g.psch, g.qsch) / 0.8)
    set_dynamics(ps, psd)
    if zip_alpha is not None:
        for dev in psd.devices:
            if dev.dtype.kind == "zipload":
                dev.dtype.alpha = zip_alpha
    build_network(ps)
    if not pf_done:
        _solve_pf_scaling(ps, meta, pf_max_step, **(pf_kw or {}))
    else:
        runpf_inplace(ps, max_step=pf_max_step)
    # ZIP loads were created from ps.loads in order by set_dynamics; carry a
    # load scaling applied by the power-flow fallback over to them.
    zips = [dev.dtype for dev in psd.devices if dev.dtype.kind == "zipload"]
    for z, ld in zip(zips, ps.loads):
        z.pinj, z.qinj = ld.pd, ld.qd
    dp = DynamicProblem(ps)
    initialize_dynamics(dp, ps)
    desc = build_desc(ps, dp)
    return Case(name, ps, dp, desc, M, DT, events, np.asarray(tf, dtype=float), meta)


def case_2bus_testdata(fault=True, tf=1.0):
    """test/testdata/2bus.raw + 2bus.dyr: slack bus 1 with one GENROU, PQ bus 2
    with a 103.13 MW / 37.617 Mvar load, one line.  Values typed from those
    files (PSS/E v33 fields; gen mbase 100 MVA)."""
    buses = [Bus(1, "BUS1", SLACK, 345.0, 1.04, 0.0),
             Bus(2, "BUS2", PQ, 345.0, 1.01613, math.radians(-3.3252))]
    gens = [Gen(0, "1", 106.496 / 100.0, 45.996 / 100.0, 100.0)]
    loads = [Load(1, "11", 103.130 / 100.0, -37.617 / 100.0)]
    branches = [Branch(0, 1, "1", 1.0e-4, 5.76e-2, 0.0001)]
    ps = PowerSystem(100.0, buses, gens, loads, branches, [])
    psd = PowerSystemDynamics()
    psd.add_device(Genrou(1, "1", 1.575, 1.512, 0.291, 0.39, 0.1733, 0.0787, 3.38, 0.0,
                          6.1, 1.0, 0.05, 0.15, 0.0, 0.0))
    events = [[ContingencyEvent(1, 0.2, 0.2, 0.3)]] if fault else [[]]
    return _finish("2bus_testdata", ps, psd, 1, events, [tf], {})


def case0_2bus(tf=10.0):
    rng = np.random.default_rng(1000)
    meta = {"seed": 1000, "ieesgo_K2": 0.5, "ieesgo_K3": 0.5, "ieesgo_pmin": 0.0}
    buses = [Bus(1, "BUS1", SLACK, 100.0, 1.0, 0.0), Bus(2, "BUS2", PQ, 100.0, 1.0, 0.0)]
    gens = [Gen(0, "1", 0.8, 0.2, 100.0)]
    loads = [Load(1, "1", 0.8, -0.2)]
    branches = [Branch(0, 1, "1", 0.01, 0.1, 0.0)]
    ps = PowerSystem(100.0, buses, gens, loads, branches, [])
    psd = PowerSystemDynamics()
    psd.add_device(_genrou(1))
    psd.add_device(_ieesgo(1, "1", rng, meta))
    events = [[ContingencyEvent(1, 0.2, 0.2, 0.3)]]
    return _finish("c0_2bus", ps, psd, 1, events, [tf], meta)


def case1_9bus(M=64, tf=1.0, gensal=False):
    """gensal: the machine at bus 4 is a GENSAL (the spec's machine data with
    the GENSAL substitutions) instead of a GENROU."""
    rng = np.random.default_rng(1001)
    meta = {"seed": 1001, "zip_alpha": 0.4,
            "zip_note": "Z/I/P 0.4/0.3/0.3 mapped to alpha=0.4 constant-Z, 0.6 constant-P"}
    nb = 9
    buses = [Bus(i + 1, f"B{i+1}", PQ, 230.0, 1.0, 0.0) for i in range(nb)]
    edges = [(i, (i + 1) % nb) for i in range(nb)]
    have = {tuple(sorted(e)) for e in edges}
    while len(edges) < nb + 3:
        a, b = rng.choice(nb, 2, replace=False)
        e = tuple(sorted((int(a), int(b))))
        if e not in have:
            have.add(e)
            edges.append(e)
    branches = []
    for (a, b) in edges:
        x = rng.uniform(0.05, 0.2)
        r = x / rng.uniform(5.0, 15.0)
        branches.append(Branch(a, b, "1", r, x, rng.uniform(0.0, 0.2)))
    gen_buses = [0, 3, 6]
    loads = []
    for i in range(nb):
        if i in gen_buses:
            continue
        P = rng.uniform(0.5, 1.5)
        loads.append(Load(i, "1", P, -0.3 * P))
    total = sum(l.pd for l in loads)
    gens = []
    for k, b in enumerate(gen_buses):
        buses[b].type = SLACK if k == 0 else PV
        buses[b].v0m = 1.02
        gens.append(Gen(b, "1", total / 3.0, 0.0, 300.0))
    ps = PowerSystem(100.0, buses, gens, loads, branches, [])
    psd = PowerSystemDynamics()
    for b in gen_buses:
        if gensal and b == 3:
            g = C0_GENROU
            psd.add_device(Gensal(b + 1, "1", g["x_d"], g["x_q"], g["x_dp"], g["x_ddp"], g["xl"], g["H"], g["D"],
                                  g["T_d0p"], g["T_d0dp"], g["T_q0dp"], g["S1"], g["S2"]))
        else:
            psd.add_device(_genrou(b + 1))
        psd.add_device(_sexs(b + 1))
        psd.add_device(_tgov1(b + 1))
    events = []
    for m in range(M):
        events.append([ContingencyEvent(m % nb, float(rng.uniform(0.01, 0.2)), 0.02,
                                        float(rng.uniform(0.1, 0.2)))])
    if gensal:
        meta["gensal_bus"] = 4
    return _finish("c1_9bus_gensal" if gensal else "c1_9bus", ps, psd, M, events, [tf] * M, meta, zip_alpha=0.4)


def _rgg_edges(rng, n, mean_degree):
    from scipy.spatial import cKDTree
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import minimum_spanning_tree
    pts = rng.random((n, 2))
    tree = cKDTree(pts)
    k = 6
    dd, ii = tree.query(pts, k + 1)
    G = coo_matrix((dd[:, 1:].ravel(), (np.repeat(np.arange(n), k), ii[:, 1:].ravel())),
                   shape=(n, n)).tocsr()
    G = G.maximum(G.T)
    T = minimum_spanning_tree(G).tocoo()
    mst = {tuple(sorted((int(a), int(b)))) for a, b in zip(T.row, T.col)}
    need = max(0, int(round(mean_degree * n / 2)) - len(mst))
    r = math.sqrt(mean_degree / (n * math.pi))
    pairs = tree.query_pairs(r, output_type="ndarray")
    cand = [tuple(sorted((int(a), int(b)))) for a, b in pairs]
    cand = [e for e in cand if e not in mst]
    d = np.array([np.linalg.norm(pts[a] - pts[b]) for a, b in cand]) if cand else np.zeros(0)
    order = np.argsort(d)
    extra = [cand[i] for i in order[:need]]
    return sorted(mst | set(extra)), pts


def _spread_buses(rng, edges, n, k):
    """Pick k buses spread over the graph: greedy farthest-first in hop
    distance from a random start, so every load bus has generation a few hops
    away (random placement leaves long generator-free corridors on sparse
    random geometric graphs, where the power flow has no solution)."""
    from collections import deque
    adj = [[] for _ in range(n)]
    for a, b in edges:
        adj[a].append(b)
        adj[b].append(a)
    # hop distance to the nearest chosen bus, updated by a breadth-first
    # search from each new choice that stops where the distance does not
    # fall (the same distances as a full search from every chosen bus)
    dist = np.full(n, np.inf)

    def lower(src):
        dist[src] = 0.0
        dq = deque([src])
        while dq:
            u = dq.popleft()
            du = dist[u] + 1.0
            for v in adj[u]:
                if du < dist[v]:
                    dist[v] = du
                    dq.append(v)
    chosen = [int(rng.integers(n))]
    lower(chosen[0])
    while len(chosen) < k:
        far = np.flatnonzero(dist == dist.max())
        nxt = int(rng.choice(far))
        chosen.append(nxt)
        lower(nxt)
    return np.array(chosen)


def _regional_dispatch(edges, n, gen_buses, loads, floor=0.1):
    """Scheduled generation that serves nearby load: each load is assigned to
    the generator bus nearest in hops (multi-source BFS) and a generator's
    output is its assigned load (at least `floor`), rescaled so the total
    matches the total load.  Flows then stay regional, as in planned grids."""
    from collections import deque
    adj = [[] for _ in range(n)]
    for a, b in edges:
        adj[a].append(b)
        adj[b].append(a)
    owner = np.full(n, -1)
    dq = deque()
    for k, g in enumerate(gen_buses):
        owner[g] = k
        dq.append(g)
    while dq:
        u = dq.popleft()
        for v in adj[u]:
            if owner[v] < 0:
                owner[v] = owner[u]
                dq.append(v)
    share = np.zeros(len(gen_buses))
    for ld in loads:
        share[owner[ld.bus]] += ld.pd
    share = np.maximum(share, floor)
    total = sum(ld.pd for ld in loads)
    return share * (total / share.sum())


def case2_grid(M=256, nbus=2000, n_genrou=400, n_pss=100, n_static=300, seed=1002,
               load_fraction=0.2, tf_range=(1.0, 5.0), name="c2_grid", charging_scale=1.0,
               zip_alpha=1.0, static_gen_type=PV, static_gen_p=1.0, fault_sg_hops=3, size_machines=True,
               pf_max_step=None):
    rng = np.random.default_rng(seed)
    meta = {"seed": seed, "zip_alpha": zip_alpha, "nbus": nbus,
            "zip_note": "constant-impedance loads (alpha=1, the reference default of "
                        "src/GradPower.jl:382): with a constant-power share, bolted faults on this "
                        "weak synthetic grid leave ~10% of the faulted algebraic systems without "
                        "a Newton-reachable solution, in the oracle as on the GPU",
            "static_gen_note": "static generators regulate voltage in the power flow, then become "
                               "constant P/Q injections (PQ stubs) at that operating point: a PV stub "
                               "holds |v| with unlimited reactive power, and faults 2-3 buses away "
                               "then defeat the undamped Newton re-solve (oracle and GPU alike)"}
    edges, pts = _rgg_edges(rng, nbus, 2.6)
    meta["mean_degree"] = 2 * len(edges) / nbus
    meta["charging_scale"] = charging_scale
    buses = [Bus(i + 1, f"B{i+1}", PQ, 230.0, 1.0, 0.0) for i in range(nbus)]
    branches = []
    for (a, b) in edges:
        x = rng.uniform(0.05, 0.2)
        r = x / rng.uniform(5.0, 15.0)
        branches.append(Branch(a, b, "1", r, x, charging_scale * rng.uniform(0.0, 0.2)))
    spread = _spread_buses(rng, edges, nbus, n_genrou + n_static)
    spread = spread[rng.permutation(len(spread))]
    gbus = spread[:n_genrou]
    sbus = spread[n_genrou:]
    rest = np.setdiff1d(np.arange(nbus), spread)
    lbus = rest[rng.random(len(rest)) < load_fraction / max(1e-9, len(rest) / nbus)]
    loads = []
    for b in sorted(int(x) for x in lbus):
        P = rng.uniform(0.5, 1.5)
        loads.append(Load(b, "1", P, -0.3 * P))
    allg = [int(b) for b in gbus] + [int(b) for b in sbus]
    pg = _regional_dispatch(edges, nbus, allg, loads)
    pg[n_genrou:] *= static_gen_p
    gens = []
    for k, b in enumerate(allg):
        if k < n_genrou:
            buses[b].type = SLACK if k == 0 else PV
            buses[b].v0m = 1.02
        else:
            buses[b].type = PV
            buses[b].v0m = 1.01
        gens.append(Gen(b, "1", float(pg[k]), 0.0, max(100.0, 100.0 * pg[k] / 0.7)))
    ps = PowerSystem(100.0, buses, gens, loads, branches, [])
    psd = PowerSystemDynamics()
    exc_kind = rng.random(n_genrou) < 0.5
    gov_kind = rng.random(n_genrou) < 0.5
    pss_on = np.zeros(n_genrou, bool)
    pss_on[rng.choice(n_genrou, n_pss, replace=False)] = True
    H = rng.uniform(2.0, 8.0, n_genrou)
    D = rng.uniform(0.0, 2.0, n_genrou)
    for k, b in enumerate(gbus):
        eb = int(b) + 1
        psd.add_device(_genrou(eb, H=float(H[k]), D=float(D[k])))
        psd.add_device(_ieesgo(eb, "1", rng, meta) if gov_kind[k] else _tgov1(eb))
        psd.add_device(_sexs(eb) if exc_kind[k] else _esdc1a(eb))
        if pss_on[k]:
            psd.add_device(_ieeest(eb))
    meta.pop("ieesgo_T", None)
    # Fault buses are drawn uniformly among buses more than fault_sg_hops - 1
    # hops from every static generator: a PV stub holds |v| = vset with
    # unlimited reactive power, and a resistive fault at, next to or two hops
    # from one leaves the faulted algebraic system without a Newton-reachable
    # solution in about 8% of draws (the reference Newton fails there too).
    adj = [[] for _ in range(nbus)]
    for (a, b) in edges:
        adj[a].append(b)
        adj[b].append(a)
    excl = set(int(b) for b in sbus)
    for _ in range(fault_sg_hops - 1):
        for b in list(excl):
            excl.update(adj[b])
    fault_buses = np.array([b for b in range(nbus) if b not in excl])
    meta["fault_bus_pool"] = int(len(fault_buses))
    events, tf = [], []
    for m in range(M):
        events.append([ContingencyEvent(int(fault_buses[rng.integers(len(fault_buses))]),
                                        float(rng.uniform(0.01, 0.2)), 0.02, float(rng.uniform(0.05, 0.25)))])
        tf.append(float(rng.uniform(*tf_range)))
    return _finish(name, ps, psd, M, events, tf, meta, zip_alpha=zip_alpha,
                   pq_after_pf=[int(b) for b in sbus] if static_gen_type == PQ else None,
                   size_machines=size_machines, pf_max_step=pf_max_step)


def case4_grid(M=8, tiles=(7, 5), tile_bus=2000, tile_genrou=343, tile_static=260, tile_pss=86,
               gensal_frac=0.25, ties=4, tf=2.0, seed=1004, name="c4_grid"):
    """The 70,000-bus workload: 35 tiles of the 2,000-bus generator's grid
    (random geometric graph, mean degree 2.6, the same line, load, machine and
    controller draws) on a 7 x 5 lattice, neighbouring tiles joined by `ties`
    lines between their closest buses.  12,005 machines, a quarter GENSAL,
    with SEXS/ESDC1A exciters, TGOV1/IEESGO governors and IEEEST stabilisers
    on a quarter; 9,100 static generator stubs.  One slack machine per tile
    (a single slack for 70,000 buses leaves the angle block of the power-flow
    Jacobian nearly singular, and flat-start Newton diverges even at zero
    load).  Faults as in the 2,000-bus case: bus uniform over the buses at
    least 3 hops from every static generator, r ~ U(0.01, 0.2), on at 0.02 s,
    cleared at U(0.05, 0.25) s; M scenarios over `tf` seconds."""
    rng = np.random.default_rng(seed)
    tx, ty = tiles
    nt = tx * ty
    nbus = nt * tile_bus
    meta = {"seed": seed, "zip_alpha": 1.0, "nbus": nbus, "tiles": [tx, ty], "ties_per_neighbour": ties}
    buses = [Bus(i + 1, f"B{i+1}", PQ, 230.0, 1.0, 0.0) for i in range(nbus)]
    branches, loads, gens, edges_all = [], [], [], []
    psd = PowerSystemDynamics()
    gbus_all, sbus_all, pts_all = [], [], np.zeros((nbus, 2))
    for t in range(nt):
        o = t * tile_bus
        edges, pts = _rgg_edges(rng, tile_bus, 2.6)
        pts_all[o:o + tile_bus] = pts + np.array([t % tx, t // tx], float)
        for (a, b) in edges:
            x = rng.uniform(0.05, 0.2)
            branches.append(Branch(o + a, o + b, "1", x / rng.uniform(5.0, 15.0), x, rng.uniform(0.0, 0.2)))
            edges_all.append((o + a, o + b))
        spread = _spread_buses(rng, edges, tile_bus, tile_genrou + tile_static)
        spread = spread[rng.permutation(len(spread))]
        gb, sb = spread[:tile_genrou], spread[tile_genrou:]
        rest = np.setdiff1d(np.arange(tile_bus), spread)
        lb = rest[rng.random(len(rest)) < 0.2 / max(1e-9, len(rest) / tile_bus)]
        tl = []
        for b in sorted(int(x) for x in lb):
            P = rng.uniform(0.5, 1.5)
            tl.append(Load(o + b, "1", P, -0.3 * P))
        loads += tl
        allg = [int(b) for b in gb] + [int(b) for b in sb]
        pg = _regional_dispatch(edges, tile_bus, allg, [Load(ld.bus - o, ld.id, ld.pd, ld.qd) for ld in tl])
        for k, b in enumerate(allg):
            bb = o + b
            if k < tile_genrou:
                buses[bb].type = SLACK if k == 0 else PV
                buses[bb].v0m = 1.02
            else:
                buses[bb].type = PV
                buses[bb].v0m = 1.01
            gens.append(Gen(bb, "1", float(pg[k]), 0.0, max(100.0, 100.0 * pg[k] / 0.7)))
        gbus_all += [o + int(b) for b in gb]
        sbus_all += [o + int(b) for b in sb]
        n_m = tile_genrou
        exc_kind = rng.random(n_m) < 0.5
        gov_kind = rng.random(n_m) < 0.5
        sal = rng.random(n_m) < gensal_frac
        pss_on = np.zeros(n_m, bool)
        pss_on[rng.choice(n_m, tile_pss, replace=False)] = True
        H = rng.uniform(2.0, 8.0, n_m)
        D = rng.uniform(0.0, 2.0, n_m)
        for k, b in enumerate(gb):
            eb = o + int(b) + 1
            if sal[k]:
                g = C0_GENROU
                psd.add_device(Gensal(eb, "1", g["x_d"], g["x_q"], g["x_dp"], g["x_ddp"], g["xl"], float(H[k]),
                                      float(D[k]), g["T_d0p"], g["T_d0dp"], g["T_q0dp"], g["S1"], g["S2"]))
            else:
                psd.add_device(_genrou(eb, H=float(H[k]), D=float(D[k])))
            psd.add_device(_ieesgo(eb, "1", rng, meta) if gov_kind[k] else _tgov1(eb))
            psd.add_device(_sexs(eb) if exc_kind[k] else _esdc1a(eb))
            if pss_on[k]:
                psd.add_device(_ieeest(eb))
    meta.pop("ieesgo_T", None)
    # tie lines between neighbouring tiles: the closest bus pairs across each boundary
    from scipy.spatial import cKDTree
    for t in range(nt):
        i, j = t % tx, t // tx
        for u in ([t + 1] if i + 1 < tx else []) + ([t + tx] if j + 1 < ty else []):
            A = pts_all[t * tile_bus:(t + 1) * tile_bus]
            B = pts_all[u * tile_bus:(u + 1) * tile_bus]
            d, k = cKDTree(B).query(A)
            used = set()
            for a in np.argsort(d):
                if len(used) >= ties:
                    break
                if int(k[a]) in used:
                    continue
                used.add(int(k[a]))
                x = rng.uniform(0.05, 0.2)
                fa, fb = t * tile_bus + int(a), u * tile_bus + int(k[a])
                branches.append(Branch(fa, fb, "1", x / rng.uniform(5.0, 15.0), x, rng.uniform(0.0, 0.2)))
                edges_all.append((fa, fb))
    meta["mean_degree"] = 2 * len(edges_all) / nbus
    meta["n_machines"] = nt * tile_genrou
    meta["n_static"] = nt * tile_static
    ps = PowerSystem(100.0, buses, gens, loads, branches, [])
    adj = [[] for _ in range(nbus)]
    for (a, b) in edges_all:
        adj[a].append(b)
        adj[b].append(a)
    excl = set(sbus_all)
    for _ in range(2):
        for b in list(excl):
            excl.update(adj[b])
    fault_buses = np.array([b for b in range(nbus) if b not in excl])
    meta["fault_bus_pool"] = int(len(fault_buses))
    events = [[ContingencyEvent(int(fault_buses[rng.integers(len(fault_buses))]), float(rng.uniform(0.01, 0.2)),
                                0.02, float(rng.uniform(0.05, 0.25)))] for _ in range(M)]
    # power-flow continuation from zero load in steps of 1/4 (halved on
    # failure, down to 1/32), at most 25 Newton iterations per attempt
    return _finish(name, ps, psd, M, events, [tf] * M, meta, zip_alpha=1.0, size_machines=True,
                   pf_kw={"maxit": 25, "first": 0.0, "min_step": 1.0 / 32})


def case3_200bus(M=128, tf=3.0):
    rng = np.random.default_rng(1003)
    meta = {"seed": 1003, "zip_alpha": 0.4}
    nbus, ng = 200, 40
    edges, _ = _rgg_edges(rng, nbus, 2.6)
    buses = [Bus(i + 1, f"B{i+1}", PQ, 230.0, 1.0, 0.0) for i in range(nbus)]
    branches = []
    for (a, b) in edges:
        x = rng.uniform(0.05, 0.2)
        branches.append(Branch(a, b, "1", x / rng.uniform(5.0, 15.0), x, rng.uniform(0.0, 0.2)))
    gbus = _spread_buses(rng, edges, nbus, ng)
    loads = []
    for b in sorted(int(x) for x in np.setdiff1d(np.arange(nbus), gbus) if rng.random() < 0.25):
        P = rng.uniform(0.5, 1.5)
        loads.append(Load(b, "1", P, -0.3 * P))
    pg = _regional_dispatch(edges, nbus, [int(b) for b in gbus], loads)
    gens = []
    for k, b in enumerate(gbus):
        b = int(b)
        buses[b].type = SLACK if k == 0 else PV
        buses[b].v0m = 1.02
        gens.append(Gen(b, "1", float(pg[k]), 0.0, max(100.0, 100.0 * pg[k] / 0.7)))
    ps = PowerSystem(100.0, buses, gens, loads, branches, [])
    psd = PowerSystemDynamics()
    for b in gbus:
        eb = int(b) + 1
        psd.add_device(_genrou(eb, H=float(rng.uniform(2, 8)), D=float(rng.uniform(0, 2))))
        psd.add_device(_sexs(eb))
        psd.add_device(_tgov1(eb))
    events = [[ContingencyEvent(int(rng.integers(nbus)), float(rng.uniform(0.01, 0.2)), 0.02,
                                float(rng.uniform(0.05, 0.25)))] for _ in range(M)]
    return _finish("c3_200bus", ps, psd, M, events, [tf] * M, meta, zip_alpha=0.4)


def case_psse(raw, dyr, M=16, tf=1.0, exact_slack=True, zip_alpha=None, seed=1004, xf_tap=None):
    """A system read from PSS/E files (raw + dyr) with M fault scenarios:
    scenario m puts a shunt fault on bus m mod nbus (internal order) from
    0.05 s to a clearing time drawn in [0.1, 0.2] s, fault resistance drawn
    in [0.005, 0.05] pu.  exact_slack: see network.runpf_inplace (the
    bundled IEEE 39-bus case has a load on its slack bus).  zip_alpha, when
    given, is the constant-impedance fraction of every load.  xf_tap =
    (k, WINDV1, ANG1) sets the winding-1 ratio and phase shift (degrees) of
    the k-th two-winding transformer before the conversion."""
    import os
    from .psse import from_psse, read_psse_raw
    rng = np.random.default_rng(seed)
    rd = read_psse_raw(raw)
    if xf_tap is not None:
        k, w1, a1 = xf_tap
        rd.transformers[k].WINDV1 = float(w1)
        rd.transformers[k].ANG1 = float(a1)
    ps = from_psse(rd, dyr)
    if zip_alpha is not None:
        for dev in ps.dynamic.devices:
            if dev.dtype.kind == "zipload":
                dev.dtype.alpha = zip_alpha
    build_network(ps)
    runpf_inplace(ps, exact_slack=exact_slack)
    dp = DynamicProblem(ps)
    initialize_dynamics(dp, ps)
    desc = build_desc(ps, dp)
    nb = ps.nbus
    events = [[ContingencyEvent(m % nb, float(rng.uniform(0.005, 0.05)), 0.05, float(rng.uniform(0.1, 0.2)))]
              for m in range(M)]
    name = os.path.splitext(os.path.basename(dyr))[0]
    meta = {"seed": seed, "raw": os.path.basename(raw), "dyr": os.path.basename(dyr),
            "exact_slack": exact_slack, "zip_alpha": zip_alpha, "xf_tap": xf_tap}
    return Case(name, ps, dp, desc, M, DT, events, np.full(M, float(tf)), meta)
