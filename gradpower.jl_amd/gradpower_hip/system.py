"""Host-side power-system model: buses, branches, devices, wiring, clusters.

This module mirrors the reference's setup layer so that a user of
GradPower.jl finds the same objects and the same call order:

    ps = PowerSystem(...)              # src/GradPower.jl:197-218
    psd = PowerSystemDynamics(); add_device(psd, Genrou(...))
    set_dynamics(ps, psd)              # src/GradPower.jl:297-447
    build_network(ps)                  # src/network.jl:1-13
    runpf(ps)                          # src/pflow.jl:375-418  (powerflow.py)
    dp = DynamicProblem(ps); initialize_dynamics(dp, ps)   # src/dynamics.jl:123-250

All indices here are 0-based.  After `set_dynamics` the state vector is in
the reference's cluster-contiguous order z = [w_1 | ... | w_Nc | v]
(src/clusters.jl:225-319): every device's `diff_ptr`/`alg_ptr` is an absolute
0-based z index.
"""
from __future__ import annotations

import cmath
import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

TWOPI60 = 2.0 * math.pi * 60.0

# Bus types (PSS/E): 1 = PQ, 2 = PV, 3 = slack.
PQ, PV, SLACK = 1, 2, 3


# ----------------------------------------------------------------------------
# Static network data (src/GradPower.jl:22-69)
# ----------------------------------------------------------------------------
@dataclass
class Bus:
    i: int              # external bus number
    name: str
    type: int
    baseKV: float
    v0m: float
    v0a: float          # radians


@dataclass
class Gen:
    bus: int            # internal 0-based bus index
    id: str
    psch: float         # pu on system base
    qsch: float
    mbase: float
    status: int = 1


@dataclass
class Load:
    bus: int
    id: str
    pd: float           # consumed active power, pu
    qd: float           # reactive *injection* convention: qd = -Q_load (src/parse.jl:384)


@dataclass
class Branch:
    fr: int
    to: int
    ckt: str
    r: float
    x: float
    sh: float           # total charging susceptance
    tap: float = 0.0
    shift: float = 0.0  # degrees


@dataclass
class Shunt:
    bus: int
    name: str
    gsh: float
    bsh: float


@dataclass
class ContingencyEvent:
    """Three-phase shunt fault (src/GradPower.jl:117-128): conductance 1/rfault
    on the vr and vi rows of `bus` (0-based) between ton and toff."""
    bus: int
    rfault: float
    ton: float
    toff: float
    status: bool = False


@dataclass
class TripLineEvent:
    """Trip of one line at ton, reclosed at toff (inf: never), with its
    admittance stamp (src/GradPower.jl:140-167).  Buses are 0-based."""
    from_bus: int
    to_bus: int
    ton: float
    toff: float
    yff: complex
    yft: complex
    ytf: complex
    ytt: complex

    def stamp(self):
        """The 8 reals of the C ABI: yff, yft, ytf, ytt as (re, im)."""
        return [v for y in (self.yff, self.yft, self.ytf, self.ytt) for v in (y.real, y.imag)]

    def steps(self, dt):
        """(step_trip, step_reclose) at round(t/dt); 0 = never reclosed."""
        off = 0 if math.isinf(self.toff) else int(round(self.toff / dt))
        return int(round(self.ton / dt)), off


@dataclass
class DisconnectDeviceEvent:
    """Device `device_idx` (0-based index into ps.dynamic.devices) taken
    offline for good at ton (src/GradPower.jl:130-143).  The HIP library
    takes ZIP loads and SEXS / ESDC1A exciters (target gives the device
    type and its table row)."""
    device_idx: int
    ton: float

    DROPPABLE = {"zipload": 6, "sexs": 3, "esdc1a": 4}   # gp_device_type of each kind

    def target(self, ps):
        devs = ps.dynamic.devices
        kind = devs[self.device_idx].dtype.kind
        if kind not in self.DROPPABLE:
            raise NotImplementedError(f"a {kind} cannot be disconnected: a device with algebraic states "
                                      "left out of F and J leaves them undetermined")
        return self.DROPPABLE[kind], sum(1 for d in devs[:self.device_idx] if d.dtype.kind == kind)

    def zip_row(self, ps):
        t, row = self.target(ps)
        if t != self.DROPPABLE["zipload"]:
            raise NotImplementedError("not a ZIP load")
        return row


def create_trip_line_event(ps, from_bus_ext, to_bus_ext, ton, toff=math.inf):
    """The first branch between the two buses (either orientation) and the
    admittance it adds to Y, by the same formulas as the Y builder
    (src/GradPower.jl:509-542)."""
    a, b = ps.busmap[from_bus_ext], ps.busmap[to_bus_ext]
    br = next((x for x in ps.branches if (x.fr, x.to) in ((a, b), (b, a))), None)
    if br is None:
        raise ValueError(f"No branch found between buses {from_bus_ext} and {to_bus_ext}")
    tap, shift = br.tap, br.shift
    if tap > 0.0:
        tpsh = tap * cmath.exp(1j * math.pi / 180.0 * shift)
    else:
        tpsh, tap = 1.0 + 0j, 1.0
    y = 1.0 / complex(br.r, br.x)
    return TripLineEvent(br.fr, br.to, ton, toff,
                         y / (tap * tap) + 1j * 0.5 * br.sh,
                         -y / tpsh.conjugate(),
                         -y / tpsh,
                         y + (1j * 0.5 * br.sh) / (tap * tap))


# ----------------------------------------------------------------------------
# Device models.  Sizes and parameter-vector layouts follow
# src/generators.jl:31-34,131-146, src/governors.jl:28-33,163-165,
# src/exciters.jl:26-29,333-337, src/stabilizers.jl:93-99,
# src/loads.jl:22-25, src/static_gen.jl:36-42.
# ----------------------------------------------------------------------------
class Device:
    kind = "device"
    diff_size = 0
    alg_size = 0
    ctrl_size = 0
    par_size = 0

    def fill_pvec(self, p):
        raise NotImplementedError


@dataclass
class Genrou(Device):
    bus: int            # external bus number until set_dynamics remaps it
    id: str
    x_d: float
    x_q: float
    x_dp: float
    x_qp: float
    x_ddp: float
    xl: float
    H: float
    D: float
    T_d0p: float
    T_q0p: float
    T_d0dp: float
    T_q0dp: float
    S1: float = 0.0
    S2: float = 0.0
    e_fd0: float = 0.0
    kind = "genrou"
    diff_size, alg_size, ctrl_size, par_size = 6, 4, 2, 14

    def set_ratio(self, ratio):
        # src/generators.jl:111-120
        self.x_d /= ratio; self.x_q /= ratio; self.x_dp /= ratio
        self.x_qp /= ratio; self.x_ddp /= ratio; self.xl /= ratio
        self.H *= ratio; self.D *= ratio

    def fill_pvec(self, p):
        p[:] = [self.x_d, self.x_q, self.x_dp, self.x_qp, self.x_ddp, self.xl,
                self.H, self.D, self.T_d0p, self.T_q0p, self.T_d0dp, self.T_q0dp,
                self.S1, self.S2]


def Gensal(bus, id, x_d, x_q, x_dp, x_ddp, xl, H, D, T_d0p, T_d0dp, T_q0dp,
           S1=0.0, S2=0.0):
    """GENSAL is GENROU with x_qp = x_dp and T_q0p = T_d0p (src/generators.jl:47-66)."""
    return Genrou(bus, id, x_d, x_q, x_dp, x_dp, x_ddp, xl, H, D, T_d0p, T_d0p,
                  T_d0dp, T_q0dp, S1, S2)


@dataclass
class IEESGO(Device):
    bus: int
    id: str
    T1: float
    T2: float
    T3: float
    T4: float
    T5: float
    T6: float
    K1: float
    K2: float
    K3: float
    pmax: float
    pmin: float
    pref: float = 0.0
    kind = "ieesgo"
    diff_size, alg_size, ctrl_size, par_size = 5, 1, 1, 12

    def set_ratio(self, ratio):
        pass

    def fill_pvec(self, p):
        p[:] = [self.T1, self.T2, self.T3, self.T4, self.T5, self.T6,
                self.K1, self.K2, self.K3, self.pmax, self.pmin, self.pref]


@dataclass
class TGOV1(Device):
    bus: int
    id: str
    R: float
    T1: float
    VMAX: float
    VMIN: float
    T2: float
    T3: float
    DT: float
    pref: float = 0.0
    kind = "tgov1"
    diff_size, alg_size, ctrl_size, par_size = 2, 1, 1, 8

    def set_ratio(self, ratio):
        # src/governors.jl:194-199
        self.R /= ratio; self.VMAX /= ratio; self.VMIN /= ratio; self.DT *= ratio

    def fill_pvec(self, p):
        p[:] = [self.R, self.T1, self.VMAX, self.VMIN, self.T2, self.T3, self.DT, self.pref]


@dataclass
class SEXS(Device):
    bus: int
    id: str
    TA_TB: float
    TB: float
    K: float
    TE: float
    EMIN: float
    EMAX: float
    vref: float = 0.0
    kind = "sexs"
    diff_size, alg_size, ctrl_size, par_size = 2, 0, 1, 7

    def set_ratio(self, ratio):
        pass

    def fill_pvec(self, p):
        p[:] = [self.TA_TB, self.TB, self.K, self.TE, self.EMIN, self.EMAX, self.vref]


def esdc1a_sat_coefficients(E1, SE1, E2, SE2):
    """Quadratic exciter saturation fit (src/exciters.jl:40-52)."""
    if E1 <= 0.0 or E2 <= 0.0 or SE1 <= 0.0 or SE2 <= 0.0:
        return 0.0, 0.0
    a = math.sqrt(SE1 * E1 / (SE2 * E2))
    if a == 1.0:
        return 0.0, 0.0
    sat_a = E2 - (E1 - E2) / (a - 1.0)
    sat_b = SE2 * E2 * (a - 1.0) ** 2 / (E1 - E2) ** 2
    return sat_a, sat_b


@dataclass
class ESDC1A(Device):
    bus: int
    id: str
    Ka: float
    Ta: float
    Kf: float
    Tf: float
    Ke: float
    Te: float
    Tr: float
    Ae: float           # sat_a
    Be: float           # sat_b
    vref: float = 0.0
    kind = "esdc1a"
    diff_size, alg_size, ctrl_size, par_size = 3, 0, 0, 10

    def set_ratio(self, ratio):
        pass

    def fill_pvec(self, p):
        p[:] = [self.Ka, self.Ta, self.Kf, self.Tf, self.Ke, self.Te, self.Tr,
                self.Ae, self.Be, self.vref]


@dataclass
class IEEEST(Device):
    bus: int
    id: str
    MODE: int
    BUSR: int
    A1: float
    A2: float
    A3: float
    A4: float
    A5: float
    A6: float
    T1: float
    T2: float
    T3: float
    T4: float
    T5: float
    T6: float
    KS: float
    LSMAX: float = 0.0
    LSMIN: float = 0.0
    VCU: float = 0.0
    VCL: float = 0.0
    kind = "ieeest"
    diff_size, alg_size, ctrl_size, par_size = 7, 1, 1, 19

    def __post_init__(self):
        # MODE 1 only; zero time constants clamped to 1e-3 (src/stabilizers.jl:128-146)
        if self.MODE != 1:
            raise ValueError(f"IEEEST MODE={self.MODE} not supported; only MODE=1")
        for name in ("A1", "A2", "A3", "A4", "A5", "A6", "T1", "T2", "T3", "T4", "T5", "T6"):
            if getattr(self, name) == 0.0:
                setattr(self, name, 1e-3)

    def set_ratio(self, ratio):
        pass

    def fill_pvec(self, p):
        p[:] = [self.A1, self.A2, self.A3, self.A4, self.A5, self.A6, self.T1,
                self.T2, self.T3, self.T4, self.T5, self.T6, self.KS, self.LSMAX,
                self.LSMIN, self.VCU, self.VCL, float(self.MODE), float(self.BUSR)]


@dataclass
class ZIPLoad(Device):
    bus: int            # internal bus index (loads come from the static case)
    id: str
    pinj: float
    qinj: float
    alpha: float = 1.0
    beta: float = 0.0
    gamma: float = 0.0
    weight: float = 1.0
    v0mag: float = 1.0
    yreal: float = 0.0
    yimag: float = 0.0
    kind = "zipload"
    diff_size, alg_size, ctrl_size, par_size = 0, 0, 0, 9

    def fill_pvec(self, p):
        p[:] = [self.pinj, self.qinj, self.alpha, self.beta, self.gamma,
                self.weight, self.v0mag, self.yreal, self.yimag]


@dataclass
class StaticGenerator(Device):
    bus: int            # internal bus index
    bus_type: int
    gen_idxs: list
    vset: float
    aset: float
    p0: float = 0.0
    q0: float = 0.0
    kind = "static_gen"
    ctrl_size, par_size = 0, 4

    @property
    def alg_size(self):
        return 1 if self.bus_type == PV else (2 if self.bus_type == SLACK else 0)

    def fill_pvec(self, p):
        p[:] = [self.p0, self.q0, self.vset, self.aset]


CONTROLLER_TARGET = {"ieesgo": "genrou", "tgov1": "genrou", "sexs": "genrou",
                     "esdc1a": "genrou", "ieeest": "exciter"}
GOVERNORS = ("ieesgo", "tgov1")
EXCITERS = ("sexs", "esdc1a")


def _norm_id(s):
    return str(s).replace("'", "").replace(" ", "").strip()


@dataclass
class DynamicDevice:
    dtype: Device
    diff_ptr: int       # before reordering: offset in the diff block; after: absolute z index
    alg_ptr: int        # before reordering: offset in the alg block; after: absolute z index
    ctrl_ptr: int
    par_ptr: int
    online: bool = True


@dataclass
class Cluster:
    """One generator with its governor/exciter/PSS (src/clusters.jl:16-28)."""
    gen_idx: int
    gov_idx: int
    exc_idx: int
    pss_idx: int
    bus: int
    d_k: int
    a_k: int
    w_size: int
    w_start: int = 0
    trivial: bool = False


class PowerSystemDynamics:
    def __init__(self):
        self.devices: List[DynamicDevice] = []
        self.diff_dim = 0
        self.alg_dim = 0
        self.ctrl_dim = 0
        self.par_dim = 0
        self.uvec_idx: Optional[np.ndarray] = None
        self.events: List[ContingencyEvent] = []
        self.trip_events: list = []          # TripLineEvent
        self.disconnect_events: list = []    # DisconnectDeviceEvent
        self.clusters: List[Cluster] = []
        self.is_diff: Optional[np.ndarray] = None
        self.dev_bus: List[int] = []     # internal bus of each device
        self.dev_gen: List[int] = []     # static gen index backing each device (-1 none)
        self.reordered = False

    def add_device(self, dtype: Device):
        """src/GradPower.jl:271-282."""
        self.devices.append(DynamicDevice(dtype, self.diff_dim, self.alg_dim,
                                          self.ctrl_dim, self.par_dim))
        self.diff_dim += dtype.diff_size
        self.alg_dim += dtype.alg_size
        self.ctrl_dim += dtype.ctrl_size
        self.par_dim += dtype.par_size

    @property
    def num_devices(self):
        return len(self.devices)


def add_device(psd: PowerSystemDynamics, dtype: Device):
    psd.add_device(dtype)


class PowerSystem:
    def __init__(self, baseMVA, buses, gens, loads, branches, shunts):
        self.baseMVA = baseMVA
        self.buses: List[Bus] = buses
        self.gens: List[Gen] = gens
        self.loads: List[Load] = loads
        self.branches: List[Branch] = branches
        self.shunts: List[Shunt] = shunts
        self.busmap = {b.i: k for k, b in enumerate(buses)}
        self.network = None
        self.dynamic: Optional[PowerSystemDynamics] = None

    @property
    def nbus(self):
        return len(self.buses)


def add_event(ps: PowerSystem, event: ContingencyEvent):
    ps.dynamic.events.append(event)
    return event


def add_trip_event(ps: PowerSystem, event):
    """add_trip_event! (src/GradPower.jl:489-496)."""
    ps.dynamic.trip_events.append(event)
    return event


def add_disconnect_event(ps: PowerSystem, event):
    """add_disconnect_event! (src/GradPower.jl:477-484)."""
    ps.dynamic.disconnect_events.append(event)
    return event


# ----------------------------------------------------------------------------
# set_dynamics (src/GradPower.jl:297-447)
# ----------------------------------------------------------------------------
def _find_gen(ps, bus_internal, gen_id, index=None):
    """First generator at (internal bus, id); index: a prebuilt (bus, id) map."""
    gid = _norm_id(gen_id)
    if index is not None:
        return index.get((bus_internal, gid), -1)
    for i, g in enumerate(ps.gens):
        if g.bus == bus_internal and _norm_id(g.id) == gid:
            return i
    return -1


def _gen_index(ps):
    idx = {}
    for i, g in enumerate(ps.gens):
        idx.setdefault((g.bus, _norm_id(g.id)), i)
    return idx


def set_dynamics(ps: PowerSystem, psd: PowerSystemDynamics, add_loads=True,
                 add_static_gen_stubs=True):
    mvbase = ps.baseMVA
    dev_bus = []
    dev_gen = []
    matched = []
    gidx = _gen_index(ps)
    for dev in psd.devices:
        d = dev.dtype
        b = ps.busmap[d.bus]
        dev_bus.append(b)
        g = -1
        if d.kind == "genrou":
            g = _find_gen(ps, b, d.id, gidx)
            if g >= 0:
                d.set_ratio(ps.gens[g].mbase / mvbase)
                matched.append(g)
        dev_gen.append(g)

    # Static-generator stubs: one per bus over active gens without a machine
    # model (src/GradPower.jl:347-375).  Buses in ascending internal order.
    if add_static_gen_stubs and len(matched) != len(ps.gens):
        mset = set(matched)
        by_bus = {}
        for i, g in enumerate(ps.gens):
            if i in mset:
                continue
            by_bus.setdefault(g.bus, []).append(i)
        for b in sorted(by_bus):
            bt = ps.buses[b].type
            sg = StaticGenerator(b, bt, by_bus[b], ps.buses[b].v0m, ps.buses[b].v0a)
            psd.add_device(sg)
            dev_bus.append(b)
            dev_gen.append(by_bus[b][0])

    if add_loads:
        for i, ld in enumerate(ps.loads):
            psd.add_device(ZIPLoad(ld.bus, ld.id, ld.pd, ld.qd, 1.0, 0.0, 0.0, 1.0,
                                   ps.buses[ld.bus].v0m, 0.0, 0.0))
            dev_bus.append(ld.bus)
            dev_gen.append(-1)

    psd.dev_bus = dev_bus
    psd.dev_gen = dev_gen
    psd.uvec_idx = np.full(psd.ctrl_dim, -1, dtype=np.int64)
    psd.target_index = _target_index(psd)
    _wire_controls(psd)

    for i, dev in enumerate(psd.devices):
        d = dev.dtype
        if d.kind in CONTROLLER_TARGET and psd.dev_gen[i] >= 0:
            d.set_ratio(ps.gens[psd.dev_gen[i]].mbase / mvbase)

    ps.dynamic = psd
    _build_clusters(psd)
    _reorder_state(psd)


def _target_index(psd):
    """(bus, id) -> [(device index, kind)] in device order, for the devices a
    controller can attach to (loads and static generator stubs excluded)."""
    idx = {}
    for j, c in enumerate(psd.devices):
        k = c.dtype.kind
        if k in ("zipload", "static_gen") or getattr(c.dtype, "bus", None) is None:
            continue
        idx.setdefault((c.dtype.bus, _norm_id(c.dtype.id)), []).append((j, k))
    return idx


def _find_target(psd, dev, kinds):
    """First device of a kind in `kinds` at the same (bus, id) as dev."""
    d = dev.dtype
    gid = _norm_id(d.id)
    idx = getattr(psd, "target_index", None)
    if idx is not None:
        for j, k in idx.get((d.bus, gid), ()):
            if k in kinds:
                return j
        return -1
    for j, c in enumerate(psd.devices):
        if c.dtype.kind in kinds and getattr(c.dtype, "bus", None) == d.bus \
                and c.dtype.kind not in ("zipload", "static_gen") and _norm_id(c.dtype.id) == gid:
            return j
    return -1


def _wire_controls(psd):
    """Trait-driven wiring (src/coupling.jl:154-256).  uvec_idx holds
    pre-reorder z indices (diff block first, alg block after diff_dim)."""
    dd = psd.diff_dim
    for i, dev in enumerate(psd.devices):
        k = dev.dtype.kind
        if k not in CONTROLLER_TARGET:
            continue
        if k == "ieeest":
            exc = _find_target(psd, dev, EXCITERS)
            if exc < 0:
                raise ValueError(f"IEEEST device {i}: no exciter at the same (bus, id)")
            gen = _find_target(psd, psd.devices[exc], ("genrou",))
            if gen < 0:
                raise ValueError(f"IEEEST device {i}: exciter has no generator")
            psd.dev_gen[i] = psd.dev_gen[gen]
            psd.uvec_idx[dev.ctrl_ptr] = psd.devices[gen].diff_ptr + 4
            continue
        tgt = _find_target(psd, dev, ("genrou",))
        if tgt < 0:
            raise ValueError(f"{k} device {i}: no GENROU at the same (bus, id)")
        target = psd.devices[tgt]
        psd.dev_gen[i] = psd.dev_gen[tgt]
        if k in GOVERNORS:
            psd.uvec_idx[target.ctrl_ptr + 1] = dd + dev.alg_ptr      # p_m <- governor alg
            psd.uvec_idx[dev.ctrl_ptr] = target.diff_ptr + 4          # w -> governor
        elif k == "sexs":
            psd.uvec_idx[target.ctrl_ptr] = dev.diff_ptr + 1          # e_fd
        elif k == "esdc1a":
            psd.uvec_idx[target.ctrl_ptr] = dev.diff_ptr + 2


def _build_clusters(psd):
    """src/clusters.jl:73-199: generator + governor/exciter/PSS, sorted by the
    type tuple so same-shape clusters are contiguous."""
    devs = psd.devices
    gov_map, exc_map, pss_map = {}, {}, {}
    for i, dev in enumerate(devs):
        k = dev.dtype.kind
        if k in GOVERNORS:
            gov_map[_find_target(psd, dev, ("genrou",))] = i
        elif k in EXCITERS:
            exc_map[_find_target(psd, dev, ("genrou",))] = i
        elif k == "ieeest":
            exc = _find_target(psd, dev, EXCITERS)
            pss_map[_find_target(psd, devs[exc], ("genrou",))] = i
    raw = []
    for gi, dev in enumerate(devs):
        if dev.dtype.kind not in ("genrou", "static_gen"):
            continue
        gov, exc, pss = gov_map.get(gi, -1), exc_map.get(gi, -1), pss_map.get(gi, -1)
        d_k = dev.dtype.diff_size
        a_k = dev.dtype.alg_size
        for c in (gov, exc, pss):
            if c >= 0:
                d_k += devs[c].dtype.diff_size
                a_k += devs[c].dtype.alg_size
        raw.append(Cluster(gi, gov, exc, pss, psd.dev_bus[gi], d_k, a_k, d_k + a_k,
                           trivial=dev.dtype.kind == "static_gen"))

    def ttuple(cl):
        t = [devs[cl.gen_idx].dtype.kind]
        for c in (cl.gov_idx, cl.exc_idx, cl.pss_idx):
            if c >= 0:
                t.append(devs[c].dtype.kind)
        return tuple(t)

    raw.sort(key=ttuple)   # stable, lexicographic like Julia's tuple isless
    psd.clusters = raw


def _reorder_state(psd):
    """Cluster-contiguous z order (src/clusters.jl:225-319).  After this call
    every device pointer is an absolute 0-based z index."""
    devs = psd.devices
    dd, ad = psd.diff_dim, psd.alg_dim
    n_da = dd + ad
    old_to_new = np.full(n_da, -1, dtype=np.int64)
    is_diff = np.zeros(n_da, dtype=bool)
    pos = 0
    for cl in psd.clusters:
        cl.w_start = pos
        for di in (cl.gen_idx, cl.gov_idx, cl.exc_idx, cl.pss_idx):
            if di < 0:
                continue
            dev = devs[di]
            for j in range(dev.dtype.diff_size):
                old_to_new[dev.diff_ptr + j] = pos
                is_diff[pos] = True
                pos += 1
            for j in range(dev.dtype.alg_size):
                old_to_new[dd + dev.alg_ptr + j] = pos
                pos += 1
        assert pos - cl.w_start == cl.w_size
    if pos != n_da:
        raise ValueError("devices without a cluster (orphan controllers) are not supported")
    for dev in devs:
        if dev.dtype.diff_size > 0:
            dev.diff_ptr = int(old_to_new[dev.diff_ptr])
        else:
            dev.diff_ptr = -1
        if dev.dtype.alg_size > 0:
            dev.alg_ptr = int(old_to_new[dd + dev.alg_ptr])
        else:
            dev.alg_ptr = -1
    u = psd.uvec_idx
    for i in range(len(u)):
        if 0 <= u[i] < n_da:
            u[i] = old_to_new[u[i]]
    psd.is_diff = is_diff
    psd.reordered = True


def net_ptr(ps):
    return ps.dynamic.diff_dim + ps.dynamic.alg_dim


def system_size(ps):
    return net_ptr(ps) + 2 * ps.nbus


class DynamicProblem:
    """z (states), u (controls), p (parameters) (src/GradPower.jl:449-455)."""

    def __init__(self, ps: PowerSystem):
        d = ps.dynamic
        assert d is not None and d.reordered, "call set_dynamics first"
        self.zvec = np.zeros(system_size(ps))
        self.uvec = np.zeros(d.ctrl_dim)
        self.pvec = np.zeros(d.par_dim)
