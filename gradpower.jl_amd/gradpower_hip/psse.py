"""PSS/E v33 reader: `.raw` static network data and `.dyr` dynamic models.

Same entry points and results as the reference's reader:

    raw = read_psse_raw("case.raw")            # src/parse_psse_raw.jl:570-584
    ps = raw_to_grad(raw)                      # src/parse.jl:212-408
    psd = dynamics_from_dyr("case.dyr")        # src/GradPower.jl:253-262
    set_dynamics(ps, psd)
    # or both at once, with inactive generators filtered:
    ps = from_psse("case.raw", "case.dyr")     # src/parse.jl:81-98

Bus indices in the returned PowerSystem are 0-based (the reference's are
1-based); device records keep the external PSS/E bus number until
`set_dynamics` maps it, as the reference's do.

Supported `.raw` sections: buses, loads, fixed shunts, generators, branches,
two- and three-winding transformers, switched shunts (initial susceptance);
every other section is skipped.  Supported `.dyr` models: GENROU, GENSAL,
IEESGO, TGOV1, SEXS, ESDC1A, IEEEST (MODE 1); other records are skipped with
a warning.
"""
from __future__ import annotations

import math
import re
import warnings
from dataclasses import dataclass, field
from typing import List, Optional, Set, Tuple

from .system import (Bus, Gen, Load, Branch, Shunt, PowerSystem, PowerSystemDynamics,
                     Genrou, Gensal, IEESGO, TGOV1, SEXS, ESDC1A, IEEEST,
                     esdc1a_sat_coefficients, set_dynamics, _norm_id,
                     CONTROLLER_TARGET, EXCITERS)


# ----------------------------------------------------------------------------
# .raw records (src/parse_psse_raw.jl).  Each record keeps the fields the
# conversion uses; field numbers below are PSS/E v33's, 1-based.
# ----------------------------------------------------------------------------
def _fields(line):
    return line.rstrip("\n").split(",")


def _unquote(s):
    return s.strip().strip("'")


@dataclass
class BusRaw:           # src/parse_psse_raw.jl:1-35
    busn: int
    name: str
    baseKV: float
    type: int
    vm: float
    va: float           # degrees

    @classmethod
    def parse(cls, line):
        f = _fields(line)
        return cls(int(f[0]), f[1], float(f[2]), int(f[3]), float(f[7]), float(f[8]))


@dataclass
class LoadRaw:          # src/parse_psse_raw.jl:37-78
    busn: int
    name: str
    status: int
    pl: float           # MW
    ql: float           # Mvar

    @classmethod
    def parse(cls, line):
        f = _fields(line)
        return cls(int(f[0]), _unquote(f[1]), int(f[2]), float(f[5]), float(f[6]))


@dataclass
class ShuntRaw:         # src/parse_psse_raw.jl:80-100
    busn: int
    name: str
    status: int
    gshunt: float       # MW at 1 pu
    bshunt: float       # Mvar at 1 pu

    @classmethod
    def parse(cls, line):
        f = _fields(line)
        return cls(int(f[0]), f[1], int(f[2]), float(f[3]), float(f[4]))


@dataclass
class GenRaw:           # src/parse_psse_raw.jl:102-160
    busn: int
    name: str
    pg: float
    qg: float
    vs: float
    mbase: float
    status: int

    @classmethod
    def parse(cls, line):
        f = _fields(line)
        return cls(int(f[0]), f[1], float(f[2]), float(f[3]), float(f[6]), float(f[8]), int(f[14]))


@dataclass
class BranchRaw:        # src/parse_psse_raw.jl:162-215
    fbus: int
    tbus: int
    ckt: str
    r: float
    x: float
    b: float
    status: int

    @classmethod
    def parse(cls, line):
        f = _fields(line)
        return cls(int(f[0]), int(f[1]), f[2], float(f[3]), float(f[4]), float(f[5]), int(f[13]))


@dataclass
class TransformerRaw:   # two-winding, 4 lines (src/parse_psse_raw.jl:217-310)
    fbus: int
    tbus: int
    ckt: str
    CW: int
    CZ: int
    MAG1: float
    MAG2: float
    status: int
    r: float
    x: float
    sbase12: float
    WINDV1: float
    NOMV1: float
    ANG1: float
    COD1: int
    WINDV2: float
    NOMV2: float

    @classmethod
    def parse(cls, l1, l2, l3, l4):
        a, b, c, d = _fields(l1), _fields(l2), _fields(l3), _fields(l4)
        return cls(int(a[0]), int(a[1]), a[3], int(a[4]), int(a[5]), float(a[7]), float(a[8]),
                   int(a[11]), float(b[0]), float(b[1]), float(b[2]), float(c[0]), float(c[1]),
                   float(c[2]), int(c[6]), float(d[0]), float(d[1]))


@dataclass
class ThreeTransformerRaw:  # three-winding, 5 lines (src/parse_psse_raw.jl:312-410)
    ibus: int
    jbus: int
    kbus: int
    ckt: str
    CW: int
    CZ: int
    status: int
    r12: float
    x12: float
    sbase12: float
    r23: float
    x23: float
    sbase23: float
    r13: float
    x13: float
    sbase31: float
    vmstar: float
    anstar: float
    WINDV1: float
    ANG1: float
    WINDV2: float
    ANG2: float
    WINDV3: float
    ANG3: float

    @classmethod
    def parse(cls, l1, l2, l3, l4, l5):
        a, b = _fields(l1), _fields(l2)
        w = [_fields(x) for x in (l3, l4, l5)]
        return cls(int(a[0]), int(a[1]), int(a[2]), a[3], int(a[4]), int(a[5]), int(a[11]),
                   *[float(v) for v in b[:11]],
                   float(w[0][0]), float(w[0][2]), float(w[1][0]), float(w[1][2]),
                   float(w[2][0]), float(w[2][2]))


@dataclass
class SwitchedShuntRaw:     # src/parse_psse_raw.jl:412-425
    busn: int
    status: int
    binit: float

    @classmethod
    def parse(cls, line):
        f = _fields(line)
        status = int(f[3]) if len(f) > 3 else 0
        binit = float(f[9]) if len(f) > 9 else 0.0
        return cls(int(f[0]), status, binit)


@dataclass
class PsystemRaw:
    baseMVA: float
    buses: List[BusRaw] = field(default_factory=list)
    loads: List[LoadRaw] = field(default_factory=list)
    shunts: List[ShuntRaw] = field(default_factory=list)
    branches: List[BranchRaw] = field(default_factory=list)
    gens: List[GenRaw] = field(default_factory=list)
    transformers: List[TransformerRaw] = field(default_factory=list)
    transthree: List[ThreeTransformerRaw] = field(default_factory=list)
    switched_shunts: List[SwitchedShuntRaw] = field(default_factory=list)


# Section terminators, in file order (src/parse_psse_raw.jl:441-540).
_ENDS = (("buses", BusRaw, "0 / END OF BUS DATA, BEGIN LOAD DATA"),
         ("loads", LoadRaw, "0 / END OF LOAD DATA, BEGIN FIXED SHUNT DATA"),
         ("shunts", ShuntRaw, "0 / END OF FIXED SHUNT DATA, BEGIN GENERATOR DATA"),
         ("gens", GenRaw, "0 / END OF GENERATOR DATA, BEGIN BRANCH DATA"),
         ("branches", BranchRaw, "0 / END OF BRANCH DATA, BEGIN TRANSFORMER DATA"))
_END_XF = "0 / END OF TRANSFORMER DATA, BEGIN AREA DATA"


def read_psse_raw(filename) -> PsystemRaw:
    """Parse a PSS/E v33 .raw file (src/parse_psse_raw.jl:570-584).

    The header's second field is the system MVA base; two title lines follow.
    Each record section ends at its "0 / END OF ..." marker or at a blank
    line; a transformer record spans four lines when its third bus is 0 and
    five otherwise.  Sections between transformers and switched shunts are
    skipped."""
    with open(filename, "r") as fh:
        lines = fh.read().split("\n")
    pos = 0

    def nxt():
        nonlocal pos
        if pos >= len(lines):
            return None
        s = lines[pos]
        pos += 1
        return s

    head = nxt()
    raw = PsystemRaw(float(head.split("/")[0].split(",")[1]))
    nxt()
    nxt()
    for attr, rec, end in _ENDS:
        while True:
            s = nxt()
            if s is None or s == "" or end in s:
                break
            getattr(raw, attr).append(rec.parse(s))
    while True:
        s = nxt()
        if s is None or s == "" or _END_XF in s:
            break
        if int(_fields(s)[2]) == 0:
            raw.transformers.append(TransformerRaw.parse(s, nxt(), nxt(), nxt()))
        else:
            raw.transthree.append(ThreeTransformerRaw.parse(s, nxt(), nxt(), nxt(), nxt()))
    while True:
        s = nxt()
        if s is None or "BEGIN SWITCHED SHUNT DATA" in s:
            break
    while True:
        s = nxt()
        if s is None or "END OF SWITCHED SHUNT DATA" in s:
            break
        if s.strip() == "":
            continue
        raw.switched_shunts.append(SwitchedShuntRaw.parse(s))
    return raw


def raw_to_grad(raw: PsystemRaw) -> PowerSystem:
    """PsystemRaw -> PowerSystem in per unit (src/parse.jl:212-408).

    Lines and transformers in service become pi branches (transformer
    impedance rebased by CZ/CW, tap WINDV1/WINDV2, phase shift ANG1 in
    degrees, magnetising MAG2 as charging); a three-winding transformer
    becomes a star bus plus one branch per winding in service.  Generators
    in service fix the voltage set point of their PV/slack bus; a PV bus with
    none left becomes PQ.  Powers are divided by the system base; loads keep
    the reactive-injection sign convention (-Q)."""
    base = raw.baseMVA
    buses: List[Bus] = []
    busmap = {}
    for k, b in enumerate(raw.buses):
        buses.append(Bus(b.busn, b.name, b.type, b.baseKV, b.vm, math.pi / 180.0 * b.va))
        busmap[b.busn] = k
    branches: List[Branch] = []
    shunts: List[Shunt] = []
    for br in raw.branches:
        if br.status != 1:
            continue
        branches.append(Branch(busmap[br.fbus], busmap[br.tbus], br.ckt, br.r, br.x, br.b, 0.0, 0.0))
    for t in raw.transformers:
        if t.status != 1:
            continue
        fr, to = busmap[t.fbus], busmap[t.tbus]
        if t.CW == 2:
            raise NotImplementedError("transformer winding-voltage code CW=2 is not supported")
        v1, v2 = t.WINDV1, t.WINDV2
        zratio = 1.0
        if t.CW == 1 and t.NOMV1 > 0.0:
            zratio = (t.NOMV1 / buses[fr].baseKV) ** 2
        if t.CZ == 1:
            r12, x12 = t.r * v2 ** 2 * zratio, t.x * v2 ** 2 * zratio
        elif t.CZ == 2:
            r12 = t.r * (base / t.sbase12) * v2 ** 2
            x12 = t.x * (base / t.sbase12) * v2 ** 2
        else:
            raise NotImplementedError("two-winding transformer impedance code CZ=3 is not supported")
        mag = t.MAG2 if abs(t.MAG2) > 0.0 else 0.0
        branches.append(Branch(fr, to, t.ckt, r12, x12, mag, v1 / v2, t.ANG1))
        if t.COD1 == 1:
            # the reference scales the magnetising admittance by the MVA base
            # here (src/parse.jl:266-268); kept as is for identical results
            shunts.append(Shunt(fr, "tran", t.MAG1 * base, t.MAG2 * base))
    if raw.transthree:
        top = max(busmap) if busmap else 0
        for nd, t in enumerate(raw.transthree):
            ib, jb, kb = busmap[t.ibus], busmap[t.jbus], busmap[t.kbus]
            star_n = top + 1 + nd
            star = len(buses)
            buses.append(Bus(star_n, "STAR", 1, 0.0, t.vmstar, math.pi / 180.0 * t.anstar))
            busmap[star_n] = star
            if t.CW == 2:
                v1 = t.WINDV1 / buses[ib].baseKV
                v2 = t.WINDV2 / buses[jb].baseKV
                v3 = t.WINDV3 / buses[kb].baseKV
            else:
                v1, v2, v3 = t.WINDV1, t.WINDV2, t.WINDV3
            if t.CZ == 1:
                r12, x12, r23, x23, r13, x13 = t.r12, t.x12, t.r23, t.x23, t.r13, t.x13
            elif t.CZ == 2:
                s12, s23, s31 = base / t.sbase12, base / t.sbase23, base / t.sbase31
                r12, x12 = t.r12 * s12, t.x12 * s12
                r23, x23 = t.r23 * s23, t.x23 * s23
                r13, x13 = t.r13 * s31, t.x13 * s31
            else:
                # load loss in W and |Z| in pu on the winding base
                r12 = t.r12 / 1e6 / t.sbase12
                r23 = t.r23 / 1e6 / t.sbase23
                r13 = t.r13 / 1e6 / t.sbase31
                x12 = math.sqrt(t.x12 ** 2 - r12 ** 2)
                x23 = math.sqrt(t.x23 ** 2 - r23 ** 2)
                x13 = math.sqrt(t.x13 ** 2 - r13 ** 2)
                s12, s23, s31 = base / t.sbase12, base / t.sbase23, base / t.sbase31
                r12, x12, r23, x23, r13, x13 = r12 * s12, x12 * s12, r23 * s23, x23 * s23, r13 * s31, x13 * s31
            # star equivalent of the three pairwise impedances
            r1, x1 = 0.5 * (r12 + r13 - r23), 0.5 * (x12 + x13 - x23)
            r2, x2 = 0.5 * (r12 - r13 + r23), 0.5 * (x12 - x13 + x23)
            r3, x3 = 0.5 * (r13 + r23 - r12), 0.5 * (x13 + x23 - x12)
            # status: 1 all in, 2 winding 2 out, 3 winding 3 out, 4 winding 1 out, else all out
            on = {1: (1, 1, 1), 2: (1, 0, 1), 3: (1, 1, 0), 4: (0, 1, 1)}.get(t.status, (0, 0, 0))
            if on[0]:
                branches.append(Branch(ib, star, t.ckt, r1, x1, 0.0, v1, t.ANG1))
            if on[1]:
                branches.append(Branch(star, jb, t.ckt, r2, x2, 0.0, v2, t.ANG2))
            if on[2]:
                branches.append(Branch(star, kb, t.ckt, r3, x3, 0.0, v3, t.ANG3))
    gens: List[Gen] = []
    with_gen = {busmap[g.busn] for g in raw.gens if g.status == 1}
    for g in raw.gens:
        if g.status != 1:
            continue
        b = busmap[g.busn]
        gens.append(Gen(b, g.name, g.pg / base, g.qg / base, g.mbase, g.status))
        if buses[b].type in (2, 3):
            buses[b].v0m = g.vs
    for k, b in enumerate(buses):
        if b.type == 2 and k not in with_gen:
            b.type = 1
    loads = [Load(busmap[l.busn], l.name, l.pl / base, -l.ql / base) for l in raw.loads if l.status == 1]
    for s in raw.shunts:
        if s.status == 1:
            shunts.append(Shunt(busmap[s.busn], s.name, s.gshunt / base, s.bshunt / base))
    for s in raw.switched_shunts:
        if s.status == 1:
            shunts.append(Shunt(busmap[s.busn], "swsh", 0.0, s.binit / base))
    ps = PowerSystem(base, buses, gens, loads, branches, shunts)
    assert ps.busmap == busmap
    return ps


# ----------------------------------------------------------------------------
# .dyr records
# ----------------------------------------------------------------------------
def read_psse_dyr(filename) -> List[List[str]]:
    """Split a .dyr file into records (src/parse.jl:115-156).

    A record starts on a line whose first token is not a `//` comment and
    runs over continuation lines until its last token is `/`.  Fields are
    separated by commas (when the record's first line has one) or by white
    space; the closing `/` is kept as the record's last field, as the
    reference keeps it (its field count decides whether GENROU's optional
    saturation pair is present)."""
    with open(filename, "r") as fh:
        data = fh.read().split("\n")
    recs = []
    i = 0
    n = len(data)
    while i < n:
        s = data[i].strip()
        dev = re.split(r"\s*,\s*", s) if "," in data[i] else re.split(r"\s+", s)
        if not dev or dev[0] == "":
            i += 1
            continue
        if dev[0].startswith("//"):
            i += 1
            continue
        i += 1
        while dev[-1] != "/":
            if i >= n:
                raise ValueError(f"{filename}: record {dev[:3]} has no closing '/'")
            dev += re.split(r"\s*,\s*|\s+", data[i].strip())
            i += 1
        recs.append(dev)
    return recs


def _num(f, k):
    return float(f[k - 1])


def _int(f, k):
    return int(f[k - 1])


def _genrou(f):         # src/generators.jl:87-108
    S1 = _num(f, 16) if len(f) >= 17 else 0.0
    S2 = _num(f, 17) if len(f) >= 18 else 0.0
    return Genrou(_int(f, 1), f[2], _num(f, 10), _num(f, 11), _num(f, 12), _num(f, 13), _num(f, 14),
                  _num(f, 15), _num(f, 8), _num(f, 9), _num(f, 4), _num(f, 6), _num(f, 5), _num(f, 7), S1, S2)


def _gensal(f):         # src/generators.jl:46-66
    S1 = _num(f, 14) if len(f) >= 15 else 0.0
    S2 = _num(f, 15) if len(f) >= 16 else 0.0
    return Gensal(_int(f, 1), f[2], _num(f, 9), _num(f, 10), _num(f, 11), _num(f, 12), _num(f, 13),
                  _num(f, 7), _num(f, 8), _num(f, 4), _num(f, 5), _num(f, 6), S1, S2)


def _ieesgo(f):         # src/governors.jl:35-51
    return IEESGO(_int(f, 1), f[2], *[_num(f, k) for k in range(4, 15)])


def _tgov1(f):          # src/governors.jl:167-178
    return TGOV1(_int(f, 1), f[2], *[_num(f, k) for k in range(4, 11)])


def _sexs(f):           # src/exciters.jl:339-349
    return SEXS(_int(f, 1), f[2], *[_num(f, k) for k in range(4, 10)])


def _esdc1a(f):         # src/exciters.jl:52-74: Tb, Tc, Vrmax, Vrmin, Sw are not modelled
    sat_a, sat_b = esdc1a_sat_coefficients(_num(f, 16), _num(f, 17), _num(f, 18), _num(f, 19))
    return ESDC1A(_int(f, 1), f[2], _num(f, 5), _num(f, 6), _num(f, 13), _num(f, 14),
                  _num(f, 11), _num(f, 12), _num(f, 4), sat_a, sat_b)


def _ieeest(f):         # src/stabilizers.jl:101-150 (zero time constants clamped by IEEEST)
    return IEEEST(_int(f, 1), f[2], _int(f, 4), _int(f, 5), *[_num(f, k) for k in range(6, 23)])


DEVICE_TYPE_MAP = {     # src/parse.jl:104-113
    "GENROU": _genrou,
    "GENSAL": _gensal,
    "IEESGO": _ieesgo,
    "TGOV1": _tgov1,
    "SEXS": _sexs,
    "ESDC1A": _esdc1a,
    "IEEEST": _ieeest,
}


def create_device_vector(records, active_gen_keys: Optional[Set[Tuple[int, str]]] = None):
    """Device models from .dyr records (src/parse.jl:410-491), in three passes:
    machines whose (bus, id) is an active static generator (all machines when
    `active_gen_keys` is None); governors and exciters whose machine was
    kept; stabilisers whose machine and exciter were kept.  Unknown model
    names are skipped with one warning that counts them."""
    parsed = []
    unknown = {}
    for f in records:
        name = f[1].strip("'")
        make = DEVICE_TYPE_MAP.get(name)
        if make is None:
            unknown[name] = unknown.get(name, 0) + 1
            continue
        parsed.append(make(f))
    out = []
    kept_gen = set()
    for d in parsed:
        if d.kind != "genrou":
            continue
        key = (d.bus, _norm_id(d.id))
        if active_gen_keys is None or key in active_gen_keys:
            kept_gen.add(key)
            out.append(d)
    kept_exc = set()
    for d in parsed:
        if d.kind == "genrou" or d.kind == "ieeest":
            continue
        key = (d.bus, _norm_id(d.id))
        if d.kind in CONTROLLER_TARGET and key not in kept_gen:
            continue
        if d.kind in EXCITERS:
            kept_exc.add(key)
        out.append(d)
    for d in parsed:
        if d.kind != "ieeest":
            continue
        key = (d.bus, _norm_id(d.id))
        if key in kept_gen and key in kept_exc:
            out.append(d)
    if unknown:
        warnings.warn(f"Unknown device type(s) in .dyr (skipped): {unknown}")
    return out


def dynamics_from_dyr(filename, active_gen_keys=None) -> PowerSystemDynamics:
    """PowerSystemDynamics(dyr_file) (src/GradPower.jl:253-262)."""
    psd = PowerSystemDynamics()
    for d in create_device_vector(read_psse_dyr(filename), active_gen_keys):
        psd.add_device(d)
    return psd


def from_psse(raw_file, dyr_file=None, add_static_gen_stubs=True) -> PowerSystem:
    """Network from `raw_file` (a path, or a PsystemRaw already read); with
    `dyr_file`, its device models for the generators in service, wired by
    set_dynamics (src/parse.jl:81-98)."""
    raw = raw_file if isinstance(raw_file, PsystemRaw) else read_psse_raw(raw_file)
    ps = raw_to_grad(raw)
    if dyr_file is not None:
        active = {(ps.buses[g.bus].i, _norm_id(g.id)) for g in ps.gens}
        psd = dynamics_from_dyr(dyr_file, active)
        set_dynamics(ps, psd, add_static_gen_stubs=add_static_gen_stubs)
    return ps
