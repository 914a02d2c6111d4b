"""PSS/E v33 reader (gradpower_hip.psse) on the reference's bundled cases.

The .raw/.dyr files under tests/golden/psse are the reference's own data
files (examples/*.raw, examples/*.dyr, test/testdata/2bus.*), copied as
fixtures; 2bus_esdc1a.npz is the reference's stored trajectory of the 2-bus
GENROU+ESDC1A case from an independent simulator (examples/refs/, written by
scripts/gen_ref_2bus_esdc1a.py with the uqgrid package).

Checks, each chosen so that a misread field fails it:
  - 2bus.raw + 2bus.dyr read into exactly the values of the hand-typed
    2-bus test system (test/test_dynamics.jl:3-28 builds it this way);
  - the power flow on the read IEEE 9-bus network reproduces the solved
    voltages that PSS/E stored in ieee9_v33.raw (5 digits);
  - every bundled raw/dyr pair initialises to a steady state (residual
    < 1e-9, test/test_ieesgo_kernel.jl:13-43);
  - the 2-bus ESDC1A case's initial state equals the independent
    simulator's to 1e-9 and its rotor speed follows that simulator's within
    2e-4 pu over 5 s (observed 1e-4: that simulator's ESDC1A keeps two
    states and the limits that the reduced model here leaves out; a 10 %
    change of H, or D = 0.5 instead of 0, moves the speed by 1.3e-3 or more);
  - a three-winding transformer becomes a star whose winding impedances add
    up pairwise to the given ones; inactive generators drop their machine
    and controllers; unknown models are skipped with a warning.
"""
import math
import os
import warnings

import numpy as np
import pytest

from conftest import get_case

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "psse")

PAIRS = [("2bus.raw", "2bus.dyr"), ("2bus.raw", "2bus_ESDC1A.dyr"), ("2bus.raw", "2bus_IEESGO.dyr"),
         ("2bus.raw", "2bus_SEXS.dyr"), ("2bus.raw", "2bus_TGOV1.dyr"),
         ("ieee9_v33.raw", "ieee9bus.dyr"), ("ieee9_v33.raw", "ieee9bus_SEXS.dyr"),
         ("ieee9_v33.raw", "ieee9bus_TGOV1.dyr"), ("ieee9_v33.raw", "ieee9bus_gov.dyr"),
         ("IEEE39.raw", "IEEE39.dyr"), ("IEEE39.raw", "IEEE39_gov.dyr")]


def _p(name):
    return os.path.join(G, name)


def _setup(raw, dyr, exact_slack=False, zip_alpha=None):
    from gradpower_hip import psse
    from gradpower_hip.network import build_network, runpf_inplace
    from gradpower_hip.system import DynamicProblem
    from gradpower_hip.init import initialize_dynamics
    from gradpower_hip.desc import build_desc
    ps = psse.from_psse(_p(raw), _p(dyr))
    if zip_alpha is not None:
        for dev in ps.dynamic.devices:
            if dev.dtype.kind == "zipload":
                dev.dtype.alpha = zip_alpha
    build_network(ps)
    runpf_inplace(ps, exact_slack=exact_slack)
    dp = DynamicProblem(ps)
    initialize_dynamics(dp, ps)
    return ps, build_desc(ps, dp)


def _oracle(desc):
    from oracle.oracle import Oracle
    return Oracle(desc)


def test_2bus_reads_as_the_typed_system():
    from gradpower_hip import psse
    ref = get_case("case_2bus_testdata").ps
    ps = psse.from_psse(_p("2bus.raw"), _p("2bus.dyr"))
    assert ps.baseMVA == ref.baseMVA == 100.0
    for a, b in zip(ps.buses, ref.buses):
        assert (a.i, a.type, a.baseKV) == (b.i, b.type, b.baseKV)
    # before the power flow, bus 2 carries the file's stored voltage
    raw = psse.read_psse_raw(_p("2bus.raw"))
    assert ps.buses[1].v0m == 1.01613 and ps.buses[1].v0a == pytest.approx(math.radians(-3.3252), abs=1e-15)
    assert [b.vm for b in raw.buses] == [1.04, 1.01613]
    # the typed system's generator table before its power flow rewrote it
    (g,) = ps.gens
    assert (g.bus, g.psch, g.qsch, g.mbase) == (0, 106.496 / 100.0, 45.996 / 100.0, 100.0)
    (l,), (lr,) = ps.loads, ref.loads
    assert (l.bus, l.pd, l.qd) == (lr.bus, lr.pd, lr.qd)
    (br,), (brr,) = ps.branches, ref.branches
    assert (br.fr, br.to, br.r, br.x, br.sh, br.tap) == (brr.fr, brr.to, brr.r, brr.x, brr.sh, brr.tap)
    gm = [d.dtype for d in ps.dynamic.devices if d.dtype.kind == "genrou"][0]
    gmr = [d.dtype for d in ref.dynamic.devices if d.dtype.kind == "genrou"][0]
    pa, pb = np.zeros(14), np.zeros(14)
    gm.fill_pvec(pa)
    gmr.fill_pvec(pb)
    assert np.array_equal(pa, pb)


def test_tables_of_bundled_cases():
    from gradpower_hip import psse
    r9 = psse.read_psse_raw(_p("ieee9_v33.raw"))
    assert (len(r9.buses), len(r9.loads), len(r9.gens), len(r9.branches), len(r9.transformers)) == (9, 3, 3, 6, 3)
    r39 = psse.read_psse_raw(_p("IEEE39.raw"))
    assert (len(r39.buses), len(r39.loads), len(r39.gens), len(r39.branches), len(r39.transformers)) == (39, 19, 10, 34, 12)
    assert [(s.busn, s.status, s.binit) for s in r39.switched_shunts] == [(4, 1, 100.0), (5, 1, 200.0)]
    ps = psse.raw_to_grad(r39)
    assert len(ps.branches) == 46
    # switched shunts at their initial susceptance on the system base; the
    # transformer 31-6 with COD1 = 1 adds its (zero) magnetising shunt
    assert [(s.bus, s.bsh) for s in ps.shunts] == [(ps.busmap[31], 0.0), (ps.busmap[4], 1.0), (ps.busmap[5], 2.0)]
    t = [b for b in ps.branches if b.tap != 0.0]
    assert len(t) == 12 and t[0].x == 1.81e-2 and t[0].tap == 1.025
    # .dyr: comma records on one line and white-space records over three
    recs = psse.read_psse_dyr(_p("IEEE39_gov.dyr"))
    assert [r[1] for r in recs] == ["'GENROU'"] * 10 + ["'IEESGO'"] * 10
    assert all(r[-1] == "/" for r in recs) and len(recs[0]) == 18
    recs9 = psse.read_psse_dyr(_p("ieee9bus_gov.dyr"))
    assert [len(r) for r in recs9] == [18, 18, 18, 15, 15, 15]
    psd = psse.dynamics_from_dyr(_p("ieee9bus_gov.dyr"))
    g = psd.devices[0].dtype
    assert (g.bus, g.T_d0p, g.T_d0dp, g.T_q0p, g.T_q0dp, g.H, g.D) == (1, 7.729, 0.047, 0.859, 0.068, 4.31, 0.01)
    assert (g.x_d, g.x_q, g.x_dp, g.x_qp, g.x_ddp, g.xl) == (1.9266, 1.8442, 0.3812, 0.5469, 0.2889, 0.2443)
    assert (g.S1, g.S2) == (0.115, 0.627)
    gv = psd.devices[5].dtype
    assert gv.kind == "ieesgo" and (gv.K1, gv.pmax, gv.pmin) == (16.0, 100.0, -100.0)


def test_power_flow_reproduces_stored_ieee9_solution():
    """ieee9_v33.raw stores PSS/E's solved bus voltages (5 digits)."""
    from gradpower_hip import psse
    raw = psse.read_psse_raw(_p("ieee9_v33.raw"))
    ps, _ = _setup("ieee9_v33.raw", "ieee9bus.dyr")
    vm = np.array([b.v0m for b in ps.buses])
    va = np.array([b.v0a for b in ps.buses])
    assert np.max(np.abs(vm - [b.vm for b in raw.buses])) < 1e-5
    assert np.max(np.abs(np.degrees(va) - [b.va for b in raw.buses])) < 1e-4


@pytest.mark.parametrize("raw,dyr", PAIRS)
def test_bundled_cases_initialise_to_steady_state(raw, dyr):
    _, desc = _setup(raw, dyr, exact_slack=True)
    f = _oracle(desc).residual(desc.z0, desc.z0, 1.0 / 120.0)
    assert np.max(np.abs(f)) < 1e-9


def test_ieee39_slack_load_as_the_reference_writes_it_back():
    """Without exact_slack the slack machine of bus 39 is set 2 x 11.04 pu
    short, as src/pflow.jl:393-414 sets it: the initial state balances every
    row but the two network rows of bus 39."""
    ps, desc = _setup("IEEE39.raw", "IEEE39.dyr", exact_slack=False)
    f = _oracle(desc).residual(desc.z0, desc.z0, 1.0 / 120.0)
    bad = np.nonzero(np.abs(f) > 1e-9)[0]
    k = ps.busmap[39]
    nz = len(desc.z0)
    assert list(bad) == [nz - 2 * ps.nbus + 2 * k, nz - 2 * ps.nbus + 2 * k + 1]
    assert np.max(np.abs(f)) > 1.0


def test_2bus_esdc1a_against_stored_trajectory():
    gold = np.load(_p("2bus_esdc1a.npz"))
    H = gold["history"]
    ps, desc = _setup("2bus.raw", "2bus_ESDC1A.dyr", zip_alpha=float(gold["zipload_alpha"]))
    # that simulator's state order -> ours: machine diff (6), machine alg (4),
    # exciter (its last 3 of 5 states), bus voltages (vr, vi per bus)
    mp = list(range(6)) + [11, 12, 13, 14] + [8, 9, 10] + [17, 18, 19, 20]
    assert len(desc.z0) == len(mp)
    assert np.max(np.abs(desc.z0 - H[mp, 0])) < 1e-9
    h = float(gold["tvec"][1])
    n = H.shape[1] - 1
    assert np.allclose(np.diff(gold["tvec"]), h, rtol=1e-9, atol=0)
    on, off = int(round(float(gold["ton"]) / h)), int(round(float(gold["toff"]) / h))
    tr, _, _ = _oracle(desc).integrate(h, n, [(int(gold["fault_bus"]), float(gold["rfault"]), on, off)], tol=1e-12)
    w = tr[:, 4]
    assert int(gold["speed_idx"][0]) == 4
    assert np.max(np.abs(w - H[4])) < 2e-4
    assert np.max(np.abs(w)) > 1e-2


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_three_winding_transformer_star(tmp_path):
    """A three-winding transformer (CZ = 2: impedances on the winding base
    sbase) added to the 9-bus case between buses 4, 5 and 6."""
    from gradpower_hip import psse
    txt = open(_p("ieee9_v33.raw")).read()
    xf = ("     4,     5,     6,'X3',1,2,1, 0.00000E+0, 0.00000E+0,2,'            ',1,   1,1.0000,   0,1.0000,   0,1.0000,   0,1.0000,'            '\n"
          " 0.002, 0.08, 200.0, 0.003, 0.12, 150.0, 0.004, 0.10, 250.0, 1.01, -2.5\n"
          "1.02,   0.000,   1.500,     0.00,     0.00,     0.00, 0,      0, 1.10000, 0.90000, 1.10000, 0.90000,  33, 0, 0.00000, 0.00000,  0.000\n"
          "0.99,   0.000,   0.000,     0.00,     0.00,     0.00, 0,      0, 1.10000, 0.90000, 1.10000, 0.90000,  33, 0, 0.00000, 0.00000,  0.000\n"
          "1.00,   0.000,  -1.000,     0.00,     0.00,     0.00, 0,      0, 1.10000, 0.90000, 1.10000, 0.90000,  33, 0, 0.00000, 0.00000,  0.000\n")
    txt = txt.replace("0 / END OF TRANSFORMER DATA", xf + "0 / END OF TRANSFORMER DATA")
    raw = psse.read_psse_raw(_write(tmp_path, "x3.raw", txt))
    assert len(raw.transthree) == 1 and len(raw.transformers) == 3
    ps = psse.raw_to_grad(raw)
    assert ps.nbus == 10 and ps.buses[9].i == 10 and ps.buses[9].v0m == 1.01
    assert ps.buses[9].v0a == pytest.approx(math.radians(-2.5), abs=1e-15)
    w = ps.branches[-3:]
    star = 9
    assert [(b.fr, b.to) for b in w] == [(ps.busmap[4], star), (star, ps.busmap[5]), (star, ps.busmap[6])]
    assert [b.tap for b in w] == [1.02, 0.99, 1.00] and [b.shift for b in w] == [1.5, 0.0, -1.0]
    s12, s23, s31 = 100.0 / 200.0, 100.0 / 150.0, 100.0 / 250.0
    assert w[0].r + w[1].r == pytest.approx(0.002 * s12, abs=1e-15)
    assert w[1].r + w[2].r == pytest.approx(0.003 * s23, abs=1e-15)
    assert w[0].r + w[2].r == pytest.approx(0.004 * s31, abs=1e-15)
    assert w[0].x + w[1].x == pytest.approx(0.08 * s12, abs=1e-15)
    assert w[1].x + w[2].x == pytest.approx(0.12 * s23, abs=1e-15)
    assert w[0].x + w[2].x == pytest.approx(0.10 * s31, abs=1e-15)
    # status 2: winding 2 out of service
    raw.transthree[0].status = 2
    assert len(psse.raw_to_grad(raw).branches) == len(ps.branches) - 1


def test_inactive_generator_drops_machine_and_controllers(tmp_path):
    from gradpower_hip import psse
    from gradpower_hip.system import PQ
    lines = open(_p("ieee9_v33.raw")).read().split("\n")
    k = next(i for i, s in enumerate(lines) if "BEGIN GENERATOR DATA" in s) + 3   # third generator (bus 3)
    f = lines[k].split(",")
    assert int(f[0]) == 3 and int(f[14]) == 1
    f[14] = "0"
    lines[k] = ",".join(f)
    ps = psse.from_psse(_write(tmp_path, "g3off.raw", "\n".join(lines)), _p("ieee9bus_gov.dyr"))
    assert len(ps.gens) == 2 and ps.buses[ps.busmap[3]].type == PQ
    kinds = [(d.dtype.kind, d.dtype.bus) for d in ps.dynamic.devices if d.dtype.kind in ("genrou", "ieesgo")]
    assert kinds == [("genrou", 1), ("genrou", 2), ("ieesgo", 1), ("ieesgo", 2)]


def test_dyr_models_unknown_records_and_stabiliser_filter(tmp_path):
    from gradpower_hip import psse
    text = (
        "// comment line\n"
        "1 'GENSAL' 1 5.0 0.05 0.1 3.0 0.5 1.8 1.1 0.3 0.2 0.15 0.1 0.4 /\n"
        "1 'SEXS' 1 0.1 10 100 0.05 -999 999 /\n"
        "1 'IEEEST' 1 1 0 0 0 0 0 0 0 0.2 0.02 0.0 0.0 1.0 1.0 10.0 0.1 -0.1 0 0 /\n"
        "1 'XYZ12' 1 1.0 /\n"
        "1,'ESDC1A',2,0.0,25.0,0.2,0.0,0.0,1.0,-1.0,-0.05,0.57,0.091,0.35,0.0,2.846,0.0811,3.795,0.3162,/\n"
        "1 'IEEEST' 2 1 0 0 0 0 0 0 0 0.2 0.02 0.0 0.0 1.0 1.0 10.0 0.1 -0.1 0 0 /\n"
    )
    path = _write(tmp_path, "m.dyr", text)
    with pytest.warns(UserWarning, match="XYZ12"):
        devs = psse.create_device_vector(psse.read_psse_dyr(path))
    kinds = [(d.kind, d.id) for d in devs]
    # machine 2 does not exist: its exciter and stabiliser are dropped
    assert kinds == [("genrou", "1"), ("sexs", "1"), ("ieeest", "1")]
    g, _, pss = devs
    assert g.x_qp == g.x_dp == 0.3 and g.T_q0p == g.T_d0p == 5.0
    assert (g.T_d0dp, g.T_q0dp, g.H, g.D, g.x_d, g.x_q, g.x_ddp, g.xl, g.S1, g.S2) == \
        (0.05, 0.1, 3.0, 0.5, 1.8, 1.1, 0.2, 0.15, 0.1, 0.4)
    # zero lead-lag constants are clamped to 1e-3; KS, limits in place
    assert (pss.A1, pss.T1, pss.T3, pss.T4, pss.KS, pss.LSMAX, pss.LSMIN) == (1e-3, 0.2, 1e-3, 1e-3, 10.0, 0.1, -0.1)
    with pytest.warns(UserWarning, match="XYZ12"):
        e = psse.create_device_vector(psse.read_psse_dyr(path), active_gen_keys={(1, "1"), (1, "2")})
    assert [(d.kind, d.id) for d in e] == [("genrou", "1"), ("sexs", "1"), ("ieeest", "1")]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        esd = [d for d in psse.create_device_vector(
            psse.read_psse_dyr(path) + [["1", "'GENROU'", "2"] + ["1.0"] * 14 + ["/"]]) if d.kind == "esdc1a"]
    assert len(esd) == 1 and (esd[0].Ka, esd[0].Ta, esd[0].Ke, esd[0].Te, esd[0].Kf, esd[0].Tf) == \
        (25.0, 0.2, -0.05, 0.57, 0.091, 0.35)
