"""Steady-state initialisation of the dynamic states (src/dynamics.jl:123-250).

Each machine/controller is initialised by solving its own small steady-state
system (the reference uses NLsolve with ftol=1e-12, src/dynamics.jl:49); here
a damped Newton iteration with a central-difference Jacobian solves the same
equations to |f| < 1e-13.  The equations are restated from
src/generators.jl:214-285 (GENROU), src/governors.jl:71-138 (IEESGO),
src/governors.jl:205-256 (TGOV1), src/exciters.jl:123-187 (ESDC1A),
src/exciters.jl:365-412 (SEXS), src/stabilizers.jl:178-246 (IEEEST).
"""
from __future__ import annotations

import cmath
import math

import numpy as np

from .system import (PowerSystem, DynamicProblem, TWOPI60, PV, SLACK,
                     _norm_id, net_ptr)


def genrou_sat_coefficients(s1, s2):
    """Quadratic open-circuit saturation fit (src/generators.jl:71-80)."""
    if s1 <= 0.0 or s2 <= 0.0:
        return 0.0, 0.0
    e1, e2 = 1.0, 1.2
    a = math.sqrt(s1 * e1 / (s2 * e2))
    if a == 1.0:
        return 0.0, 0.0
    return e2 - (e1 - e2) / (a - 1.0), s2 * e2 * (a - 1.0) ** 2 / (e1 - e2) ** 2


def genrou_sat_se(psi, sat_a, sat_b):
    if sat_b == 0.0 or psi <= sat_a or psi == 0.0:
        return 0.0
    return sat_b * (psi - sat_a) ** 2 / psi


def _newton(fun, x0, tol=1e-13, maxit=60):
    x = np.array(x0, dtype=float)
    f = fun(x)
    for _ in range(maxit):
        if np.max(np.abs(f), initial=0.0) < tol:
            return x, f
        n = len(x)
        J = np.empty((n, n))
        for j in range(n):
            h = 1e-7 * max(1.0, abs(x[j]))
            xp = x.copy(); xp[j] += h
            xm = x.copy(); xm[j] -= h
            J[:, j] = (fun(xp) - fun(xm)) / (2 * h)
        dx = np.linalg.lstsq(J, -f, rcond=None)[0]
        lam = 1.0
        fn0 = np.max(np.abs(f))
        while True:
            xn = x + lam * dx
            fnew = fun(xn)
            if np.max(np.abs(fnew)) < fn0 or lam < 1e-4:
                break
            lam *= 0.5
        x, f = xn, fnew
    return x, f


# --- per-device steady-state equations ----------------------------------------

def _genrou_init_res(x, p, pg, qg, vm, va):
    (x_d, x_q, x_dp, x_qp, x_ddp, xl, H, D, T_d0p, T_q0p, T_d0dp, T_q0dp, S1, S2) = p
    x_qdp = x_ddp
    e_qp, e_dp, phi_1d, phi_2q, w, delta, v_q, v_d, i_q, i_d, e_fd, p_m = x
    psi_de = (x_ddp - xl) / (x_dp - xl) * e_qp + (x_dp - x_ddp) / (x_dp - xl) * phi_1d
    psi_qe = -(x_ddp - xl) / (x_qp - xl) * e_dp + (x_qp - x_ddp) / (x_qp - xl) * phi_2q
    sa, sb = genrou_sat_coefficients(S1, S2)
    Se = genrou_sat_se(math.sqrt(psi_de * psi_de + psi_qe * psi_qe), sa, sb)
    f = np.empty(12)
    f[0] = (-e_qp + e_fd - (i_d - (-x_ddp + x_dp) * (-e_qp + i_d * (x_dp - xl) + phi_1d)
                            / ((x_dp - xl) ** 2.0)) * (x_d - x_dp) - Se * psi_de) / T_d0p
    f[1] = (-e_dp + (i_q - (-x_qdp + x_qp) * (e_dp + i_q * (x_qp - xl) + phi_2q)
                     / ((x_qp - xl) ** 2.0)) * (x_q - x_qp)) / T_q0p
    f[2] = (e_qp - i_d * (x_dp - xl) - phi_1d) / T_d0dp
    f[3] = (-e_dp - i_q * (x_qp - xl) - phi_2q) / T_q0dp
    f[4] = (p_m - psi_de * i_q + psi_qe * i_d) / (2.0 * H)
    f[5] = TWOPI60 * w
    f[6] = i_d - ((x_ddp - xl) / (x_dp - xl) * e_qp + (x_dp - x_ddp) / (x_dp - xl) * phi_1d - v_q) / x_ddp
    f[7] = i_q - (-(x_qdp - xl) / (x_qp - xl) * e_dp + (x_qp - x_qdp) / (x_qp - xl) * phi_2q + v_d) / x_qdp
    f[8] = v_d - vm * math.sin(delta - va)
    f[9] = v_q - vm * math.cos(delta - va)
    f[10] = v_d * i_d + v_q * i_q - pg
    f[11] = v_q * i_d - v_d * i_q - qg
    return f


def _genrou_init_guess(p, pg, qg, vm, va):
    (x_d, x_q, x_dp, x_qp, x_ddp, xl) = p[:6]
    vt = vm * cmath.exp(1j * va)
    ig = (pg - 1j * qg) / vt.conjugate()
    delta = cmath.phase(vt + 1j * x_q * ig)
    v_d = vm * math.sin(delta - va)
    v_q = vm * math.cos(delta - va)
    i_d = (pg * v_d + qg * v_q) / (v_d ** 2 + v_q ** 2)
    i_q = (pg * v_q - qg * v_d) / (v_d ** 2 + v_q ** 2)
    e_dp = -x_qp * i_q + v_d
    e_qp = x_dp * i_d + v_q
    phi_1d = e_qp - (x_dp - xl) * i_d
    phi_2q = -e_dp - (x_qp - xl) * i_q
    e_fd = e_qp + (x_d - x_dp) * i_d
    return [e_qp, e_dp, phi_1d, phi_2q, 0.0, delta, v_q, v_d, i_q, i_d, e_fd, pg]


def _ieesgo_init_res(x, p, pg):
    T1, T2, T3, T4, T5, T6, K1, K2, K3 = p[:9]
    PF0, PLL, TP1, TP2, TP3, p_m, pref = x
    w = 0.0
    SatP = pref - (T2 / T3) * PF0 - PLL
    return np.array([(1.0 / T1) * (K1 * w - PF0),
                     (1.0 / T3) * ((1.0 - T2 / T3) * PF0 - PLL),
                     (1.0 / T4) * (SatP - TP1),
                     (1.0 / T5) * (K2 * TP1 - TP2),
                     (1.0 / T6) * (K3 * TP2 - TP3),
                     TP1 * (1 - K2) + TP2 * (1 - K3) + TP3 - p_m,
                     p_m - pg])


def _tgov1_init_res(x, p, pg):
    R, T1, _, _, T2, T3, DT = p[:7]
    x1, x2, p_m, pref = x
    w = 0.0
    return np.array([(-x1 + (1.0 - T2 / T3) * x2) / T3,
                     ((pref - w) / R - x2) / T1,
                     x1 + (T2 / T3) * x2 - DT * w - p_m,
                     p_m - pg])


def _sexs_init_res(x, p, vm, efd0):
    TA_TB, TB, K, TE = p[:4]
    x1, e_fd, vref = x
    return np.array([(-x1 + (1.0 - TA_TB) * (vref - vm)) / TB,
                     (-e_fd + K * (x1 + TA_TB * (vref - vm))) / TE,
                     e_fd - efd0])


def _ieeest_init_res(x, p):
    A1, A2, A3, A4, A5, A6, T1, T2, T3, T4, T5, T6, KS = p[:13]
    s0, s1, s2, s3, s4, s5, s6, vs, w = x
    y2 = s3 + A5 * s2 + (A6 / A4) * (s1 - s3 - A3 * s2)
    y3 = s4 + (T1 / T2) * (y2 - s4)
    y4 = s5 + (T3 / T4) * (y3 - s5)
    return np.array([(w - s1 - A1 * s0) / A2, s0, (s1 - s3 - A3 * s2) / A4, s2,
                     (y2 - s4) / T2, (y3 - s5) / T4, (KS * y4 - s6) / T6,
                     vs - (T5 / T6) * (KS * y4 - s6), w])


class InitError(RuntimeError):
    pass


def initialize_dynamics(dp: DynamicProblem, ps: PowerSystem, check_tol=1e-9):
    """Fill dp.zvec / dp.uvec / dp.pvec at the power-flow operating point."""
    psd = ps.dynamic
    z, u, p = dp.zvec, dp.uvec, dp.pvec
    npt = net_ptr(ps)
    for i, b in enumerate(ps.buses):
        z[npt + 2 * i] = b.v0m * math.cos(b.v0a)
        z[npt + 2 * i + 1] = b.v0m * math.sin(b.v0a)

    for i, dev in enumerate(psd.devices):
        d = dev.dtype
        if d.kind == "zipload":
            d.v0mag = ps.buses[psd.dev_bus[i]].v0m
            y = complex(d.pinj, d.qinj) / d.v0mag ** 2
            d.yreal, d.yimag = y.real, y.imag
    for dev in psd.devices:
        dev.dtype.fill_pvec(p[dev.par_ptr:dev.par_ptr + dev.dtype.par_size])

    def place(dev, x):
        d = dev.dtype
        ds, as_, cs = d.diff_size, d.alg_size, d.ctrl_size
        if ds:
            z[dev.diff_ptr:dev.diff_ptr + ds] = x[:ds]
        if as_:
            z[dev.alg_ptr:dev.alg_ptr + as_] = x[ds:ds + as_]
        if cs:
            u[dev.ctrl_ptr:dev.ctrl_ptr + cs] = x[ds + as_:ds + as_ + cs]

    def check(i, f):
        if np.max(np.abs(f), initial=0.0) >= check_tol:
            raise InitError(f"init residual {np.max(np.abs(f)):.3e} exceeds {check_tol} "
                            f"for device {i} ({psd.devices[i].dtype.kind})")

    def op_point(i):
        g = ps.gens[psd.dev_gen[i]]
        b = ps.buses[psd.dev_bus[i]]
        return g.psch, g.qsch, b.v0m, b.v0a

    for i, dev in enumerate(psd.devices):
        d = dev.dtype
        if d.kind != "genrou":
            continue
        if psd.dev_gen[i] < 0:
            raise InitError(f"GENROU device {i} has no static generator")
        pv = p[dev.par_ptr:dev.par_ptr + 14].copy()
        pg, qg, vm, va = op_point(i)
        x, f = _newton(lambda x: _genrou_init_res(x, pv, pg, qg, vm, va),
                       _genrou_init_guess(pv, pg, qg, vm, va))
        check(i, f)
        place(dev, x)
        d.e_fd0 = float(x[10])

    for i, dev in enumerate(psd.devices):
        d = dev.dtype
        if d.kind != "static_gen":
            continue
        ps_ = sum(ps.gens[g].psch for g in d.gen_idxs)
        qs_ = sum(ps.gens[g].qsch for g in d.gen_idxs)
        d.p0, d.q0 = ps_, qs_
        if d.bus_type == PV:
            z[dev.alg_ptr] = qs_
        elif d.bus_type == SLACK:
            z[dev.alg_ptr] = ps_
            z[dev.alg_ptr + 1] = qs_
        p[dev.par_ptr] = ps_
        p[dev.par_ptr + 1] = qs_

    gen_of = {}   # (bus, id) -> the first GENROU there, in device order
    for c in psd.devices:
        if c.dtype.kind == "genrou":
            gen_of.setdefault((c.dtype.bus, _norm_id(c.dtype.id)), c.dtype)

    def matched_gen_efd0(d):
        g = gen_of.get((d.bus, _norm_id(d.id)))
        if g is None:
            raise InitError("exciter without generator")
        return g.e_fd0

    for i, dev in enumerate(psd.devices):
        d = dev.dtype
        if d.kind not in ("ieesgo", "tgov1", "sexs", "esdc1a", "ieeest"):
            continue
        pv = p[dev.par_ptr:dev.par_ptr + d.par_size]
        pg, qg, vm, va = op_point(i)
        if d.kind == "ieesgo":
            K2, K3 = pv[7], pv[8]
            x, f = _newton(lambda x: _ieesgo_init_res(x, pv, pg),
                           [0.0, 0.0, pg, K2 * pg, K2 * K3 * pg, pg, pg])
            check(i, f)
            place(dev, x)
            d.pref = float(x[6]); pv[11] = d.pref
        elif d.kind == "tgov1":
            R, T2, T3 = pv[0], pv[4], pv[5]
            x, f = _newton(lambda x: _tgov1_init_res(x, pv, pg),
                           [(1.0 - T2 / T3) * pg, pg, pg, R * pg])
            check(i, f)
            place(dev, x)
            d.pref = float(x[3]); pv[7] = d.pref
        elif d.kind == "sexs":
            efd0 = matched_gen_efd0(d)
            TA_TB, K = pv[0], pv[2]
            vg = efd0 / K + vm
            x, f = _newton(lambda x: _sexs_init_res(x, pv, vm, efd0),
                           [(1.0 - TA_TB) * (vg - vm), efd0, vg])
            check(i, f)
            place(dev, x)
            d.vref = float(x[2]); pv[6] = d.vref
        elif d.kind == "esdc1a":
            efd0 = matched_gen_efd0(d)
            Ka, Kf, Tf, Ke, sat_a, sat_b = pv[0], pv[2], pv[3], pv[4], pv[7], pv[8]
            sat = 0.0 if (sat_b == 0.0 or efd0 <= sat_a) else sat_b * (efd0 - sat_a) ** 2
            vr1 = Ke * efd0 + sat
            vr2 = -(Kf / Tf) * efd0
            d.vref = vm + vr1 / Ka
            pv[9] = d.vref
            place(dev, np.array([vr1, vr2, efd0]))
        elif d.kind == "ieeest":
            x, f = _newton(lambda x: _ieeest_init_res(x, pv), np.zeros(9))
            check(i, f)
            place(dev, x)
    return dp


def genrou_u0(ps: PowerSystem, dp: DynamicProblem):
    """Constant (e_fd, p_m) per GENROU, used when no exciter/governor is wired:
    the residual then reads the init values left in u (src/dynamics.jl:283-288)."""
    psd = ps.dynamic
    out = []
    for dev in psd.devices:
        if dev.dtype.kind == "genrou":
            out.append((dp.uvec[dev.ctrl_ptr], dp.uvec[dev.ctrl_ptr + 1]))
    return np.array(out, dtype=float).reshape(-1, 2)
