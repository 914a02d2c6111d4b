"""Benchmark: contingency-steps/s of the batched backward-Euler integrator on
the 2,000-bus, M=256 uneven-contingency workload (Schur + FP32 factors,
beside the plain FP64 monolithic path and Schur FP64), one MI355X.

A "step" of this benchmark is one pass of the integrator over the whole
batch: all 256 scenarios integrated from t=0 to their own horizons
(U(1,5) s at dt=1/120 s).  value = completed contingency-steps / wall
seconds of that pass (median of the timed passes), inputs resident on the
GPU; system setup and symbolic analysis are excluded (done in the warmup
pass).  Prints one JSON line.

The line also carries the roofline of the dominant kernel (the batched
triangular solve or the batched numeric factorisation, whichever has the
larger device time in the headline method), timed with HIP events by the
library around every launch inside the timed passes, and the CPU oracle
timed on a bounded sample of the same workload.

The "adjoint" object measures the discrete adjoint on the 200-bus workload
(40 GENROU+SEXS+TGOV1, M=128 faults, 3 s): one step is one gp_adjoint call
(checkpointed forward sweep, recomputation, transposed solves) giving the
gradient of the batch objective w.r.t. 120 parameters; it carries a central
difference check on a few random parameters and the CPU oracle's adjoint
timed on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "gradpower.jl_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, published
# the library times every k-th launch of the factorisation and of the
# triangular solve (an event pair costs the round's stream a few us)
SOLVE_TIMING_STRIDE = 8
KERNEL_NAMES = {("solve", 2): "k_splu_solve<float>", ("solve", 1): "k_splu_solve<double>",
                ("solve", 0): "k_splu_solve<double>", ("factor", 2): "k_splu_factor_res<float>",
                ("factor", 1): "k_splu_factor<double>", ("factor", 0): "k_splu_factor<double>",
                ("solve", 3): "k_splu_solve<double>", ("factor", 3): "k_splu_factor_res<double>"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def events_of(case):
    ev = []
    for m in range(case.M):
        for e in case.events[m]:
            ev.append((m, e.bus, e.rfault, int(round(e.ton / case.dt)), int(round(e.toff / case.dt))))
    return ev


def run_method(case, method, steps, warmup, record_idx, tol, itermax, opts):
    from gradpower_hip.lib import HipContext
    ctx = HipContext(case.desc, case.M, record_idx=record_idx, **opts)
    ev = events_of(case)
    times, stats, kern = [], None, {"solve_ms": 0.0, "solve_n": 0, "solve_sc": 0, "factor_ms": 0.0,
                                    "factor_n": 0, "factor_sc": 0}
    for it in range(warmup + steps):
        ctx.set_state(case.desc.z0, case.desc.p)
        ctx.set_events(ev, case.nsteps)
        t0 = time.perf_counter()
        stats = ctx.integrate(method, case.dt, tol, itermax, raise_on_fail=False)
        t1 = time.perf_counter()
        if it >= warmup:
            times.append(t1 - t0)
            kern["solve_ms"] += stats["solve_kernel_ms"]
            kern["solve_n"] += stats["solve_kernel_launches"]
            kern["solve_sc"] += stats["solve_kernel_scenarios"]
            kern["factor_ms"] += stats["factor_kernel_ms"]
            kern["factor_n"] += stats["factor_kernel_launches"]
            kern["factor_sc"] += stats["factor_kernel_scenarios"]
        log(f"  method {method} pass {it}: {t1 - t0:.3f} s, device {stats['elapsed_ms']:.1f} ms, "
            f"steps {stats['scenario_steps']}, rounds {stats['newton_rounds']}, exact retries "
            f"{stats['exact_retries']}, repivots {stats['repivots']} "
            f"(factor {stats['factor_launches']}), corrections {stats['newton_iters']}, "
            f"refactorizations {stats['refactorizations']}, not converged {stats['n_not_converged']}, "
            f"solve kernel {stats['solve_kernel_ms']:.1f} ms / {stats['solve_kernel_launches']}, "
            f"factor kernel {stats['factor_kernel_ms']:.1f} ms / {stats['factor_kernel_launches']}")
    speeds = [ctx.trajectory(m) for m in range(case.M)] if record_idx is not None else None
    status, steps_done = ctx.scenario_status()
    info = ctx.symbolic_info()
    ctx.close()
    return {"times": np.array(times), "stats": stats, "speeds": speeds, "info": info,
            "status": status, "steps_done": steps_done, "kern": kern}


_CPU_CASE = None
_CPU_ORACLE = None


def _oracle_scenario(m):
    """One scenario of the CPU baseline: the oracle (the reference CPU path
    restated in C: monolithic Newton, KLU-style sparse LU, one thread) to the
    scenario's horizon or its first Newton solve that does not converge,
    where the library stops a scenario too.  Returns (m, completed steps, s)."""
    global _CPU_ORACLE
    from oracle.oracle import Oracle
    case = _CPU_CASE
    if _CPU_ORACLE is None:
        _CPU_ORACLE = Oracle(case.desc)
    ev = [(e.bus, e.rfault, int(round(e.ton / case.dt)), int(round(e.toff / case.dt))) for e in case.events[m]]
    t0 = time.perf_counter()
    _, _, st = _CPU_ORACLE.integrate(case.dt, int(case.nsteps[m]), ev, tol=1e-10, itermax=30, traj=False,
                                     stop_on_fail=True)
    done = int(case.nsteps[m]) if st["first_fail_step"] < 0 else st["first_fail_step"]
    return m, done, time.perf_counter() - t0


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, budget_s=20.0):
    """Completed contingency-steps/s of the CPU oracle on the GPU box's host:
    (1) one core, the batch's scenarios one after another from the first,
    until the time budget; (2) every host core this job may use (at most
    16), one scenario per worker process, on the first scenarios of the
    batch that hold about budget_s of each worker's work.  Run before the
    GPU is touched (the workers are forked)."""
    import multiprocessing as mp
    global _CPU_CASE
    _CPU_CASE = case
    steps, t1, n1 = 0, 0.0, 0
    for m in range(case.M):
        _, done, t = _oracle_scenario(m)
        steps += done
        t1 += t
        n1 += 1
        log(f"  cpu oracle scenario {m}: {done} of {int(case.nsteps[m])} steps, {t1:.1f} s so far")
        if t1 > budget_s:
            break
    rate1 = steps / t1
    one = {"value": rate1, "unit": "contingency-steps/s", "cores": 1, "kind": "port",
           "sample": f"oracle (restated reference CPU path: monolithic Newton with a KLU-style sparse LU that "
                     f"pivots at every factorisation, 1 thread, -O3) on the first {n1} scenarios of the batch, "
                     f"each to its horizon or its first Newton solve that does not converge ({steps} completed "
                     f"steps, {t1:.1f} s)"}
    workers = max(1, min(16, os.cpu_count() or 1))
    target = workers * rate1 * budget_s   # steps that keep every worker busy about budget_s
    k, acc = 0, 0
    while k < case.M and acc < target:
        acc += int(case.nsteps[k])
        k += 1
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(workers) as pool:
        res = pool.map(_oracle_scenario, range(k), chunksize=1)
    tw = time.perf_counter() - t0
    stepsw = sum(r[1] for r in res)
    log(f"  cpu oracle, {workers} workers: {k} scenarios, {stepsw} completed steps, {tw:.1f} s")
    allc = {"value": stepsw / tw, "unit": "contingency-steps/s", "cores": workers, "kind": "port",
            "sample": f"the same oracle, one scenario per worker process, {workers} workers, on the first {k} "
                      f"scenarios of the batch ({stepsw} completed steps, {tw:.1f} s wall)"}
    return one, allc


def adjoint_params(case):
    """All GENROU H and D and all SEXS K of a case (parameter indices)."""
    d = case.desc.dev
    return ([int(r[2] + 6) for r in d[0]] + [int(r[2] + 7) for r in d[0]] + [int(r[1] + 2) for r in d[3]])


def adjoint_objective(case, p, rows, scale, tol, scen):
    """The adjoint's objective from gp_integrate's trajectories (for the
    central differences): scale * sum_m sum_{k>=1} dt sum_r z_k[r]^2."""
    from gradpower_hip.lib import HipContext, METHOD_MONO_FP64
    ctx = HipContext(case.desc, case.M, record_idx=rows)
    ctx.set_state(case.desc.z0, p)
    ctx.set_events(events_of(case), case.nsteps)
    ctx.integrate(METHOD_MONO_FP64, case.dt, tol, 50, raise_on_fail=False)
    st, _ = ctx.scenario_status()
    if any(st[m] != 0 for m in scen):
        raise RuntimeError("a scenario of the difference quotient did not converge")
    tot = sum(scale * case.dt * float((ctx.trajectory(m)[:, 1:] ** 2).sum()) for m in scen)
    ctx.close()
    return tot


def run_adjoint(steps, warmup, n_fd, cpu_budget, no_cpu):
    """Discrete adjoint on the 200-bus workload: 40 GENROU+SEXS+TGOV1,
    M=128 faults, 3 s at dt=1/120, objective sum_m sum_k dt mean_g w_g^2,
    gradient w.r.t. all H, D and SEXS K (120 parameters), 30-step
    checkpoint segments.  One "step" is one gp_adjoint call over the batch.
    The forward sweep and its recomputation run the Schur method with FP32
    factors (Newton converges to 1e-12 on the FP64 residual, so the
    trajectory, and the gradient, are those of the FP64 methods to 1e-11);
    the transposed solves are FP64."""
    from gradpower_hip import synth
    from gradpower_hip.lib import HipContext, METHOD_SCHUR_FP32
    case = synth.case3_200bus()
    params = adjoint_params(case)
    rows = [int(r[0] + 4) for r in case.desc.dev[0]]
    scale = 1.0 / len(rows)
    tol = 1e-12
    ctx = HipContext(case.desc, case.M)
    times, r = [], None
    for it in range(warmup + steps):
        ctx.set_state(case.desc.z0, case.desc.p)
        ctx.set_events(events_of(case), case.nsteps)
        t0 = time.perf_counter()
        r = ctx.adjoint(METHOD_SCHUR_FP32, case.dt, params, tol=tol, itermax=50, segment=30, scale=scale,
                        raise_on_fail=False)
        t1 = time.perf_counter()
        if it >= warmup:
            times.append(t1 - t0)
        a = r["stats"]
        log(f"  adjoint pass {it}: {t1 - t0:.3f} s, device {a['elapsed_ms']:.1f} ms (forward "
            f"{a['forward_ms']:.1f}, recompute+backward {a['backward_ms']:.1f}), adjoint steps "
            f"{a['adjoint_steps']}, excluded {a['n_excluded']}")
    st, _ = ctx.scenario_status()
    info = ctx.symbolic_info()
    ctx.close()
    ok = [m for m in range(case.M) if st[m] == 0]
    t = float(np.median(times))
    out = {"workload": "200-bus synthetic grid, 40 GENROU+SEXS+TGOV1, M=128 faults, 3 s at dt=1/120, "
                       "objective sum_m sum_k dt mean_g w_g^2, d/d(all H, D, SEXS K) = 120 parameters, "
                       "30-step checkpoint segments, forward and recomputation by the Schur method with FP32 "
                       "factors (FP64 residuals, Newton tol 1e-12), transposed FP64 LU",
           "grads_per_s": 1.0 / t, "unit": "batch gradients/s", "ms_per_gradient": 1e3 * t,
           "scenario_grads_per_s": len(ok) / t, "adjoint_steps_per_s": r["stats"]["adjoint_steps"] / t,
           "forward_ms": r["stats"]["forward_ms"], "backward_ms": r["stats"]["backward_ms"],
           "excluded_scenarios": case.M - len(ok), "n_params": len(params), "newton_tol": tol,
           "dtype": "f64 adjoint and residuals, f32 forward factors", "lu_dense_block_j": info["nd_j"], "lu_dense_block_jt": info["nd_jt"],
           "lu_factor_lds_resident_j": bool(info["res_j"])}
    if n_fd > 0:
        rng = np.random.default_rng(2024)
        worst, p0 = 0.0, case.desc.p.copy()
        for j in rng.choice(len(params), size=n_fd, replace=False):
            q = params[j]
            h = max(1e-6 * abs(p0[q]), 1e-5)
            pp, pm = p0.copy(), p0.copy()
            pp[q] += h
            pm[q] -= h
            fd = (adjoint_objective(case, pp, rows, scale, tol, ok) -
                  adjoint_objective(case, pm, rows, scale, tol, ok)) / (2 * h)
            worst = max(worst, abs(r["grad"][j] - fd) / abs(fd))
            log(f"  fd check p[{q}]: adjoint {r['grad'][j]:.10e} fd {fd:.10e}")
        out["fd_check"] = {"params": int(n_fd), "h": "max(1e-6 |p|, 1e-5)", "max_rel_err": worst}
    if not no_cpu:
        from oracle.oracle import Oracle
        o = Oracle(case.desc)
        tt, done = 0.0, 0
        for m in ok:
            ev = [(e.bus, e.rfault, int(round(e.ton / case.dt)), int(round(e.toff / case.dt)))
                  for e in case.events[m]]
            t0 = time.perf_counter()
            o.adjoint(case.dt, int(case.nsteps[m]), params, ev, rows=rows, scale=scale, tol=tol, itermax=50)
            tt += time.perf_counter() - t0
            done += 1
            log(f"  cpu oracle adjoint scenario {m}: {tt:.1f} s so far")
            if tt > cpu_budget:
                break
        out["cpu_baseline"] = {"value": done / tt, "unit": "scenario gradients/s", "cores": 1, "kind": "port",
                               "sample": f"oracle adjoint (restated reference CPU path, sparse LU of J^T, "
                                         f"residual differences for dF/dp, 1 thread) on {done} scenarios of "
                                         f"the batch, 120 parameters, full 3 s ({tt:.1f} s)"}
    return out


def run_c4(tol, itermax, base, no_cpu, cpu_budget):
    """The 70,000-bus workload (synth.case4_grid): M=8 faults over 2 s, the
    Schur method with FP32 factors and asynchronous refactorisation.  The
    symbolic analysis and the dense-block tuning run in an untimed two-step
    integration first; one full pass is timed.  CPU: the oracle on scenario 0
    for as many steps as the budget allows (one core)."""
    from gradpower_hip import synth
    from gradpower_hip.desc import gen_speed_indices
    from gradpower_hip.lib import HipContext, METHOD_SCHUR_FP32, REFACTOR_ASYNC
    t0 = time.perf_counter()
    case = synth.case4_grid()
    build_s = time.perf_counter() - t0
    ev = events_of(case)
    ctx = HipContext(case.desc, case.M, record_idx=gen_speed_indices(case.desc).astype(np.int32),
                     **dict(base, refactor_policy=REFACTOR_ASYNC))
    ctx.set_state(case.desc.z0, case.desc.p)
    ctx.set_events(ev, np.minimum(case.nsteps, 2))
    ctx.integrate(METHOD_SCHUR_FP32, case.dt, tol, itermax, raise_on_fail=False)
    ctx.set_state(case.desc.z0, case.desc.p)
    ctx.set_events(ev, case.nsteps)
    t1 = time.perf_counter()
    st = ctx.integrate(METHOD_SCHUR_FP32, case.dt, tol, itermax, raise_on_fail=False)
    wall = time.perf_counter() - t1
    status, done = ctx.scenario_status()
    info = ctx.symbolic_info()
    ctx.close()
    log(f"  c4: build {build_s:.1f} s, pass {wall:.2f} s, steps {st['scenario_steps']}, rounds {st['newton_rounds']}, "
        f"solve kernel {st['solve_kernel_ms']:.0f} ms / {st['solve_kernel_launches']}, "
        f"factor kernel {st['factor_kernel_ms']:.0f} ms / {st['factor_kernel_launches']}")
    out = {"workload": f"70,000-bus synthetic grid (35 tiles of the 2,000-bus generator), {case.desc.n} states, "
                       f"M={case.M} faults over 2 s, dt=1/120 s, schur32 with asynchronous refactorisation",
           "value": st["scenario_steps"] / wall, "unit": "contingency-steps/s", "ms_per_pass": 1e3 * wall,
           "completed_steps": st["scenario_steps"], "failed_scenarios": int(np.sum(status != 0)),
           "newton_rounds": st["newton_rounds"], "newton_corrections": st["newton_iters"],
           "n_sys": case.desc.n, "n_red": info["n_red"], "nnz_lu_s": info["nnz_lu_s"],
           "lu_dense_block": info["nd_s"], "solve_in_global_memory": True,
           "solve_kernel_ms_per_launch": st["solve_kernel_ms"] / max(1, st["solve_kernel_launches"]),
           "factor_kernel_ms_per_launch": st["factor_kernel_ms"] / max(1, st["factor_kernel_launches"]),
           "build_s": build_s}
    if not no_cpu:
        from oracle.oracle import Oracle
        o = Oracle(case.desc)
        e0 = [(e.bus, e.rfault, int(round(e.ton / case.dt)), int(round(e.toff / case.dt))) for e in case.events[0]]
        t2 = time.perf_counter()
        ns = 0
        for k in (4, 8, 16, 32, 64):   # grow until the budget is spent
            t3 = time.perf_counter()
            _, _, sto = o.integrate(case.dt, k, e0, tol=tol, itermax=itermax, traj=False)
            ns, el = k, time.perf_counter() - t3
            if sto["first_fail_step"] >= 0:
                ns = sto["first_fail_step"]
            if time.perf_counter() - t2 + 2.2 * el > cpu_budget:
                break
        out["cpu_baseline"] = {"value": ns / el, "unit": "contingency-steps/s", "cores": 1, "kind": "port",
                               "sample": f"the oracle on scenario 0, its first {ns} steps (fault on at step 2)"}
        out["vs_cpu_one_core"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def roofline(run, method):
    """Dominant kernel of one method: HIP-event device time per launch and
    algorithmic bytes per launch (bytes per scenario x scenarios served)."""
    k, st = run["kern"], run["stats"]
    stride = {"solve": SOLVE_TIMING_STRIDE, "factor": SOLVE_TIMING_STRIDE}
    which = "solve" if k["solve_ms"] >= k["factor_ms"] else "factor"
    n = k[which + "_n"]
    if n == 0:
        return None
    avg_s = k[which + "_ms"] / n * 1e-3
    per_scen = st["solve_bytes_per_scenario" if which == "solve" else "factor_bytes_per_scenario"]
    bytes_per_launch = per_scen * k[which + "_sc"] / n
    achieved = bytes_per_launch / avg_s / 1e9
    name = KERNEL_NAMES[(which, method)]
    if which == "factor" and method != 2 and run["info"]["res_j" if method == 0 else "res_s"]:
        name = "k_splu_factor_res<double>"   # FP64 factorisation with its multipliers in LDS
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_" + name.replace("<", "_").replace(">", "") + ".json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            tj = json.load(fh)
        # the counter runs' launches serve another mix of scenarios: scale
        # their measured bytes per scenario to this run's launches
        t_scen = tj.get("traffic_bytes_per_scenario")
        traffic = t_scen * k[which + "_sc"] / n if t_scen else tj.get("traffic_bytes_per_launch")
    return {"kernel": name, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "avg_launch_us": avg_s * 1e6,
            "launches": n, "algorithmic_bytes_per_launch": bytes_per_launch,
            "scenarios_per_launch": k[which + "_sc"] / n, "bytes_per_scenario": per_scen,
            "share_of_wall": k[which + "_ms"] * stride[which] * 1e-3 / float(np.sum(run["times"])),
            "timed_launches": f"every {stride[which]}-th" if stride[which] > 1 else "every"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nbus", type=int, default=2000)
    ap.add_argument("--M", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--horizon-cap", type=float, default=0.0,
                    help="cap every scenario horizon (s); profiling runs only, not the benchmark workload")
    ap.add_argument("--methods", default="schur32,schur64,mono,shared")
    ap.add_argument("--policy", default="async", choices=["async", "adaptive"],
                    help="refactorisation policy of the Schur methods")
    ap.add_argument("--theta", type=float, default=30.0,
                    help="refactorise when |F| falls by less than this factor in an iteration")
    ap.add_argument("--async-below", type=float, default=0.0,
                    help="async policy: refactorisations asked for below this |F| run on the second stream")
    ap.add_argument("--stall-window", type=int, default=0,
                    help="stop a scenario whose |F| has not halved in this many Newton iterations "
                         "(0, the reference's rule: run to itermax)")
    ap.add_argument("--retry-window", type=int, default=0,
                    help="corrections without halving |F| after which an inexact solve restarts as exact FP64 "
                         "Newton (0: the library's default, 8; -1: at itermax only)")
    ap.add_argument("--predictor", type=int, default=0,
                    help="first iterate of a step: 0 z_k (the reference's start), 1 linear extrapolation, "
                         "2 quadratic (under the Schur methods' policies)")
    ap.add_argument("--solve-levels", type=int, default=0,
                    help="levels per partition of the LU plans' triangular solves (0 default, 1 unpartitioned)")
    ap.add_argument("--no-profile", action="store_true", help="do not time kernels with HIP events")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the discrete-adjoint measurement")
    ap.add_argument("--no-c4", action="store_true", help="skip the 70,000-bus M=8 measurement")
    ap.add_argument("--adjoint-fd", type=int, default=4,
                    help="parameters checked against central differences in the adjoint measurement")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the integrator shards by scenario; "
              f"--gpus {args.gpus} is not supported on this box)", file=sys.stderr)
        sys.exit(2)
    from gradpower_hip import synth
    from gradpower_hip.desc import gen_speed_indices
    from gradpower_hip.lib import (METHOD_MONO_FP64, METHOD_SCHUR_FP64, METHOD_SCHUR_FP32, METHOD_SCHUR_SHARED,
                                   REFACTOR_ADAPTIVE, REFACTOR_ASYNC, REFACTOR_EVERY_ITER)

    t0 = time.perf_counter()
    case = synth.case2_grid(M=args.M, nbus=args.nbus)
    if args.horizon_cap > 0:
        case.tf = np.minimum(case.tf, args.horizon_cap)
    log(f"setup {time.perf_counter() - t0:.1f} s: n={case.desc.n} nbus={case.desc.nbus} M={case.M} "
        f"steps/scenario {case.nsteps.min()}..{case.nsteps.max()} load_scale={case.meta.get('load_scale')}")
    total_steps = int(case.nsteps.sum())   # scheduled; the value counts completed steps
    widx = gen_speed_indices(case.desc).astype(np.int32)
    tol, itermax = 1e-10, 30
    cpu = None
    if not args.no_cpu:
        log(f"cpu oracle ({cpu_model()}, {os.cpu_count()} logical cpus)")
        cpu = cpu_baseline(case, args.cpu_budget)

    spol = {"async": REFACTOR_ASYNC, "adaptive": REFACTOR_ADAPTIVE}[args.policy]
    base = {"stall_window": args.stall_window, "refactor_theta": args.theta, "async_below": args.async_below,
            "profile": 0 if args.no_profile else SOLVE_TIMING_STRIDE, "retry_window": args.retry_window,
            "predictor": args.predictor, "solve_levels": args.solve_levels}
    methods = args.methods.split(",")
    runs, mcode = {}, {}
    for name, meth, pol in (("schur32", METHOD_SCHUR_FP32, spol), ("schur64", METHOD_SCHUR_FP64, spol),
                            ("mono", METHOD_MONO_FP64, REFACTOR_EVERY_ITER),
                            ("shared", METHOD_SCHUR_SHARED, spol)):
        if name not in methods:
            continue
        log(f"method {name}")
        runs[name] = run_method(case, meth, args.steps, args.warmup, widx if name != "mono" else None, tol,
                                itermax, dict(base, refactor_policy=pol))
        mcode[name] = meth
    head_name = "schur32" if "schur32" in runs else next(iter(runs))
    head = runs[head_name]
    info = head["info"]
    v = {k: r["stats"]["scenario_steps"] / float(np.median(r["times"])) for k, r in runs.items()}
    out = {
        "metric": "contingency-steps/sec at 2k-bus M=256 (Schur+FP32 LU vs plain FP64); adjoint grads/s",
        "value": v[head_name],
        "unit": "contingency-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * float(np.median(head["times"])),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 residual/state, f32 factors" if head_name == "schur32" else "f64",
        "data": "synthetic (seeded random geometric 2,000-bus grid, 256 uneven faults)",
        "config": {"workload": f"{args.nbus}-bus synthetic grid, M={args.M} uneven contingencies, dt=1/120 s, "
                               f"horizons U(1,5) s, {head_name}",
                   "n_sys": case.desc.n, "n_red": info["n_red"], "nnz_lu_s": info["nnz_lu_s"],
                   "nnz_lu_j": info["nnz_lu_j"], "contingency_steps_scheduled": total_steps,
                   "contingency_steps_completed": int(head["stats"]["scenario_steps"]),
                   "scenarios_stopped": f"{int(np.sum(head['status'] != 0))} of {case.M}",
                   "newton_tol": tol, "itermax": itermax, "stall_window": args.stall_window,
                   "retry_window": args.retry_window or 8,
                   "refactor_policy": args.policy, "refactor_theta": args.theta, "predictor": args.predictor,
                   "solve_levels": args.solve_levels or "autotuned",
                   "horizon_cap": args.horizon_cap or None,
                   "deviations": {
                       "loads": f"constant impedance (ZIP alpha={case.meta['zip_alpha']}), not the 0.4/0.3/0.3 "
                                "Z/I/P split of the 9-bus configuration",
                       "fault_buses": f"uniform over the {case.meta['fault_bus_pool']} of {args.nbus} buses at least "
                                      "3 hops from every static generator (PV stub), not over all buses",
                       "load_scale": case.meta["load_scale"],
                       "machine_mva_base": "sized to at most 80% loading at the solved operating point"}},
    }
    rl = roofline(head, mcode[head_name]) if not args.no_profile else None
    if rl:
        out["roofline"] = rl
    for k, r in runs.items():
        st = r["stats"]
        out[k] = {"value": v[k], "ms_per_pass": 1e3 * float(np.median(r["times"])),
                  "completed_steps": st["scenario_steps"], "newton_rounds": st["newton_rounds"],
                  "factor_launches": st["factor_launches"], "newton_corrections": st["newton_iters"],
                  "corrections_per_step": st["newton_iters"] / max(1, st["scenario_steps"]),
                  "refactorizations": st["refactorizations"], "exact_retries": st["exact_retries"],
                  "repivots": st["repivots"], "status": st["status"],
                  "failed_scenarios": int(np.sum(r["status"] != 0)),
                  "lu_dense_block": r["info"]["nd_j" if k == "mono" else "nd_s"],
                  "lu_solve_levels": r["info"]["solve_levels_j" if k == "mono" else "solve_levels_s"],
                  "lu_factor_lds_resident": bool(r["info"]["res_j" if k == "mono" else "res_s"])}
        if not args.no_profile and k != head_name:
            out[k]["roofline"] = roofline(r, mcode[k])
    if "schur32" in runs and "schur64" in runs:
        a, b = runs["schur32"], runs["schur64"]
        ok = [m for m in range(case.M) if a["status"][m] == 0 and b["status"][m] == 0]
        err = max((float(np.max(np.abs(a["speeds"][m] - b["speeds"][m]))) for m in ok), default=None)
        out["schur32"]["rotor_speed_err_vs_fp64_pu"] = err
        out["schur32"]["rotor_speed_err_scenarios"] = len(ok)
        out["schur32"]["extra_corrections_vs_fp64"] = (a["stats"]["newton_iters"] - b["stats"]["newton_iters"])
    if "shared" in runs and "schur64" in runs:
        a, b = runs["shared"], runs["schur64"]
        ok = [m for m in range(case.M) if a["status"][m] == 0 and b["status"][m] == 0]
        err = max((float(np.max(np.abs(a["speeds"][m] - b["speeds"][m]))) for m in ok), default=None)
        out["shared"]["rotor_speed_err_vs_fp64_pu"] = err
        out["shared"]["rotor_speed_err_scenarios"] = len(ok)
    if "schur32" in v and "mono" in v:
        out["speedup_vs_mono_fp64"] = v["schur32"] / v["mono"]
    if cpu is not None:
        one, allc = cpu
        out["cpu_baseline"] = dict(allc, cpu_model=cpu_model(), one_core=one)
    if not args.no_adjoint:
        log("adjoint (200-bus, M=128)")
        out["adjoint"] = run_adjoint(args.steps, args.warmup, args.adjoint_fd, args.cpu_budget / 2, args.no_cpu)
    if not args.no_c4:
        log("70,000-bus grid, M=8")
        out["c4"] = run_c4(tol, itermax, base, args.no_cpu, args.cpu_budget / 2)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
