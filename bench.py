"""Benchmark: contingency-steps/s of the batched backward-Euler integrator on
the 2,000-bus, M=256 uneven-contingency workload (Schur + FP32 factors,
beside the plain FP64 monolithic path and Schur FP64), one MI355X.

A "step" of this benchmark is one pass of the integrator over the whole
batch: all 256 scenarios integrated from t=0 to their own horizons
(U(1,5) s at dt=1/120 s).  value = completed contingency-steps / wall
seconds of that pass, inputs resident on the GPU, system setup and symbolic
analysis excluded (done in warmup).  Prints one JSON line.
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def events_of(case):
    ev = []
    for m in range(case.M):
        for e in case.events[m]:
            ev.append((m, e.bus, e.rfault, int(round(e.ton / case.dt)), int(round(e.toff / case.dt))))
    return ev


def run_method(case, method, steps, warmup, record_idx, tol, refactor_policy, stall_window, theta=0.0,
               async_below=0.0):
    from gradpower_hip.lib import HipContext
    ctx = HipContext(case.desc, case.M, record_idx=record_idx, refactor_policy=refactor_policy,
                     stall_window=stall_window, refactor_theta=theta, async_below=async_below)
    ev = events_of(case)
    times, stats = [], None
    for it in range(warmup + steps):
        ctx.set_state(case.desc.z0, case.desc.p)
        ctx.set_events(ev, case.nsteps)
        t0 = time.perf_counter()
        stats = ctx.integrate(method, case.dt, tol, 30, raise_on_fail=False)
        t1 = time.perf_counter()
        if it >= warmup:
            times.append(t1 - t0)
        log(f"  method {method} pass {it}: {t1 - t0:.3f} s, device {stats['elapsed_ms']:.1f} ms, "
            f"rounds {stats['newton_rounds']} (factor {stats['factor_launches']}), "
            f"corrections {stats['newton_iters']}, refactorizations {stats['refactorizations']}, "
            f"not converged {stats['n_not_converged']}, nan {stats['n_nan']}, status {stats['status']}")
    speeds = None
    if record_idx is not None:
        speeds = [ctx.trajectory(m) for m in range(case.M)]
    status, steps_done = ctx.scenario_status()
    info = ctx.symbolic_info()
    ctx.close()
    return {"times": np.array(times), "stats": stats, "speeds": speeds, "info": info,
            "status": status, "steps_done": steps_done}


def cpu_baseline(case, budget_s=20.0):
    """The CPU oracle (restated reference CPU path, one core) on the first
    scenarios of the same batch, full horizons, until the time budget."""
    from oracle.oracle import Oracle
    o = Oracle(case.desc)
    steps, t_total, done = 0, 0.0, 0
    for m in range(case.M):
        ev = [(e.bus, e.rfault, int(round(e.ton / case.dt)), int(round(e.toff / case.dt))) for e in case.events[m]]
        t0 = time.perf_counter()
        o.integrate(case.dt, int(case.nsteps[m]), ev, tol=1e-10, itermax=50, traj=False)
        t_total += time.perf_counter() - t0
        steps += int(case.nsteps[m])
        done += 1
        if t_total > budget_s:
            break
    return steps / t_total, done, steps, t_total


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nbus", type=int, default=2000)
    ap.add_argument("--M", type=int, default=256)
    ap.add_argument("--no-mono", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--horizon-cap", type=float, default=0.0,
                    help="cap every scenario horizon (s); profiling runs only, not the benchmark workload")
    ap.add_argument("--methods", default="schur32,schur64,mono")
    ap.add_argument("--policy", default="async", choices=["async", "adaptive"],
                    help="refactorisation policy of the Schur methods")
    ap.add_argument("--theta", type=float, default=10.0,
                    help="adaptive refactorisation: refactor when |F| falls by less than this factor")
    ap.add_argument("--async-below", type=float, default=0.0,
                    help="async policy: refactorisations asked for below this |F| run on the second stream")
    ap.add_argument("--stall-window", type=int, default=8,
                    help="stop a scenario whose |F| has not halved in this many Newton iterations (0: run to itermax)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the integrator shards by scenario; "
              f"--gpus {args.gpus} is not supported on this box)", file=sys.stderr)
        sys.exit(2)
    from gradpower_hip import synth
    from gradpower_hip.desc import gen_speed_indices
    from gradpower_hip.lib import (METHOD_MONO_FP64, METHOD_SCHUR_FP64, METHOD_SCHUR_FP32, REFACTOR_ADAPTIVE,
                                   REFACTOR_ASYNC, REFACTOR_EVERY_ITER)

    t0 = time.perf_counter()
    case = synth.case2_grid(M=args.M, nbus=args.nbus)
    if args.horizon_cap > 0:
        case.tf = np.minimum(case.tf, args.horizon_cap)
    log(f"setup {time.perf_counter() - t0:.1f} s: n={case.desc.n} nbus={case.desc.nbus} M={case.M} "
        f"steps/scenario {case.nsteps.min()}..{case.nsteps.max()} load_scale={case.meta.get('load_scale')}")
    total_steps = int(case.nsteps.sum())   # scheduled; the value counts completed steps
    widx = gen_speed_indices(case.desc).astype(np.int32)
    tol = 1e-10

    methods = args.methods.split(",")
    runs = {}
    spol = {"async": REFACTOR_ASYNC, "adaptive": REFACTOR_ADAPTIVE}[args.policy]
    for name, meth, pol in (("schur32", METHOD_SCHUR_FP32, spol),
                            ("schur64", METHOD_SCHUR_FP64, spol),
                            ("mono", METHOD_MONO_FP64, REFACTOR_EVERY_ITER)):
        if name not in methods or (name == "mono" and args.no_mono):
            continue
        nst = args.steps if name != "mono" else max(1, args.steps // 3)
        runs[name] = run_method(case, meth, nst, args.warmup if name != "mono" else min(1, args.warmup),
                                widx if name != "mono" else None, tol, pol, args.stall_window, args.theta,
                                args.async_below)
    head = runs.get("schur32") or next(iter(runs.values()))
    info = head["info"]
    # value: completed contingency-steps per second of the pass (a scenario
    # stopped by a failed Newton solve contributes the steps it completed)
    v = {k: r["stats"]["scenario_steps"] / float(np.median(r["times"])) for k, r in runs.items()}
    out = {
        "metric": "contingency-steps/sec at 2k-bus M=256 (Schur+FP32 LU vs plain FP64)",
        "value": v.get("schur32", next(iter(v.values()))),
        "unit": "contingency-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * float(np.median(head["times"])),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 residual/state, f32 factors",
        "data": "synthetic (seeded random geometric 2,000-bus grid, 256 uneven faults)",
        "config": {"workload": f"{args.nbus}-bus synthetic grid, M={args.M} uneven contingencies, dt=1/120 s, "
                               f"horizons U(1,5) s, Schur + FP32 factors",
                   "n_sys": case.desc.n, "n_red": info["n_red"], "contingency_steps": total_steps,
                   "newton_tol": tol, "itermax": 30, "stall_window": args.stall_window,
                   "refactor_theta": args.theta, "refactor_policy": args.policy,
                   "async_below": args.async_below,
                   "horizon_cap": args.horizon_cap or None},
    }
    for k, r in runs.items():
        st = r["stats"]
        out[k] = {"value": v[k], "ms_per_pass": 1e3 * float(np.median(r["times"])),
                  "completed_steps": st["scenario_steps"], "newton_rounds": st["newton_rounds"],
                  "factor_launches": st["factor_launches"], "newton_corrections": st["newton_iters"],
                  "refactorizations": st["refactorizations"], "status": st["status"],
                  "failed_scenarios": int(np.sum(r["status"] != 0))}
    if "schur32" in runs and "schur64" in runs:
        a, b = runs["schur32"], runs["schur64"]
        ok = [m for m in range(case.M) if a["status"][m] == 0 and b["status"][m] == 0]
        err = max((float(np.max(np.abs(a["speeds"][m] - b["speeds"][m]))) for m in ok), default=None)
        out["schur32"]["rotor_speed_err_vs_fp64_pu"] = err
        out["schur32"]["rotor_speed_err_scenarios"] = len(ok)
    if "schur32" in v and "mono" in v:
        out["speedup_vs_mono_fp64"] = v["schur32"] / v["mono"]
    if not args.no_cpu:
        v, done, steps, tt = cpu_baseline(case, args.cpu_budget)
        out["cpu_baseline"] = {"value": v, "unit": "contingency-steps/s", "cores": 1, "kind": "port",
                               "sample": f"oracle (restated reference CPU path, monolithic sparse LU) on the first "
                                         f"{done} scenarios, full horizons ({steps} steps, {tt:.1f} s)"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
