"""Workload driver for hardware-counter and kernel-trace runs (rocprofv3):
one warm-up pass of the 2,000-bus M=256 contingency workload (or, with
--case c4, of the 70,000-bus M=8 one) -- symbolic analysis, LU-plan
autotuning -- then a plain copy kernel of a known size
(which calibrates the byte counters and marks where the measured launches
begin), then the measured pass(es), optionally with the scenario horizons
capped so that a counter run stays short; or (--adjoint) one
discrete-adjoint pass of the 200-bus M=128 3 s workload.

    rocprofv3 --kernel-trace --stats --output-format csv -d OUT -o run -- \\
        python3 prof_workload.py --horizon-cap 0.5
"""
import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "gradpower.jl_amd"))

import numpy as np  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-mb", type=float, default=256.0)
    ap.add_argument("--horizon-cap", type=float, default=0.5)
    ap.add_argument("--method", default="schur32", choices=["schur32", "schur64", "mono"])
    ap.add_argument("--case", default="c2", choices=["c2", "c4"],
                    help="c2: the 2,000-bus M=256 workload; c4: the 70,000-bus M=8 workload")
    ap.add_argument("--passes", type=int, default=1)
    ap.add_argument("--stats-out", default="",
                    help="write the batched LU launches' scenario counts of the last pass here (JSON)")
    ap.add_argument("--adjoint", action="store_true",
                    help="one gp_adjoint pass of the 200-bus, M=128, 3 s workload instead")
    args = ap.parse_args()
    if args.adjoint:
        return adjoint(args.passes)
    import ctypes as C
    from gradpower_hip import synth
    from gradpower_hip.lib import (HipContext, METHOD_MONO_FP64, METHOD_SCHUR_FP32, METHOD_SCHUR_FP64,
                                   REFACTOR_ASYNC, REFACTOR_EVERY_ITER)
    t0 = time.perf_counter()
    case = synth.case2_grid(M=256) if args.case == "c2" else synth.case4_grid()
    if args.horizon_cap > 0:   # 0: the full horizons
        case.tf = np.minimum(case.tf, args.horizon_cap)
    print(f"setup {time.perf_counter() - t0:.1f} s", flush=True)
    meth = {"schur32": METHOD_SCHUR_FP32, "schur64": METHOD_SCHUR_FP64, "mono": METHOD_MONO_FP64}[args.method]
    pol = REFACTOR_EVERY_ITER if args.method == "mono" else REFACTOR_ASYNC
    ctx = HipContext(case.desc, case.M, record_idx=np.zeros(1, np.int32), refactor_policy=pol, profile=8)
    ev = [(m, e.bus, e.rfault, int(round(e.ton / case.dt)), int(round(e.toff / case.dt)))
          for m in range(case.M) for e in case.events[m]]
    ctx.set_state(case.desc.z0, case.desc.p)
    ctx.set_events(ev, case.nsteps)
    st = ctx.integrate(meth, case.dt, 1e-10, 30, raise_on_fail=False)
    print(f"warm-up pass: {st['scenario_steps']} steps, {st['newton_rounds']} rounds", flush=True)
    # calibration copy (and marker): n doubles device to device through the library's copy kernel
    n = int(args.copy_mb * 1e6 / 8)
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(8 * n)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(8 * n)) == 0
    assert hip.hipMemset(src, 0, C.c_size_t(8 * n)) == 0
    rc = ctx.lib.gp_copy_kernel(ctx.h, src, dst, n)
    assert rc == 0, ctx.lib.gp_last_error()
    print(f"copy kernel: {8 * n} bytes read, {8 * n} bytes written", flush=True)
    hip.hipFree(src)
    hip.hipFree(dst)
    for it in range(args.passes):
        ctx.set_state(case.desc.z0, case.desc.p)
        ctx.set_events(ev, case.nsteps)
        st = ctx.integrate(meth, case.dt, 1e-10, 30, raise_on_fail=False)
        print(f"pass {it}: {st['scenario_steps']} steps, {st['newton_rounds']} rounds, "
              f"{st['elapsed_ms']:.1f} ms", flush=True)
    if args.stats_out:
        import json
        with open(args.stats_out, "w") as fh:
            json.dump({"solve_launches": st["solve_kernel_launches"], "solve_scenarios": st["solve_kernel_scenarios"],
                       "factor_launches": st["factor_kernel_launches"],
                       "factor_scenarios": st["factor_kernel_scenarios"]}, fh)
    ctx.close()


def adjoint(passes):
    """The bench's adjoint: Schur forward with FP32 factors, FP64 transposed solves."""
    from gradpower_hip import synth
    from gradpower_hip.lib import HipContext, METHOD_SCHUR_FP32
    case = synth.case3_200bus()
    d = case.desc.dev
    params = [int(r[2] + 6) for r in d[0]] + [int(r[2] + 7) for r in d[0]] + [int(r[1] + 2) for r in d[3]]
    ev = [(m, e.bus, e.rfault, int(round(e.ton / case.dt)), int(round(e.toff / case.dt)))
          for m in range(case.M) for e in case.events[m]]
    ctx = HipContext(case.desc, case.M)
    for it in range(passes):
        ctx.set_state(case.desc.z0, case.desc.p)
        ctx.set_events(ev, case.nsteps)
        r = ctx.adjoint(METHOD_SCHUR_FP32, case.dt, params, tol=1e-12, itermax=50, segment=30,
                        scale=1.0 / len(d[0]), raise_on_fail=False)
        a = r["stats"]
        print(f"adjoint pass {it}: {a['elapsed_ms']:.1f} ms (forward {a['forward_ms']:.1f}, "
              f"recompute+backward {a['backward_ms']:.1f}), adjoint steps {a['adjoint_steps']}", flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
