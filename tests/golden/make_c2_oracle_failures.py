"""Writes tests/golden/c2_oracle_failures.json: which scenarios of the
2,000-bus, 256-scenario benchmark workload (synth.case2_grid()) have a Newton
solve that does not converge in the CPU oracle, and after how many completed
steps, under the benchmark's Newton rule (|F|inf < 1e-10, at most 30
corrections per solve).

The oracle restates the reference CPU path: integrate! with newton_step!
(src/dynamics.jl:502-655, src/numerics.jl:55-176) and a KLU-style sparse LU
that pivots afresh at every factorisation.  The GPU test
tests/test_gpu_c2.py compares the library's per-scenario outcome with this
fixture.  Run from the repository root:

    python tests/golden/make_c2_oracle_failures.py [--workers 8]
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
for _p in (ROOT, os.path.join(ROOT, "gradpower.jl_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

TOL, ITERMAX = 1e-10, 30
_CASE = None


def _events(case, m):
    return [(e.bus, e.rfault, int(round(e.ton / case.dt)), int(round(e.toff / case.dt))) for e in case.events[m]]


def _run(m):
    from oracle.oracle import Oracle
    case = _CASE
    o = Oracle(case.desc)
    t0 = time.perf_counter()
    _, _, st = o.integrate(case.dt, int(case.nsteps[m]), _events(case, m), tol=TOL, itermax=ITERMAX, traj=False)
    return m, st, time.perf_counter() - t0


def main():
    global _CASE
    ap = argparse.ArgumentParser()
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--out", default=os.path.join(ROOT, "tests", "golden", "c2_oracle_failures.json"))
    args = ap.parse_args()
    from gradpower_hip import synth
    _CASE = synth.case2_grid()
    with mp.get_context("fork").Pool(args.workers) as pool:
        res = sorted(pool.map(_run, range(_CASE.M)))
    failed, first, rcs = [], {}, {}
    for m, st, _ in res:
        if st["not_converged"] > 0 or st["rc"] != 0:
            failed.append(m)
            first[str(m)] = st["first_fail_step"]
            rcs[str(m)] = st["rc"]
    out = {"workload": "synth.case2_grid() (2,000 buses, M=256)", "newton_tol": TOL, "itermax": ITERMAX,
           "nsteps": [int(x) for x in _CASE.nsteps], "failed": failed, "first_fail_step": first, "rc": rcs,
           "oracle_seconds": sum(r[2] for r in res)}
    with open(args.out, "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"{len(failed)} of {_CASE.M} scenarios fail: {failed}")


if __name__ == "__main__":
    main()
