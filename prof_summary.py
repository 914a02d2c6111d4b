"""Summaries of rocprofv3 CSV output for the profiles/ folder.

    python3 prof_summary.py kernels TRACE_DIR FETCH_DIR WRITE_DIR COPY_BYTES OUT.md --cmd "..."
        every kernel of the measured launches (those after the marker copy
        kernel k_copy_plain, which reads and writes COPY_BYTES): calls,
        average duration from a --kernel-trace run, HBM bytes per launch from
        a FETCH_SIZE run and a WRITE_SIZE run (each scaled by what the counter
        reads for the copy kernel), achieved GB/s and its fraction of the
        MI355X's 8 TB/s; and, per hardware queue, how much of the measured
        span had a kernel running

    python3 prof_summary.py stats DIR OUT.md --cmd "..."
        per-kernel table (calls, total ms, average us, share) from *kernel_stats.csv

    python3 prof_summary.py traffic FETCH_DIR WRITE_DIR KERNEL COPY_BYTES OUT.json --cmd "..." \
            [--scenarios SCENARIOS_PER_LAUNCH] [--last N]
        HBM bytes per launch of KERNEL from a FETCH_SIZE run and a WRITE_SIZE run
        (KiB per dispatch), each scaled by what the counter reads for the plain
        copy kernel k_copy_plain, which reads and writes COPY_BYTES; with the
        mean number of scenarios a launch served in those runs, also per scenario;
        with --last N over the last N dispatches only (the workload's own)
"""
import csv
import glob
import json
import os
import sys


def _rows(d, pattern):
    files = sorted(glob.glob(os.path.join(d, "**", pattern), recursive=True))
    if not files:
        raise SystemExit(f"no {pattern} under {d}")
    out = []
    for f in files:
        with open(f, newline="") as fh:
            out.extend(csv.DictReader(fh))
    return out


def _col(row, *names):
    low = {k.lower(): k for k in row}
    for n in names:
        if n.lower() in low:
            return row[low[n.lower()]]
    raise KeyError(f"none of {names} in {list(row)}")


def short(name):
    s = name.replace("void ", "").replace("gp::", "")
    return s.split("(")[0]


def stats(d, out, cmd):
    rows = _rows(d, "*kernel_stats.csv")
    tab = []
    for r in rows:
        tot = float(_col(r, "TotalDurationNs", "Total_Duration_Ns", "TotalDuration"))
        tab.append((short(_col(r, "Name", "Kernel_Name", "KernelName")), int(_col(r, "Calls")), tot,
                    float(_col(r, "AverageNs", "Average_Ns", "Average")), float(_col(r, "Percentage"))))
    tab.sort(key=lambda t: -t[2])
    with open(out, "w") as fh:
        fh.write(f"# Kernel time summary\n\nCommand (from the repository root, under `cd /tmp && export TMPDIR=/tmp`):\n\n"
                 f"```\n{cmd}\n```\n\n| kernel | calls | total ms | average us | share % |\n|---|---:|---:|---:|---:|\n")
        for name, calls, tot, avg, pct in tab[:25]:
            fh.write(f"| `{name}` | {calls} | {tot / 1e6:.2f} | {avg / 1e3:.2f} | {pct:.2f} |\n")
    print(open(out).read())


def _counter(d, kernel, counter, last=None):
    vals = []
    for r in _rows(d, "*counter_collection.csv"):
        if kernel in _col(r, "Kernel_Name", "KernelName", "Name") and \
                _col(r, "Counter_Name", "CounterName") == counter:
            vals.append((int(_col(r, "Dispatch_Id", "DispatchId")), float(_col(r, "Counter_Value", "CounterValue"))))
    if not vals:
        raise SystemExit(f"no {counter} samples for {kernel} under {d}")
    vals.sort()
    if last:
        vals = vals[-int(last):]   # the workload's own launches (earlier ones: the plan autotuner's)
    return sum(v for _, v in vals) / len(vals), len(vals)


def traffic(fdir, wdir, kernel, copy_bytes, out, cmd, scen=None, last=None):
    copy_bytes = float(copy_bytes)
    kf, nkf = _counter(fdir, kernel, "FETCH_SIZE", last)
    kw, nkw = _counter(wdir, kernel, "WRITE_SIZE", last)
    cf, _ = _counter(fdir, "k_copy_plain", "FETCH_SIZE")
    cw, _ = _counter(wdir, "k_copy_plain", "WRITE_SIZE")
    sf, sw = copy_bytes / cf, copy_bytes / cw   # bytes per counter unit
    res = {"kernel": kernel, "fetch_size_per_dispatch": kf, "write_size_per_dispatch": kw,
           "dispatches": [nkf, nkw], "copy_bytes": copy_bytes, "copy_fetch_size": cf, "copy_write_size": cw,
           "fetch_scale_bytes_per_unit": sf, "write_scale_bytes_per_unit": sw,
           "fetch_bytes_per_launch": kf * sf, "write_bytes_per_launch": kw * sw,
           "traffic_bytes_per_launch": kf * sf + kw * sw, "command": cmd}
    if scen:
        res["scenarios_per_launch"] = float(scen)
        res["traffic_bytes_per_scenario"] = (kf * sf + kw * sw) / float(scen)
    with open(out, "w") as fh:
        json.dump(res, fh, indent=1)
    print(json.dumps(res, indent=1))


HBM_PEAK_GBS = 8000.0


def _after_marker(rows, name_cols, id_cols):
    """Rows after the last dispatch of the marker kernel k_copy_plain."""
    rows = sorted(rows, key=lambda r: int(_col(r, *id_cols)))
    mk = [i for i, r in enumerate(rows) if "k_copy_plain" in _col(r, *name_cols)]
    if not mk:
        raise SystemExit("no k_copy_plain marker dispatch")
    return rows[mk[-1]], rows[mk[-1] + 1:]


def _per_kernel_counter(d, counter):
    rows = [r for r in _rows(d, "*counter_collection.csv") if _col(r, "Counter_Name", "CounterName") == counter]
    marker, rows = _after_marker(rows, ("Kernel_Name", "KernelName"), ("Dispatch_Id", "DispatchId"))
    acc = {}
    for r in rows:
        k = short(_col(r, "Kernel_Name", "KernelName"))
        v = acc.setdefault(k, [0.0, 0])
        v[0] += float(_col(r, "Counter_Value", "CounterValue"))
        v[1] += 1
    return float(_col(marker, "Counter_Value", "CounterValue")), {k: v[0] / v[1] for k, v in acc.items()}


def kernels(tdir, fdir, wdir, copy_bytes, out, cmd):
    copy_bytes = float(copy_bytes)
    trace = _rows(tdir, "*kernel_trace.csv")
    _, trace = _after_marker(trace, ("Kernel_Name",), ("Dispatch_Id",))
    tab, queues = {}, {}
    t0 = min(int(r["Start_Timestamp"]) for r in trace)
    t1 = max(int(r["End_Timestamp"]) for r in trace)
    for r in trace:
        k = short(r["Kernel_Name"])
        a, b = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
        v = tab.setdefault(k, [0, 0.0])
        v[0] += 1
        v[1] += b - a
        queues.setdefault(r["Queue_Id"], []).append((a, b))
    if fdir == "-":   # kernel times only
        cf = cw = copy_bytes
        fetch, write = {}, {}
    else:
        cf, fetch = _per_kernel_counter(fdir, "FETCH_SIZE")
        cw, write = _per_kernel_counter(wdir, "WRITE_SIZE")
    sf, sw = copy_bytes / cf, copy_bytes / cw
    ktot = sum(v[1] for v in tab.values())
    lines = []
    for k, (calls, tot) in sorted(tab.items(), key=lambda kv: -kv[1][1]):
        avg = tot / calls
        fb = fetch.get(k, 0.0) * sf
        wb = write.get(k, 0.0) * sw
        gbs = (fb + wb) / avg if avg > 0 else 0.0   # bytes per ns = GB/s
        row = f"| `{k}` | {calls} | {tot / 1e6:.2f} | {100 * tot / ktot:.1f} | {avg / 1e3:.2f} |"
        if fdir != "-":
            row += f" {fb / 1e6:.3f} | {wb / 1e6:.3f} | {gbs:.1f} | {gbs / HBM_PEAK_GBS:.4f} |"
        lines.append(row)
    qlines = []
    for q, iv in sorted(queues.items()):
        iv.sort()
        busy, cs, ce = 0, None, None
        for a, b in iv:
            if cs is None or a > ce:
                if cs is not None:
                    busy += ce - cs
                cs, ce = a, b
            else:
                ce = max(ce, b)
        busy += ce - cs
        qlines.append(f"| {q} | {len(iv)} | {busy / 1e6:.2f} | {100 * busy / (t1 - t0):.1f} |")
    with open(out, "w") as fh:
        if fdir == "-":
            head = (f"# Per-kernel time\n\nCommand (the measured launches are those after the marker copy kernel):"
                    "\n\n```\n" + cmd + f"\n```\n\nMeasured span {(t1 - t0) / 1e6:.2f} ms.\n\n"
                    "| kernel | calls | total ms | share % | avg us |\n|---|---:|---:|---:|---:|\n")
        else:
            head = ("# Per-kernel time and HBM traffic\n\nCommands (from the repository root; the measured launches "
                    "are those after the marker copy kernel):\n\n```\n" + cmd + "\n```\n\n"
                    f"Counter scale (copy kernel of {copy_bytes / 1e6:.0f} MB read and written): FETCH_SIZE "
                    f"{sf:.1f} B per unit, WRITE_SIZE {sw:.1f} B per unit.  GB/s = (fetched + written bytes per "
                    f"launch) / average launch duration; peak {HBM_PEAK_GBS:.0f} GB/s.  Measured span "
                    f"{(t1 - t0) / 1e6:.2f} ms.\n\n"
                    "| kernel | calls | total ms | share % | avg us | fetched MB/launch | written MB/launch | GB/s | "
                    "frac of peak |\n|---|---:|---:|---:|---:|---:|---:|---:|---:|\n")
        fh.write(head + "\n".join(lines) +
                 "\n\nKernel occupancy of each hardware queue over the measured span (the context's stream and the "
                 "asynchronous factorisation's stream):\n\n| queue | launches | busy ms | busy % of span |\n"
                 "|---|---:|---:|---:|\n" + "\n".join(qlines) + "\n")
    print(open(out).read())


if __name__ == "__main__":
    a = sys.argv[1:]
    cmd, scen = "", None
    if "--cmd" in a:
        i = a.index("--cmd")
        cmd = a[i + 1]
        a = a[:i] + a[i + 2:]
    if "--scenarios" in a:
        i = a.index("--scenarios")
        scen = a[i + 1]
        a = a[:i] + a[i + 2:]
    last = None
    if "--last" in a:
        i = a.index("--last")
        last = a[i + 1]
        a = a[:i] + a[i + 2:]
    if a[0] == "kernels":
        kernels(a[1], a[2], a[3], a[4], a[5], cmd)
    elif a[0] == "stats":
        stats(a[1], a[2], cmd)
    elif a[0] == "traffic":
        traffic(a[1], a[2], a[3], a[4], a[5], cmd, scen, last)
    else:
        raise SystemExit(__doc__)
