"""Summaries of rocprofv3 CSV output for the Cholesky path (used to write profiles/).

    python -m libddla_amd.prof_summary --trace DIR [--fetch DIR --write DIR]
        [--n 32768 --nb 512] [--focus NAME ...] [--command "..."] [--md OUT.md]
        [--traffic-json OUT.json]

--trace is the output directory of `rocprofv3 --kernel-trace --stats
--output-format csv`; --focus names the trailing-update kernels to report with
their algorithmic rate; --fetch / --write those of `--pmc FETCH_SIZE` and
`--pmc WRITE_SIZE` runs of the same program.  The summary lists every kernel
by total time and, for the trailing update (syrk_update_kernel), its average
launch duration and HBM traffic per launch.  FETCH_SIZE and WRITE_SIZE (KiB
per dispatch) are scaled by what they read for the copy kernel of known size
in the same run (__amd_rocclr_copyBuffer: the largest one moves n^2 doubles
each way), so traffic is reported in bytes.

    python -m libddla_amd.prof_summary --trace DIR --after NAME [--per K] ...
    python -m libddla_amd.prof_summary --trace DIR --span START END ...
    python -m libddla_amd.prof_summary --trace DIR --pmc DIR [DIR ...] --focus NAME ...

summarize one window of the kernel trace instead: --after keeps the
dispatches after the last one whose name contains NAME (through the last one
containing --until, if given); --span keeps those
from the last dispatch containing START that precedes the last one
containing END, through that END dispatch.  The window's wall time (first
start to last end), busy time and per-kernel totals are listed, divided by
--per (e.g. the number of problems in the window).  --pmc lists, for each
--focus kernel, the mean of every counter of the given `rocprofv3 --pmc`
passes per dispatch, with the shader clock, the matrix cores' busy share and
instruction ratios derived from them.
"""
import argparse
import csv
import glob
import json
import os


def _read(pattern):
    files = sorted(glob.glob(pattern))
    if not files:
        raise FileNotFoundError(pattern)
    with open(files[0], newline="") as f:
        return list(csv.DictReader(f))


def short_name(name: str) -> str:
    base = name.replace("void ", "")
    for prefix in ("(anonymous namespace)::", "ddla::kern::", "ddla::", "gemm_detail::", "panel_detail::",
                   "solve_detail::", "batched_detail::"):
        base = base.replace(prefix, "")
    return base.split("(")[0][:90]


def kernel_stats(trace_dir):
    rows = _read(os.path.join(trace_dir, "*kernel_stats.csv"))
    out = []
    for r in rows:
        out.append({"name": r["Name"], "calls": int(r["Calls"]), "total_ms": float(r["TotalDurationNs"]) / 1e6,
                    "avg_us": float(r["AverageNs"]) / 1e3, "pct": float(r["Percentage"])})
    return sorted(out, key=lambda x: -x["total_ms"])


def trace_window(trace_dir, after=None, span=None, until=None):
    """Dispatches of the kernel trace in the window selected by after (and until) / span, in start order."""
    rows = _read(os.path.join(trace_dir, "*kernel_trace.csv"))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    if after:
        hits = [i for i, r in enumerate(rows) if after in r["Kernel_Name"]]
        if not hits:
            raise ValueError(f"no dispatch contains {after!r}")
        last = len(rows)
        if until:
            ends = [i for i, r in enumerate(rows) if until in r["Kernel_Name"] and i > hits[-1]]
            if not ends:
                raise ValueError(f"no dispatch containing {until!r} follows the last {after!r}")
            last = ends[-1] + 1
        return rows[hits[-1] + 1:last]
    start, end = span
    ends = [i for i, r in enumerate(rows) if end in r["Kernel_Name"]]
    if not ends:
        raise ValueError(f"no dispatch contains {end!r}")
    starts = [i for i in range(ends[-1] + 1) if start in rows[i]["Kernel_Name"]]
    if not starts:
        raise ValueError(f"no dispatch containing {start!r} precedes the last {end!r}")
    return rows[starts[-1]:ends[-1] + 1]


def window_lines(rows, per):
    t0 = min(int(r["Start_Timestamp"]) for r in rows)
    t1 = max(int(r["End_Timestamp"]) for r in rows)
    agg = {}
    for r in rows:
        a = agg.setdefault(short_name(r["Kernel_Name"]), [0, 0.0])
        a[0] += 1
        a[1] += (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6
    busy = sum(v[1] for v in agg.values())
    lines = [f"Window: {len(rows)} dispatches, wall {(t1 - t0) / 1e6:.3f} ms, kernel busy {busy:.3f} ms"
             + (f"; per unit (/{per}): wall {(t1 - t0) / 1e6 / per:.3f} ms, busy {busy / per:.3f} ms" if per > 1 else ""),
             "", "| kernel | calls | total ms | avg us | % of busy |" + (" ms per unit |" if per > 1 else ""),
             "|---|---:|---:|---:|---:|" + ("---:|" if per > 1 else "")]
    for name, (calls, ms) in sorted(agg.items(), key=lambda x: -x[1][1])[:20]:
        lines.append(f"| `{name}` | {calls} | {ms:.3f} | {ms / calls * 1e3:.1f} | {100 * ms / busy:.1f} |"
                     + (f" {ms / per:.3f} |" if per > 1 else ""))
    return lines


def counters(pmc_dir, counter):
    rows = _read(os.path.join(pmc_dir, "*counter_collection.csv"))
    return [(r["Kernel_Name"], int(r["Grid_Size"]), float(r["Counter_Value"])) for r in rows
            if r["Counter_Name"] == counter]


def counter_table(pmc_dirs, trace_dir, focus, cus=256, xcds=8):
    """Per focus kernel: the mean of every collected counter per dispatch, and derived rates.

    GRBM_GUI_ACTIVE is summed over the XCDs, so GRBM_GUI_ACTIVE / xcds / (mean traced duration)
    is the shader clock during the kernel; SQ_VALU_MFMA_BUSY_CYCLES counts busy cycles of every
    SIMD (4 per CU), so its share of 4 * cus * GRBM_GUI_ACTIVE / xcds is the fraction of the
    matrix cores kept busy."""
    sums = {}
    for d in pmc_dirs:
        for r in _read(os.path.join(d, "*counter_collection.csv")):
            for f in focus:
                if f in r["Kernel_Name"]:
                    key = (f, short_name(r["Kernel_Name"]))
                    acc = sums.setdefault(key, {})
                    v = acc.setdefault(r["Counter_Name"], [0.0, 0])
                    v[0] += float(r["Counter_Value"])
                    v[1] += 1
    durations = {}
    if trace_dir:
        for st in kernel_stats(trace_dir):
            durations[short_name(st["name"])] = st["avg_us"]
    lines = []
    for (f, name), acc in sorted(sums.items()):
        mean = {c: v[0] / v[1] for c, v in acc.items()}
        lines += ["", f"### `{name}`", "", "| counter | mean per dispatch | dispatches |", "|---|---:|---:|"]
        for c in sorted(mean):
            lines.append(f"| {c} | {mean[c]:.4g} | {acc[c][1]} |")
        grbm = mean.get("GRBM_GUI_ACTIVE")
        derived = []
        if grbm:
            per_xcd = grbm / xcds
            if name in durations and durations[name] > 0:
                derived.append(f"shader clock {per_xcd / (durations[name] * 1e-6) / 1e9:.2f} GHz "
                               f"(GRBM_GUI_ACTIVE / {xcds} over the traced mean {durations[name]:.1f} us)")
            if "SQ_VALU_MFMA_BUSY_CYCLES" in mean:
                derived.append(f"matrix cores busy {mean['SQ_VALU_MFMA_BUSY_CYCLES'] / (4 * cus * per_xcd):.3f} "
                               f"of {4 * cus} SIMDs")
                if "SQ_VALU_MFMA_COEXEC_CYCLES" in mean:
                    derived.append(f"cycles with VALU and MFMA executing together "
                                   f"{mean['SQ_VALU_MFMA_COEXEC_CYCLES']:.0f}")
        if mean.get("SQ_INSTS_MFMA"):
            if "SQ_INSTS_LDS" in mean:
                derived.append(f"LDS instructions per MFMA {mean['SQ_INSTS_LDS'] / mean['SQ_INSTS_MFMA']:.3f}")
            if "SQ_INSTS_VALU" in mean:
                derived.append(f"VALU instructions per MFMA {mean['SQ_INSTS_VALU'] / mean['SQ_INSTS_MFMA']:.3f}")
        if mean.get("SQ_ACTIVE_INST_ANY") and "SQ_WAIT_INST_ANY" in mean:
            derived.append(f"SQ_WAIT_INST_ANY / SQ_ACTIVE_INST_ANY {mean['SQ_WAIT_INST_ANY'] / mean['SQ_ACTIVE_INST_ANY']:.2f}")
        if "SQ_LDS_BANK_CONFLICT" in mean:
            derived.append(f"LDS bank-conflict cycles {mean['SQ_LDS_BANK_CONFLICT']:.0f}")
        if derived:
            lines += ["", "Derived: " + "; ".join(derived) + "."]
    return lines


def traffic(fetch_dir, write_dir, n):
    """(fetch scale, write scale, per-launch fetch bytes, per-launch write bytes) of the trailing update."""
    fetch, write = counters(fetch_dir, "FETCH_SIZE"), counters(write_dir, "WRITE_SIZE")
    copy_bytes = 8.0 * n * n
    cf = max(v for k, g, v in fetch if "copyBuffer" in k)
    cw = max(v for k, g, v in write if "copyBuffer" in k)
    f_scale = copy_bytes / (cf * 1024.0)
    w_scale = copy_bytes / (cw * 1024.0)
    fs = [v for k, g, v in fetch if "syrk_update_kernel" in k]
    ws = [v for k, g, v in write if "syrk_update_kernel" in k]
    return f_scale, w_scale, sum(fs) * 1024.0 * f_scale / len(fs), sum(ws) * 1024.0 * w_scale / len(ws), len(fs)


def rest_updates(n, nb):
    """(M, kb) of the square trailing updates of one factorization: step p applies panel p to every
    column beyond panel p + 2 (the next two panels' columns are look-ahead updates, another kernel)."""
    np_ = (n + nb - 1) // nb
    out = []
    for p in range(np_):
        c = min(n, (p + 3) * nb)
        if c >= n:
            break
        out.append((n - c, min(nb, n - p * nb)))
    return out


def trailing_flops_per_launch(n, nb):
    """Mean algorithmic flops over the square trailing-update launches of one factorization: M (M + 1) kb."""
    ups = rest_updates(n, nb)
    return sum(m * (m + 1.0) * kb for m, kb in ups) / len(ups), len(ups)


def all_updates(n, nb):
    """(flops, launches) of every update of one factorization -- the look-ahead updates of the next two
    panels' columns and the square trailing updates: panel p reaches the lower triangle of every column
    beyond it, M (M + 1) kb flops with M = n - (p + 1) nb."""
    np_ = (n + nb - 1) // nb
    flops, launches = 0.0, 0
    for p in range(np_ - 1):
        m = n - min(n, (p + 1) * nb)
        flops += m * (m + 1.0) * min(nb, n - p * nb)
        launches += 1 + (p + 2 < np_) + (p + 3 < np_)
    return flops, launches


def algorithmic_bytes_per_launch(n, nb):
    """Mean over the square trailing-update launches: the panel rows read once (M x kb) and the lower
    triangle of C read and written."""
    ups = rest_updates(n, nb)
    return sum(8.0 * (m * kb + m * (m + 1.0)) for m, kb in ups) / len(ups), len(ups)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--nb", type=int, default=512)
    ap.add_argument("--command", default="")
    ap.add_argument("--focus", nargs="+", default=["syrk_update_kernel"],
                    help="name substrings of the trailing-update kernels to report")
    ap.add_argument("--focus-all", nargs="+", default=[],
                    help="name substrings of kernels that run every update of the factorization (look-ahead and "
                         "trailing): reported as all update flops over their summed time per factorization")
    ap.add_argument("--md")
    ap.add_argument("--traffic-json")
    ap.add_argument("--after", help="summarize the dispatches after the last one whose name contains this")
    ap.add_argument("--until", help="with --after: end the window at the last dispatch whose name contains this")
    ap.add_argument("--span", nargs=2, metavar=("START", "END"),
                    help="summarize the last START .. END window of dispatches")
    ap.add_argument("--per", type=int, default=1, help="units (problems, calls) in the window")
    ap.add_argument("--title", default="rocprofv3 kernel summary")
    ap.add_argument("--pmc", nargs="+", help="counter-collection output directories (rocprofv3 --pmc passes); "
                                             "summarizes the --focus kernels' counters")
    args = ap.parse_args()

    lines = [f"# {args.title}", ""]
    if args.command:
        lines += ["Command:", "", "```", args.command, "```", ""]
    if args.pmc:
        lines += counter_table(args.pmc, args.trace, args.focus)
        text = "\n".join(lines) + "\n"
        if args.md:
            with open(args.md, "w") as f:
                f.write(text)
        else:
            print(text)
        return
    if args.after or args.span:
        lines += window_lines(trace_window(args.trace, args.after, args.span, args.until), args.per)
        text = "\n".join(lines) + "\n"
        if args.md:
            with open(args.md, "w") as f:
                f.write(text)
        else:
            print(text)
        return
    stats = kernel_stats(args.trace)
    lines += ["| kernel | calls | total ms | avg us | % |", "|---|---:|---:|---:|---:|"]
    for s in stats[:16]:
        lines.append(f"| `{short_name(s['name'])}` | {s['calls']} | {s['total_ms']:.3f} | {s['avg_us']:.1f} | "
                     f"{s['pct']:.2f} |")
    flops, per_fact = trailing_flops_per_launch(args.n, args.nb)
    for focus in args.focus:
        rows = [s for s in stats if focus in s["name"]]
        if not rows:
            continue
        calls = sum(s["calls"] for s in rows)
        avg_ms = sum(s["total_ms"] for s in rows) / calls
        names = ", ".join(f"`{short_name(s['name'])}` ({s['calls']})" for s in rows)
        lines += ["", f"Trailing update {names}: {calls} launches, average {avg_ms:.4f} ms per launch; "
                      f"{per_fact} launches per factorization at n = {args.n}, nb = {args.nb}, mean "
                      f"{flops / 1e9:.1f} GFLOP (FP32- or FP64-equivalent) per launch -> "
                      f"{flops / (avg_ms * 1e-3) / 1e12:.1f} TFLOP/s."]
    all_flops, all_launches = all_updates(args.n, args.nb)
    for focus in args.focus_all:
        for s in [s for s in stats if focus in s["name"]]:
            fact = s["calls"] / all_launches
            per_fact_ms = s["total_ms"] / fact
            lines += ["", f"Updates `{short_name(s['name'])}`: {s['calls']} launches = {fact:.2f} factorizations of "
                          f"{all_launches} launches (look-ahead and trailing) at n = {args.n}, nb = {args.nb}; "
                          f"{per_fact_ms:.2f} ms of summed launch time per factorization for "
                          f"{all_flops / 1e12:.3f} TFLOP (FP32-equivalent) -> at least "
                          f"{all_flops / (per_fact_ms * 1e-3) / 1e12:.1f} TFLOP/s (the look-ahead launches overlap "
                          f"the trailing ones, so summed durations exceed the time they occupy)."]
    result = {}
    if args.fetch and args.write:
        f_scale, w_scale, fb, wb, nl = traffic(args.fetch, args.write, args.n)
        alg, launches = algorithmic_bytes_per_launch(args.n, args.nb)
        result = {"hbm_bytes_per_launch": fb + wb, "fetch_bytes_per_launch": fb, "write_bytes_per_launch": wb,
                  "algorithmic_bytes_per_launch": alg, "launches_counted": nl, "fetch_scale": f_scale,
                  "write_scale": w_scale, "n": args.n, "nb": args.nb}
        lines += ["",
                  f"HBM traffic per trailing-update launch (mean of {nl} dispatches): FETCH {fb / 1e9:.3f} GB, "
                  f"WRITE {wb / 1e9:.3f} GB, total {(fb + wb) / 1e9:.3f} GB; algorithmic "
                  f"{alg / 1e9:.3f} GB (panel read once, lower C read and written).",
                  f"Scale factors from the {8 * args.n * args.n / 1e9:.2f} GB copy in the same run: FETCH_SIZE x "
                  f"{f_scale:.3f}, WRITE_SIZE x {w_scale:.3f}."]
    text = "\n".join(lines) + "\n"
    if args.md:
        with open(args.md, "w") as f:
            f.write(text)
    else:
        print(text)
    if args.traffic_json and result:
        with open(args.traffic_json, "w") as f:
            json.dump(result, f, indent=1)


if __name__ == "__main__":
    main()
