"""Summarizes rocprofv3 output of bench.py runs into profiles/.

Inputs are the CSV directories of four runs of the same command
(`bench.py --steps S --warmup W --no-cpu-baseline --copy-probe`):

    rocprofv3 --kernel-trace --stats --output-format csv -d <kt> -o kt -- python3 bench.py ...
    rocprofv3 --pmc FETCH_SIZE --output-format csv -d <fetch> -o fetch -- python3 bench.py ...
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d <write> -o write -- python3 bench.py ...
    rocprofv3 --pmc TCC_HIT_sum TCC_MISS_sum --output-format csv -d <l2> -o l2 -- python3 bench.py ...

The copy probe (one 256 MiB device-to-device copy at start-up) calibrates the
byte counters: each counter is scaled by (true bytes of the copy) / (what the
counter reads for it).  The timed closest-hit launches on the incoherent rays
are the 2nd .. (1 + W + S)th closest-hit dispatches (the 1st traces the
camera rays).

    python -m sprout_amd.rocprof_summary <kt> <fetch> <write> <l2> --steps S --warmup W \
        --rays N --nodes M --out profiles/trace_closest_c1.md --command "<the rocprofv3 lines>"

The summary ends with "key: value" figures that bench.py reads for the
roofline's `traffic` (only when its workload matches rays and nodes).
"""

from __future__ import annotations

import argparse
import csv
import glob
import os
import statistics

KERNEL = "trace_refill_packed<false, false>"  # closest hit, kernel 3 (kernel 4: <false, true>)
COPY_BYTES = 256 << 20


def _rows(directory: str, suffix: str):
    paths = glob.glob(os.path.join(directory, f"*{suffix}"))
    if not paths:
        raise FileNotFoundError(f"no *{suffix} under {directory}")
    with open(paths[0]) as f:
        return list(csv.DictReader(f))


def _counter(rows, name_part, counter):
    out = []
    for r in rows:
        if name_part in r["Kernel_Name"] and r["Counter_Name"] == counter:
            out.append((int(r["Dispatch_Id"]), float(r["Counter_Value"])))
    out.sort()
    return out


def _copy_kib(rows, counter):
    """Counter value of the 256 MiB calibration copy (the largest copyBuffer)."""
    vals = [v for _, v in _counter(rows, "copyBuffer", counter)]
    return max(vals)


def summarize(kt, fetch, write, l2, steps, warmup, kernel=KERNEL):
    """The timed launches follow one camera-ray launch of the same kernel
    (bench.py with a fixed --shape)."""
    KERNEL = kernel
    timed = slice(1, 1 + warmup + steps)

    trace = [r for r in _rows(kt, "kernel_trace.csv") if KERNEL in r["Kernel_Name"]]
    trace.sort(key=lambda r: int(r["Start_Timestamp"]))
    durations = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6 for r in trace][timed]

    frows = _rows(fetch, "counter_collection.csv")
    wrows = _rows(write, "counter_collection.csv")
    lrows = _rows(l2, "counter_collection.csv")

    fetch_factor = (COPY_BYTES / 1024) / _copy_kib(frows, "FETCH_SIZE")
    write_factor = (COPY_BYTES / 1024) / _copy_kib(wrows, "WRITE_SIZE")

    f = [v for _, v in _counter(frows, KERNEL, "FETCH_SIZE")][timed]
    w = [v for _, v in _counter(wrows, KERNEL, "WRITE_SIZE")][timed]
    hit = [v for _, v in _counter(lrows, KERNEL, "TCC_HIT_sum")][timed]
    miss = [v for _, v in _counter(lrows, KERNEL, "TCC_MISS_sum")][timed]

    grid = int(trace[1]["Grid_Size_X"]) if len(trace) > 1 else 0
    read_bytes = statistics.mean(f) * 1024 * fetch_factor
    write_bytes = statistics.mean(w) * 1024 * write_factor
    stats = {
        "kernel": KERNEL,
        "launches": len(durations),
        "grid_threads": grid,
        "avg_ms": statistics.mean(durations),
        "min_ms": min(durations),
        "max_ms": max(durations),
        "fetch_kib_raw": statistics.mean(f),
        "write_kib_raw": statistics.mean(w),
        "fetch_factor": fetch_factor,
        "write_factor": write_factor,
        "read_bytes": read_bytes,
        "write_bytes": write_bytes,
        "hbm_bytes_per_launch": read_bytes + write_bytes,
        "hbm_gbs": (read_bytes + write_bytes) / (statistics.mean(durations) * 1e-3) / 1e9,
        "l2_hit_rate": sum(hit) / (sum(hit) + sum(miss)),
    }
    return stats


def kernel_table(kt, top=12):
    rows = _rows(kt, "kernel_stats.csv")
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    lines = ["| kernel | calls | avg us | total % |", "|---|---|---|---|"]
    for r in rows[:top]:
        name = r["Name"].replace("mi::(anonymous namespace)::", "").replace("void ", "")
        name = name.split("(")[0]
        lines.append(f"| `{name[:70]}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                     f"{float(r['Percentage']):.1f} |")
    return "\n".join(lines)


def read_figures(path):
    """key: value figures at the end of a summary written by main()."""
    out = {}
    with open(path) as f:
        for line in f:
            if ": " in line and not line.startswith(("|", "#", " ", "`")):
                k, v = line.strip().split(": ", 1)
                try:
                    out[k] = float(v)
                except ValueError:
                    out[k] = v
    return out


def stats_only(argv):
    """Kernel table of one --kernel-trace --stats run:
    python -m sprout_amd.rocprof_summary --stats-only <kt> --title T --out F --command C"""
    ap = argparse.ArgumentParser()
    ap.add_argument("--stats-only", dest="kt", required=True)
    ap.add_argument("--title", required=True)
    ap.add_argument("--out", default="")
    ap.add_argument("--command", default="")
    ap.add_argument("--top", type=int, default=16)
    ap.add_argument("--note", default="")
    args = ap.parse_args(argv)
    text = (f"# rocprofv3: {args.title}\n\n"
            f"Command (run from the repository root on the GPU box, after `cd /tmp && export TMPDIR=/tmp`):\n\n"
            f"```\n{args.command}\n```\n\n"
            f"Kernel statistics of the `--kernel-trace --stats` run:\n\n{kernel_table(args.kt, args.top)}\n"
            + (f"\n{args.note}\n" if args.note else ""))
    if args.out:
        with open(args.out, "w") as f:
            f.write(text)
    print(text)


def _pmc_means(directory, kernel):
    """Mean per dispatch of every counter the run collected for kernels whose
    name contains `kernel` (function name match, ignoring templates)."""
    rows = _rows(directory, "counter_collection.csv")
    values = {}
    for r in rows:
        name = r["Kernel_Name"].replace("mi::(anonymous namespace)::", "").replace("void ", "")
        if name.split("(")[0] != kernel:
            continue
        values.setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    return {k: statistics.mean(v) for k, v in values.items()}, {k: len(v) for k, v in values.items()}


def _copy_factor(directory, counter):
    try:
        return (COPY_BYTES / 1024) / _copy_kib(_rows(directory, "counter_collection.csv"), counter)
    except (ValueError, FileNotFoundError):
        return None


def counters(argv):
    """Per-kernel counters of one workload:
    python -m sprout_amd.rocprof_summary --counters <kt> <pmc dir>... --kernels K... --simds N
        --title T --out F --command C
    Derived: HBM bytes (FETCH_SIZE / WRITE_SIZE scaled by the 256 MiB copy
    probe of the same run), L2 hit rate, MFMA busy fraction
    (SQ_VALU_MFMA_BUSY_CYCLES / (GPU cycles x SIMDs)), VALU busy
    (SQ_ACTIVE_INST_VALU x 4 / (GPU cycles x SIMDs)) and lane utilisation
    (SQ_THREAD_CYCLES_VALU / (SQ_ACTIVE_INST_VALU x 64)).  The GPU cycles
    are GRBM_GUI_ACTIVE / XCCs: the counter reads the sum over the XCCs
    (8 on MI355X), where rocprof's own MfmaUtil takes their maximum."""
    ap = argparse.ArgumentParser()
    ap.add_argument("--counters", dest="dirs", nargs="+", required=True)
    ap.add_argument("--kernels", nargs="+", required=True)
    ap.add_argument("--simds", type=int, default=1024)
    ap.add_argument("--xccs", type=int, default=8)
    ap.add_argument("--title", required=True)
    ap.add_argument("--out", default="")
    ap.add_argument("--command", default="")
    ap.add_argument("--note", default="")
    args = ap.parse_args(argv)
    kt, pmc_dirs = args.dirs[0], args.dirs[1:]
    stats = {r["Name"].replace("mi::(anonymous namespace)::", "").replace("void ", "").split("(")[0]: r
             for r in _rows(kt, "kernel_stats.csv")}
    factors = {}
    for d in pmc_dirs:
        for c in ("FETCH_SIZE", "WRITE_SIZE"):
            f = _copy_factor(d, c)
            if f:
                factors[c] = f
    lines = []
    for k in args.kernels:
        means = {}
        for d in pmc_dirs:
            m, _ = _pmc_means(d, k)
            means.update(m)
        st = stats.get(k)
        avg_us = float(st["AverageNs"]) / 1e3 if st else float("nan")
        lines.append(f"### `{k}`\n")
        lines.append(f"calls: {st['Calls'] if st else 0}; mean duration {avg_us:.1f} us\n")
        lines.append("| counter | mean per dispatch |\n|---|---|")
        for c in sorted(means):
            lines.append(f"| {c} | {means[c]:.6g} |")
        derived = []
        if "FETCH_SIZE" in means and "FETCH_SIZE" in factors:
            b = means["FETCH_SIZE"] * 1024 * factors["FETCH_SIZE"]
            derived.append(f"read beyond L2 {b / 1e6:.2f} MB per dispatch ({b / (avg_us * 1e-6) / 1e9:.0f} GB/s)")
        if "WRITE_SIZE" in means and "WRITE_SIZE" in factors:
            b = means["WRITE_SIZE"] * 1024 * factors["WRITE_SIZE"]
            derived.append(f"written {b / 1e6:.2f} MB per dispatch ({b / (avg_us * 1e-6) / 1e9:.0f} GB/s)")
        if "TCC_HIT_sum" in means and "TCC_MISS_sum" in means:
            t = means["TCC_HIT_sum"] + means["TCC_MISS_sum"]
            if t > 0:
                derived.append(f"L2 hit rate {means['TCC_HIT_sum'] / t:.3f}")
        cycles = means.get("GRBM_GUI_ACTIVE", 0.0) / args.xccs
        if cycles:
            derived.append(f"GPU cycles {cycles:.4g} ({cycles / (avg_us * 1e-6) / 1e9:.2f} GHz over the mean duration)")
        if "SQ_VALU_MFMA_BUSY_CYCLES" in means and cycles:
            derived.append(f"MFMA busy {means['SQ_VALU_MFMA_BUSY_CYCLES'] / (cycles * args.simds):.3f}")
        if "SQ_ACTIVE_INST_VALU" in means and cycles:
            derived.append(f"VALU busy {4 * means['SQ_ACTIVE_INST_VALU'] / (cycles * args.simds):.3f}")
        if "SQ_THREAD_CYCLES_VALU" in means and means.get("SQ_ACTIVE_INST_VALU"):
            derived.append(f"VALU lane utilisation {means['SQ_THREAD_CYCLES_VALU'] / (64 * means['SQ_ACTIVE_INST_VALU']):.3f}")
        if derived:
            lines.append("\n" + "; ".join(derived) + "\n")
        else:
            lines.append("")
    text = (f"# rocprofv3: {args.title}\n\n"
            f"Commands (run from the repository root on the GPU box, after `cd /tmp && export TMPDIR=/tmp`):\n\n"
            f"```\n{args.command}\n```\n\n"
            f"Kernel statistics of the `--kernel-trace --stats` run:\n\n{kernel_table(kt)}\n\n"
            f"Counters per dispatch (means over all dispatches of the kernel; one `--pmc` run per "
            f"group):\n\n" + "\n".join(lines) + (f"\n{args.note}\n" if args.note else ""))
    if args.out:
        with open(args.out, "w") as f:
            f.write(text)
    print(text)


def main():
    import sys

    if "--stats-only" in sys.argv:
        stats_only(sys.argv[1:])
        return
    if "--counters" in sys.argv:
        counters(sys.argv[1:])
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("kt")
    ap.add_argument("fetch")
    ap.add_argument("write")
    ap.add_argument("l2")
    ap.add_argument("--steps", type=int, required=True)
    ap.add_argument("--warmup", type=int, required=True)
    ap.add_argument("--rays", type=int, default=0)
    ap.add_argument("--nodes", type=int, default=0)
    ap.add_argument("--out", default="")
    ap.add_argument("--command", default="")
    ap.add_argument("--title", default="closest-hit traversal, BVH micro-benchmark")
    ap.add_argument("--kernel", default=KERNEL)
    args = ap.parse_args()
    s = summarize(args.kt, args.fetch, args.write, args.l2, args.steps, args.warmup, args.kernel)
    s["num_rays"] = args.rays
    s["num_nodes"] = args.nodes
    figures = "\n".join(f"{k}: {v:.6g}" if isinstance(v, float) else f"{k}: {v}"
                         for k, v in s.items())
    text = (f"# rocprofv3: {args.title}\n\n"
            f"Command (run from the repository root on the GPU box, after `cd /tmp && export TMPDIR=/tmp`):\n\n"
            f"```\n{args.command}\n```\n\n"
            f"Kernel statistics of the `--kernel-trace --stats` run:\n\n{kernel_table(args.kt)}\n\n"
            f"Timed closest-hit launches (incoherent rays): kernel time from the trace, HBM bytes "
            f"from FETCH_SIZE / WRITE_SIZE scaled by the 256 MiB copy probe, L2 hit rate from "
            f"TCC_HIT_sum / TCC_MISS_sum.\n\n{figures}\n")
    if args.out:
        with open(args.out, "w") as f:
            f.write(text)
    print(text)


if __name__ == "__main__":
    main()
