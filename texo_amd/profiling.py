"""Turns rocprofv3 output of the benchmark into the committed summaries.

The counter runs and the kernel-statistics run are plain rocprofv3 commands
on bench.py (listed in each summary); this module only reads their CSV files:

    python -m texo_amd.profiling STATS_DIR BENCH_JSON PMC_FETCH_DIR PMC_WRITE_DIR \\
        CAL_FETCH_DIR CAL_WRITE_DIR COPY_BYTES BATCH OUT_DIR

writes OUT_DIR/<name>.md (kernel statistics with the bench line's roofline
for comparison) and OUT_DIR/traffic.json (HBM bytes per launch per kernel);

    python -m texo_amd.profiling --stats STATS_DIR RESULT_FILE OUT_MD TITLE COMMAND

writes the kernel statistics of any other profiled run (a serving stream,
the incremental stream) with the JSON result line the run printed.
"""
from __future__ import annotations

import csv
import glob
import json
import os
import re
import sys
from collections import defaultdict


def _csv(directory: str, suffix: str):
    files = glob.glob(os.path.join(directory, "**", f"*{suffix}"), recursive=True)
    if not files:
        raise FileNotFoundError(f"no *{suffix} under {directory}")
    with open(files[0], newline="") as f:
        return list(csv.DictReader(f))


def kernel_name(full: str) -> str:
    m = re.search(r"(k_[a-z0-9_]+)(<[^>]*>)?", full)
    return m.group(0) if m else full.split("(")[0]


def counter_per_dispatch(directory: str, counter: str) -> dict:
    """Average value of `counter` per dispatch, per kernel."""
    total, count = defaultdict(float), defaultdict(int)
    for r in _csv(directory, "counter_collection.csv"):
        if r["Counter_Name"] != counter:
            continue
        k = kernel_name(r["Kernel_Name"])
        total[k] += float(r["Counter_Value"])
        count[k] += 1
    return {k: (total[k] / count[k], count[k]) for k in total}


def calibration_factor(directory: str, counter: str, copy_bytes: int) -> float:
    """Bytes per counter KiB, from the largest dispatch of a plain copy of
    copy_bytes (its reads or its writes are copy_bytes each)."""
    best = 0.0
    for r in _csv(directory, "counter_collection.csv"):
        if r["Counter_Name"] == counter:
            best = max(best, float(r["Counter_Value"]))
    if best <= 0.0:
        raise ValueError(f"no {counter} reading in {directory}")
    return copy_bytes / best


def traffic(fetch_dir, write_dir, cal_fetch_dir, cal_write_dir, copy_bytes: int, batch: int) -> dict:
    ff = calibration_factor(cal_fetch_dir, "FETCH_SIZE", copy_bytes)
    wf = calibration_factor(cal_write_dir, "WRITE_SIZE", copy_bytes)
    fetch = counter_per_dispatch(fetch_dir, "FETCH_SIZE")
    write = counter_per_dispatch(write_dir, "WRITE_SIZE")
    kernels = {}
    for k in sorted(set(fetch) & set(write)):
        name = {"k_wa": "wa", "k_deposit": "deposit", "k_gradient": "gather", "k_optimizer": "optimizer"}.get(k, k)
        kernels[name] = {"kernel": k, "fetch_bytes": fetch[k][0] * ff, "write_bytes": write[k][0] * wf,
                         "fetch_kib": fetch[k][0], "write_kib": write[k][0], "dispatches": fetch[k][1]}
    return {"batch": batch, "copy_bytes": copy_bytes, "fetch_bytes_per_kib": ff, "write_bytes_per_kib": wf,
            "kernels": kernels}


def window_average_us(stats_dir: str, kernel: str, last: int):
    """Mean duration (us) of the last `last` dispatches of `kernel` in the
    kernel trace: the timed window of a run that ends with it."""
    rows = [r for r in _csv(stats_dir, "kernel_trace.csv") if kernel_name(r["Kernel_Name"]) == kernel]
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    rows = rows[-last:] if last > 0 else rows
    if not rows:
        return None, 0
    d = [int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in rows]
    return sum(d) / len(d) / 1e3, len(d)


def window_occupancy(stats_dir: str, window_ms: float) -> dict:
    """Over the last `window_ms` of the kernel trace (the timed window of a
    run that ends with it): the time at least one kernel ran, the time with
    k kernels in flight, and each kernel's summed duration there."""
    ev = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), kernel_name(r["Kernel_Name"]))
                for r in _csv(stats_dir, "kernel_trace.csv"))
    if not ev:
        return {}
    end = max(e[1] for e in ev)
    ev = [e for e in ev if e[0] >= end - window_ms * 1e6]
    points = sorted([(s, 1) for s, _, _ in ev] + [(e, -1) for _, e, _ in ev])
    inflight, last, by_k = 0, None, defaultdict(float)
    for t, d in points:
        if last is not None and inflight > 0:
            by_k[inflight] += t - last
        inflight += d
        last = t
    per_kernel = defaultdict(float)
    for s, e, k in ev:
        per_kernel[k] += e - s
    return {"window_ms": window_ms, "busy_ms": sum(by_k.values()) / 1e6,
            "in_flight_ms": {k: v / 1e6 for k, v in sorted(by_k.items())},
            "kernel_ms": {k: v / 1e6 for k, v in sorted(per_kernel.items(), key=lambda kv: -kv[1])}}


def stats_markdown(stats_dir: str, bench_line: dict, title: str, command: str, top: int = 24) -> str:
    rows = _csv(stats_dir, "kernel_stats.csv")
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    out = [f"# {title}", "", "Command (repository root, on the MI355X):", "", "```", command, "```", "",
           "| kernel | calls | total ms | average us | % of kernel time |", "|---|---:|---:|---:|---:|"]
    for r in rows[:top]:
        t = float(r["TotalDurationNs"])
        out.append(f"| `{kernel_name(r['Name'])}` | {r['Calls']} | {t / 1e6:.2f} | "
                   f"{float(r['AverageNs']) / 1e3:.1f} | {100 * t / total:.1f} |")
    rl = bench_line.get("roofline") or {}
    out += ["", "The benchmark line of the same run:", "", "```json",
            json.dumps({k: bench_line.get(k) for k in ("metric", "value", "unit", "ms_per_step", "config")}),
            json.dumps({"roofline": rl}), "```", ""]
    if rl.get("kernel"):
        kname = {"wa": "k_wa", "deposit": "k_deposit", "gather": "k_gradient"}.get(rl["kernel"], rl["kernel"])
        avg, n = window_average_us(stats_dir, kname, int(rl.get("launches", 0)))
        event_us = rl.get("avg_launch_ms", 0) * 1e3
        out.append(f"HIP-event average of `{kname}` in the timed region: {event_us:.1f} us per launch over "
                   f"{rl.get('launches')} launches.")
        if avg:
            frac = rl["algorithmic_per_launch"] / (avg * 1e-6) / 1e9 / rl["peak"]
            out.append(f"rocprofv3 average of the same last {n} `{kname}` dispatches (the timed window): "
                       f"{avg:.1f} us, {100 * (avg / event_us - 1):+.1f}% against the event timer; "
                       f"roofline fraction from it: {frac:.4f} (the line reports {rl['frac']:.4f}).")
    window = bench_line.get("steps", 0) * bench_line.get("ms_per_step", 0.0)
    occ = window_occupancy(stats_dir, window) if window > 0 else {}
    if occ:
        total = sum(occ["kernel_ms"].values())
        out += ["", f"Timed window (the last {window:.1f} ms of the trace): a kernel runs {occ['busy_ms']:.1f} ms "
                f"({100 * occ['busy_ms'] / window:.1f}%); time with k kernels in flight: "
                + ", ".join(f"{k}: {v:.1f} ms" for k, v in occ["in_flight_ms"].items())
                + f".  Kernel durations there sum to {total:.1f} ms "
                f"({total / max(occ['busy_ms'], 1e-9):.2f} kernels in flight on average); by kernel:", "",
                "| kernel | ms in the window | share |", "|---|---:|---:|"]
        for k, v in list(occ["kernel_ms"].items())[:12]:
            out.append(f"| `{k}` | {v:.2f} | {100 * v / total:.1f}% |")
    return "\n".join(out) + "\n"


def traffic_markdown(t: dict, fetch_cmd: str, write_cmd: str, cal_cmds: list[str]) -> str:
    out = ["# HBM traffic per launch (batched ecp5_85k workload)", "",
           f"Batch of {t['batch']} jobs per launch.  Counter runs:", "", "```", fetch_cmd, write_cmd, *cal_cmds, "```",
           "", f"Calibration: a {t['copy_bytes'] >> 20} MiB device copy reads FETCH_SIZE x "
           f"{t['fetch_bytes_per_kib']:.1f} bytes per KiB and writes WRITE_SIZE x {t['write_bytes_per_kib']:.1f} bytes "
           "per KiB; the readings below are scaled by these factors.", "",
           "| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM read MB | HBM write MB | dispatches |",
           "|---|---:|---:|---:|---:|---:|"]
    for name, k in sorted(t["kernels"].items(), key=lambda kv: -(kv[1]["fetch_bytes"] + kv[1]["write_bytes"])):
        out.append(f"| `{k['kernel']}` | {k['fetch_kib']:.0f} | {k['write_kib']:.0f} | {k['fetch_bytes'] / 1e6:.1f} | "
                   f"{k['write_bytes'] / 1e6:.1f} | {k['dispatches']} |")
    return "\n".join(out) + "\n"


def plain_stats(stats_dir: str, result_file: str, out_md: str, title: str, command: str):
    rows = _csv(stats_dir, "kernel_stats.csv")
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    out = [f"# {title}", "", "Command (repository root, on the MI355X):", "", "```", command, "```", "",
           "| kernel | calls | total ms | average us | % of kernel time |", "|---|---:|---:|---:|---:|"]
    for r in rows[:24]:
        t = float(r["TotalDurationNs"])
        out.append(f"| `{kernel_name(r['Name'])}` | {r['Calls']} | {t / 1e6:.2f} | "
                   f"{float(r['AverageNs']) / 1e3:.1f} | {100 * t / total:.1f} |")
    with open(result_file) as f:
        lines = [ln for ln in f.read().splitlines() if ln.startswith("{")]
    if lines:
        out += ["", "Result line of the same run:", "", "```json", lines[-1], "```"]
    with open(out_md, "w") as f:
        f.write("\n".join(out) + "\n")


def main(argv):
    if argv and argv[0] == "--stats":
        plain_stats(*argv[1:6])
        return
    (stats_dir, bench_json, fetch_dir, write_dir, cal_f, cal_w, copy_bytes, batch, out_dir) = argv[:9]
    copy_bytes, batch = int(copy_bytes), int(batch)
    with open(bench_json) as f:
        line = json.loads([ln for ln in f.read().splitlines() if ln.startswith("{")][-1])
    # the driver's benchmark window (20 timed iterations after 5 warm-up
    # iterations), its batched form only so that the run ends with the window
    bench_args = f"--gpus 1 --steps 20 --warmup 5 --only-batched --batch {batch}"
    stats_cmd = (f"rocprofv3 --kernel-trace --stats --output-format csv -d {stats_dir} -o run -- "
                 f"python bench.py {bench_args}")
    md = stats_markdown(stats_dir, line, f"Kernel statistics: ecp5_85k, {batch} multi-start jobs in lockstep, "
                        f"{line['config'].get('precision')} transform", stats_cmd)
    with open(os.path.join(out_dir, f"kernels_ecp5_85k_batch{batch}.md"), "w") as f:
        f.write(md)
    t = traffic(fetch_dir, write_dir, cal_f, cal_w, copy_bytes, batch)
    regex = "k_wa|k_deposit|k_sample|k_gradient|k_convert|k_optimizer|k_gemm"
    pmc = lambda c, d: (f"rocprofv3 --pmc {c} --kernel-include-regex \"{regex}\" --output-format csv -d {d} -o run -- "
                        f"python bench.py {bench_args}")
    cal = [f"rocprofv3 --pmc {c} --output-format csv -d {d} -o run -- "
           f"python bench.py --copy-calibration {copy_bytes >> 20}" for c, d in (("FETCH_SIZE", cal_f), ("WRITE_SIZE", cal_w))]
    with open(os.path.join(out_dir, f"traffic_ecp5_85k_batch{batch}.md"), "w") as f:
        f.write(traffic_markdown(t, pmc("FETCH_SIZE", fetch_dir), pmc("WRITE_SIZE", write_dir),
                                 cal))
    with open(os.path.join(out_dir, "traffic.json"), "w") as f:
        json.dump(t, f, indent=1)
    print(md)


if __name__ == "__main__":
    main(sys.argv[1:])
