#!/usr/bin/env python3
"""Benchmark of the electrostatic global-placement hot loop on one MI355X.

Headline: GP iterations/s on the 84k-cell / 256x256-bin ECP5-85-scale
workload.  One step = one GP iteration (one dynamic_nesterov_step of
optimize_density_once, crates/texo-pnr/src/eplace.rs:503-574) with all
inputs resident in HBM.  The stop rules are off so exactly K iterations are
timed after W warm-up iterations (set-up and the first evaluations fall in
the warm-up); the library synchronises its streams before each timed call
returns, on both sides of the timed region.

Forms measured on the same workload:
  * plain FP32: one job, FP32 vector-ALU DCT;
  * MFMA single: one job, split-precision MFMA DCT;
  * MFMA batched (the reported value): B multi-start placements of the same
    design advanced in lockstep, one launch per kernel family covering all B;
    value = B * K / elapsed.
Serving (mixed sizes, 1k-84k cells, Poisson arrivals): full placements
through the continuous-batching server versus the same server admitting one
job at a time with the plain FP32 DCT; p50/p99 latency and jobs/s.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GP iters/s at 84k cells, 256² bins; jobs/s + p99 latency, MFMA vs FP32 plain"
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE)
# on the batched workload, scaled by the copy-kernel calibration; written by
# the profiling recipe next to the summaries it comes from
TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")


def measured_traffic(kernel: str, batch: int):
    """HBM bytes per launch of `kernel` from the committed counter runs, when
    they were taken on this batch size (else None)."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    if not k or t.get("batch") != batch:
        return None
    return k["fetch_bytes"] + k["write_bytes"]


FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0
FP16_PEAK_TFLOPS = 2500.0


def algorithmic_bytes(kernel: str, sizes: dict, problem):
    """Compulsory HBM bytes of a launch of `kernel`, as (per design, per
    evaluation): each input read once and each output written once, at
    storage precision.  The design-static arrays (netlist, pin and member
    offsets, unit tables, bounds) are read once per design in a launch, since
    the jobs of one design share one device copy of them; coordinates,
    charges, fields and results once per evaluation.  Gathered unit
    coordinates count once per coordinate, not once per use."""
    M = problem.n_members
    Fl = sizes["fillers"]
    F = sizes["fields"]
    G = problem.width * problem.height
    C = sizes["coordinates"]
    U = problem.n_units
    DU = sizes["density_units"]
    pins = problem.pin_cell.size + 2 * problem.arc_extra_weight.size
    nets = problem.n_nets + problem.arc_extra_weight.size
    pin_cells = np.concatenate([problem.pin_cell, problem.arc_cells]).astype(np.int64)
    movable_pins = int(np.count_nonzero(problem.unit_fixed[problem.cell_unit[pin_cells]] == 0))
    # offsets stored as FP32 pairs when every one is exact in FP32
    offs = np.concatenate([problem.member_offset_xy, problem.cell_offset_xy, problem.cell_fixed_xy])
    off = 8 if np.array_equal(offs.astype(np.float32).astype(np.float64), offs) else 16
    if kernel == "deposit":
        # coordinates 8 B each; per sorted charge: coordinate index 4, offset,
        # charge 8, layer 1; int64 accumulators of members and fillers per
        # field and bin: initial value read 8 + sum written 8
        return 0.0, C * 8 + (M + Fl) * (13 + off) + F * G * 2 * 16
    if kernel == "wa":
        # design: per pin unit 4 + offset / fixed coordinate + derivative slot
        # 4; per net start 4 + weight 8.  Evaluation: unit coordinates 16;
        # per movable pin its derivative written 16.  A value-only origin
        # evaluation writes no derivatives: the library's timer counts it as
        # U / (U + movable pins) of a full one, so the evaluations per launch
        # are full-evaluation equivalents
        return pins * (8 + off) + nets * 12, U * 16 + movable_pins * 16
    if kernel == "gather":
        # k_gradient.  Design: per unit pin range 4 + arc split 4 + density
        # unit 4; bounds 16 per coordinate; per density unit its member range
        # 4; per member field 1 + offset; per filler field 1.  Evaluation: per
        # coordinate position 8, metric 8, gradient written 8; per movable pin
        # its WA derivative 16; per member charge 8; the potential F*G FP32
        # once; per filler charge 8
        return (U * 12 + C * 16 + DU * 4 + M * (1 + off) + Fl,
                C * 24 + movable_pins * 16 + M * 8 + F * G * 4 + Fl * 8)
    if kernel == "optimizer":
        # per coordinate-long vector read or written (the timer counts them)
        return 0.0, C * 8
    raise ValueError(kernel)


def dct_flops(sizes: dict, problem) -> float:
    W, H, F = problem.width, problem.height, sizes["fields"]
    return F * 4.0 * W * H * (W + H)


def progress(msg: str):
    """Progress on stderr; stdout carries only the result line."""
    print(f"# {msg}", file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_all_cores(problem, iterations: int, threads: int):
    """The oracle with its parallel sections threaded as the reference's
    rayon sections are (density fields in parallel, the four WA sums
    {baseline, timing} x {x, y} in parallel): GP iterations/s over
    `iterations` after one untimed iteration."""
    import oracle
    from texo_amd import abi
    lib = oracle.load()
    lib.oracle_set_threads(threads)
    o = abi.Session(lib, "oracle", problem)
    o.iterate(1)
    t0 = time.perf_counter()
    o.iterate(iterations)
    dt = time.perf_counter() - t0
    o.close()
    lib.oracle_set_threads(1)
    return iterations / dt, dt


def cpu_baseline(problem, iterations: int, precision: int):
    """The FP64 oracle on one host core over `iterations` GP iterations of
    the workload (after one untimed iteration), and, as the checker, the
    error of the HIP path at `precision` against it: one evaluation at the
    start coordinates (potential, field at the charges, WA gradient, energy,
    overflow) and the final HPWL and overflow after the same iterations."""
    import oracle
    import texo_amd
    from texo_amd import abi
    lib = oracle.load()
    lib.oracle_set_threads(1)
    o = abi.Session(lib, "oracle", problem)
    g = texo_amd.Session(problem, abi.options(dct_precision=precision))

    def rel(a, b):
        a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    x = g.initial
    eg, eo = g.evaluate(x, 5.0, want_grids=True), o.evaluate(x, 5.0, want_grids=True)
    parity = {"precision": precision, "evaluation_at": "start coordinates, gamma 5",
              "potential_rel_l2": rel(eg["potential"], eo["potential"]),
              "field_rel_l2": rel(np.concatenate([eg["unit_gradient"].ravel(), eg["filler_gradient"].ravel()]),
                                  np.concatenate([eo["unit_gradient"].ravel(), eo["filler_gradient"].ravel()])),
              "wa_gradient_rel_l2": rel(eg["wa_gradient"], eo["wa_gradient"]),
              "energy_rel_max": float(np.max(np.abs(eg["energy"] - eo["energy"]) / np.abs(eo["energy"]))),
              "overflow_abs_max": float(np.max(np.abs(eg["overflow"] - eo["overflow"])))}
    o.iterate(1)  # set-up evaluations and the first iteration, untimed
    t0 = time.perf_counter()
    ro = o.iterate(iterations)
    dt = time.perf_counter() - t0
    rg = g.iterate(iterations + 1)
    parity["iterations"] = iterations + 1
    parity["hpwl_rel"] = abs(rg.final_hpwl - ro.final_hpwl) / ro.final_hpwl
    parity["overflow_abs_max_final"] = float(max(abs(a - b) for a, b in zip(rg.as_dict()["field_overflow"],
                                                                          ro.as_dict()["field_overflow"])))
    o.close()
    g.close()
    return iterations / dt, dt, parity


def run_single(problem, precision, steps, warmup):
    import texo_amd
    from texo_amd import abi
    sess = texo_amd.Session(problem, abi.options(dct_precision=precision))
    sess.iterate(warmup)
    rep0 = sess.iterate(0)
    t0 = time.perf_counter()
    rep = sess.iterate(steps)
    dt = time.perf_counter() - t0
    out = {"it_per_s": steps / dt, "ms_per_step": 1e3 * dt / steps,
           "evaluations_per_iteration": (rep.evaluations - rep0.evaluations) / steps,
           "trials_per_iteration": (rep.line_search_trials - rep0.line_search_trials) / steps,
           "hpwl": rep.final_hpwl, "overflow": list(rep.field_overflow[:rep.n_fields])}
    sess.close()
    return out


def baseline_table(batch: int) -> dict:
    """GP iterations/s of each form on the 1k / 64^2, 24k / 128^2 and 84k /
    256^2 configurations: the FP64 oracle on one core and on all cores
    (bounded samples of about 2-15 s each), the GPU with the FP32 transform
    (one job), with its configuration's MFMA transform (one job) and a
    lockstep batch of `batch` multi-start jobs with it."""
    from texo_amd import abi, synth
    nproc = os.cpu_count() or 1
    rows = {}
    for name, make, cpu_its, mfma in (("1: plain_1k", synth.plain_1k, 200, abi.DCT_BF16X3),
                                      ("2: ecp5_25k", synth.ecp5_25k, 40, abi.DCT_BF16X3),
                                      ("3: ecp5_85k", synth.ecp5_85k, 12, abi.DCT_FP16_SCALED)):
        p = make()
        one, t1, _ = cpu_baseline(p, cpu_its, abi.DCT_FP32)
        allc, ta = cpu_all_cores(p, cpu_its, max(nproc, 2))
        fp32 = run_single(p, abi.DCT_FP32, 200, 5)
        mf = run_single(p, mfma, 200, 5)
        starts = [synth.multistart(p, s) for s in range(batch)]
        bt, _ = run_batched(starts, mfma, 100, 5, "wa")
        rows[name] = {"cells": p.meta["cells"], "bins": f"{p.width}x{p.height}", "cpu_iterations": cpu_its,
                      "cpu_1_core_it_per_s": one, "cpu_all_cores_it_per_s": allc,
                      "cpu_seconds": {"1_core": t1, "all_cores": ta},
                      "gpu_fp32_single_it_per_s": fp32["it_per_s"],
                      "gpu_mfma_precision": {abi.DCT_BF16X3: "bf16x3", abi.DCT_FP16_SCALED: "fp16_scaled"}[mfma],
                      "gpu_mfma_single_it_per_s": mf["it_per_s"],
                      f"gpu_mfma_batch{batch}_it_per_s": bt["it_per_s"],
                      "batched_vs_fp32_single": bt["it_per_s"] / fp32["it_per_s"]}
        progress(f"table {name}: {rows[name]}")
    return {"baseline_table": rows, "nproc": nproc, "cpu_model": cpu_model(),
            "cpu_threads_note": "the oracle's parallel sections are the reference's rayon sections "
                                "(density fields, four WA sums): at most 4 threads work at once"}


def run_batched(problems, precision, steps, warmup, timing_kernel, extra_kernels=(), extra_steps=100):
    """The lockstep batch: `steps` timed iterations with `timing_kernel`'s
    launches timed, then `extra_steps` more iterations for each kernel of
    `extra_kernels` with that kernel timed (its own window)."""
    import texo_amd
    from texo_amd import abi
    b = texo_amd.Batch()
    for p in problems:
        b.add(p, abi.options(dct_precision=precision))
    b.iterate(warmup)
    r0 = [b.result(i, targets=False)[1] for i in range(len(problems))]
    b.set_kernel_timing(timing_kernel)
    t0 = time.perf_counter()
    b.iterate(steps)
    dt = time.perf_counter() - t0
    avg_ms, launches, tasks_per_launch = b.kernel_timing()
    r1 = [b.result(i, targets=False)[1] for i in range(len(problems))]
    sizes = b.sizes(0)
    extra = {}
    for k in extra_kernels:
        b.set_kernel_timing(k)
        b.iterate(extra_steps)
        extra[k] = b.kernel_timing()
    b.set_kernel_timing(None)
    n = len(problems)
    out = {"it_per_s": n * steps / dt, "per_job_it_per_s": steps / dt, "ms_per_step": 1e3 * dt / steps,
           "kernel_avg_ms": avg_ms, "kernel_launches": launches, "kernel_tasks_per_launch": tasks_per_launch,
           "evaluations_per_iteration": sum(b1["evaluations"] - a["evaluations"] for a, b1 in zip(r0, r1)) / (n * steps),
           "trials_per_iteration": sum(b1["line_search_trials"] - a["line_search_trials"]
                                       for a, b1 in zip(r0, r1)) / (n * steps),
           "hpwl": [r["final_hpwl"] for r in r1], "extra_timing": extra}
    b.close()
    return out, sizes


def run_serving(stream, precision, max_batch, at_rate: bool, clients: int, warm: bool = True):
    """Submits the stream to a server from `clients` threads (job i from
    thread i % clients, at its arrival time or as soon as the thread can) and
    waits for every job.  Each submission loads and validates its problem on
    the submitting thread, as a serving front end would.  Latency =
    submission to completion.  With `warm`, the same stream first runs once
    untimed through the same server, as a long-running service would have
    run jobs before: its device and pinned-host block pools and transform
    bases are then in place (every job of the timed pass still runs from its
    problem)."""
    import threading
    import texo_amd
    from texo_amd import abi

    def one_pass(srv):
        ids = [None] * len(stream)
        t_start = time.perf_counter()

        def client(k):
            for i in range(k, len(stream), clients):
                arrival, p = stream[i]
                if at_rate:
                    while True:
                        ahead = arrival - (time.perf_counter() - t_start)
                        if ahead <= 0:
                            break
                        time.sleep(min(ahead, 0.002))
                ids[i] = srv.submit(p, abi.options(dct_precision=precision))

        threads = [threading.Thread(target=client, args=(k,)) for k in range(clients)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        submitted = time.perf_counter() - t_start
        lat, failed = [], 0
        for i in ids:
            _, rep = srv.wait(i, raise_on_error=False)
            lat.append(rep["elapsed_seconds"])
            failed += rep["status"] != "OK"
        return time.perf_counter() - t_start, submitted, np.array(lat), failed

    with texo_amd.Server(max_batch=max_batch) as srv:
        if warm:
            one_pass(srv)
        s0 = srv.stats()
        makespan, submitted, lat, failed = one_pass(srv)
        s1 = srv.stats()
    d = {k: s1[k] - s0[k] for k in ("rounds", "busy_seconds", "admit_seconds", "retire_seconds", "wait_seconds",
                                    "launch_seconds")}
    return {"jobs": len(stream), "clients": clients, "warm": warm, "jobs_per_s": len(stream) / makespan,
            "makespan_s": makespan, "all_submitted_s": submitted,
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "failed": int(failed), "rounds": d["rounds"], "busy_s": d["busy_seconds"],
            "admit_s": d["admit_seconds"], "retire_s": d["retire_seconds"], "gpu_wait_s": d["wait_seconds"],
            "launch_s": d["launch_seconds"]}


def run_eco(n_requests: int, n_full: int, clients: int):
    """The incremental stream on its resident design (250k cells, 512x512
    bins, 300k nets): `n_requests` requests, each re-placing a connected 1-5%
    of the units for 100 iterations with the rest fixed, interleaved with
    `n_full` full placements of the design from multi-start coordinates
    (reference stop rules), all submitted at once from `clients` threads;
    BF16x3 batched against one at a time with FP32.  Latency = submission to
    completion; requests are generated before the clock starts."""
    import threading
    import texo_amd
    from texo_amd import abi, synth
    design = synth.eco_design()
    eco = synth.EcoStream(design)
    jobs = [(eco.request(1000 + k), 100) for k in range(n_requests)]
    step = max(1, n_requests // max(1, n_full))
    for k in range(n_full):
        jobs.insert(min(len(jobs), (k + 1) * step + k), (synth.multistart(design, k), 0))
    out = {"design": {k: design.meta[k] for k in ("cells", "nets", "pins")} | {"bins": "512x512"},
           "requests": n_requests, "full_jobs": n_full, "clients": clients,
           "mean_movable_units": float(np.mean([p.meta.get("movable_units", 0) for p, it in jobs if it]))}
    for name, precision, max_batch in (("batched_bf16x3", abi.DCT_BF16X3, 256), ("one_at_a_time_fp32", abi.DCT_FP32, 1)):
        with texo_amd.Server(max_batch=max_batch) as srv:
            ids = [None] * len(jobs)
            t0 = time.perf_counter()

            def client(k):
                for i in range(k, len(jobs), clients):
                    p, it = jobs[i]
                    ids[i] = srv.submit(p, abi.options(dct_precision=precision, fixed_iterations=it))

            threads = [threading.Thread(target=client, args=(k,)) for k in range(clients)]
            for t in threads:
                t.start()
            for t in threads:
                t.join()
            lat = {100: [], 0: []}
            failed = 0
            for i, (p, it) in zip(ids, jobs):
                _, rep = srv.wait(i, raise_on_error=False)
                lat[it].append(rep["elapsed_seconds"])
                failed += rep["status"] != "OK"
            dt = time.perf_counter() - t0
        r = {"jobs_per_s": len(jobs) / dt, "makespan_s": dt, "failed": failed,
             "request_p50_s": float(np.percentile(lat[100], 50)), "request_p99_s": float(np.percentile(lat[100], 99))}
        if lat[0]:
            r["full_p50_s"] = float(np.percentile(lat[0], 50))
        out[name] = r
        progress(f"eco {name}: {r}")
    out["jobs_per_s_ratio"] = out["batched_bf16x3"]["jobs_per_s"] / out["one_at_a_time_fp32"]["jobs_per_s"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--batch", type=int, default=16, help="multi-start jobs in the batched form")
    ap.add_argument("--precision", default="fp16_scaled", choices=["fp32", "bf16x3", "fp16_scaled"])
    ap.add_argument("--roofline-kernel", default="wa", choices=["deposit", "wa", "gather", "dct"])
    ap.add_argument("--serving-jobs", type=int, default=256)
    ap.add_argument("--serving-rate", type=float, default=20.0)
    ap.add_argument("--serving-clients", type=int, default=8, help="threads submitting the serving stream")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--eco-requests", type=int, default=64,
                    help="incremental-stream requests (a bounded sample of the 1,024-request stream; 0: skip)")
    ap.add_argument("--eco-full", type=int, default=1, help="full multi-start jobs interleaved with the requests")
    ap.add_argument("--config2-steps", type=int, default=200,
                    help="GP iterations of the 24k-cell / 128^2 single-job forms (FP32 vs BF16x3; 0: skip)")
    ap.add_argument("--cpu-iterations", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only-batched", action="store_true", help="time the batched form only, with no single-job, config-2, transform, serving, ECO or CPU legs (for profiling)")
    ap.add_argument("--only-serving", action="store_true",
                    help="run the saturated batched serving stream only and print its result (for profiling)")
    ap.add_argument("--only-eco", action="store_true",
                    help="run the incremental stream sample only (--eco-requests, --eco-full) and print its result")
    ap.add_argument("--baseline-table", action="store_true",
                    help="only the per-configuration table: the oracle on 1 core and on all cores, and the GPU "
                         "forms, on the 1k, 24k and 84k configurations (bounded samples)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 5 times (reference reading for the HBM counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n, "copies": 5}))
        return
    if args.baseline_table:
        print(json.dumps(baseline_table(args.batch)), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: the placement loop of one job runs on one "
                                   "device, and runs on several GPUs are not part of it", "n_gpus": args.gpus}))
        sys.exit(2)

    import texo_amd  # noqa: F401  (fails loudly when the HIP library is missing)
    from texo_amd import abi, synth

    prec = {"fp32": abi.DCT_FP32, "bf16x3": abi.DCT_BF16X3, "fp16_scaled": abi.DCT_FP16_SCALED}
    mfma = prec[args.precision]
    if args.only_serving:
        stream = synth.serving_stream(args.serving_jobs, rate=args.serving_rate)
        r = run_serving(stream, mfma, 256, False, args.serving_clients)
        print(json.dumps({"serving_batched_saturated": r, "precision": args.precision}), flush=True)
        return
    if args.only_eco:
        print(json.dumps({"eco": run_eco(max(1, args.eco_requests), args.eco_full, args.serving_clients)}), flush=True)
        return
    base = synth.ecp5_85k()

    if args.only_batched:
        plain = single = {"it_per_s": float("nan")}
    else:
        plain = run_single(base, abi.DCT_FP32, args.steps, args.warmup)
        progress(f"plain fp32 single job: {plain['it_per_s']:.1f} it/s")
        single = run_single(base, mfma, args.steps, args.warmup)
        progress(f"{args.precision} single job: {single['it_per_s']:.1f} it/s")
    starts = [synth.multistart(base, s) for s in range(args.batch)]
    extra_kernels = () if args.only_batched else tuple(
        k for k in ("gather", "deposit", "optimizer") if k != args.roofline_kernel)
    batched, sizes = run_batched(starts, mfma, args.steps, args.warmup, args.roofline_kernel, extra_kernels)
    progress(f"{args.precision} batch of {args.batch}: {batched['it_per_s']:.1f} it/s")

    k = args.roofline_kernel
    avg_s = batched["kernel_avg_ms"] * 1e-3
    # evaluations one timed launch covered (the jobs of one of the batch's
    # two job groups, one or two evaluations each)
    per_launch = batched["kernel_tasks_per_launch"]
    if k == "dct":
        flops = per_launch * dct_flops(sizes, base)
        achieved = flops / avg_s / 1e12 if avg_s > 0 else 0.0
        peak = {"fp32": FP32_PEAK_TFLOPS, "bf16x3": BF16_PEAK_TFLOPS, "fp16_scaled": FP16_PEAK_TFLOPS}[args.precision]
        roofline = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                    "traffic": None, "kernel": "dct (4 stages)", "algorithmic_per_launch": flops,
                    "avg_launch_ms": batched["kernel_avg_ms"], "evaluations_per_launch": per_launch}
    else:
        # the batch's jobs are starts of one design: its design-static arrays
        # are read once per launch
        per_design, per_eval = algorithmic_bytes(k, sizes, base)
        nbytes = per_design + per_launch * per_eval
        achieved = nbytes / avg_s / 1e9 if avg_s > 0 else 0.0
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(k, args.batch), "kernel": k,
                    "algorithmic_per_launch": nbytes, "design_bytes": per_design, "evaluation_bytes": per_eval,
                    "avg_launch_ms": batched["kernel_avg_ms"], "evaluations_per_launch": per_launch}

    roofline["launches"] = batched["kernel_launches"]
    # the other HBM-bound kernels, each timed over its own 100 iterations of
    # the same batch
    kernels = {}
    for kname, (k_ms, k_count, k_tasks) in batched["extra_timing"].items():
        per_design, per_unit = algorithmic_bytes(kname, sizes, base)
        nbytes = per_design + k_tasks * per_unit
        k_s = k_ms * 1e-3
        gbs = nbytes / k_s / 1e9 if k_s > 0 else 0.0
        kernels[kname] = {"achieved": gbs, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                          "traffic": measured_traffic(kname, args.batch),
                          "algorithmic_per_launch": nbytes, "avg_launch_ms": k_ms, "launches": k_count,
                          ("vectors_per_launch" if kname == "optimizer" else "evaluations_per_launch"): k_tasks}
        progress(f"kernel {kname}: {kernels[kname]}")

    # the 24k-cell / 128^2 configuration: FP32 plain vs BF16x3 single jobs
    config2 = None
    if args.config2_steps > 0 and not args.only_batched:
        p2 = synth.ecp5_25k()
        config2 = {"workload": f"ecp5_25k: {p2.meta['cells']} cells, {p2.width}x{p2.height} bins, "
                               f"{p2.meta['nets']} nets, one job", "steps": args.config2_steps}
        for name, prec in (("fp32", abi.DCT_FP32), ("bf16x3", abi.DCT_BF16X3)):
            config2[f"{name}_it_per_s"] = run_single(p2, prec, args.config2_steps, args.warmup)["it_per_s"]
        config2["bf16x3_vs_fp32"] = config2["bf16x3_it_per_s"] / config2["fp32_it_per_s"]
        progress(f"config 2: {config2}")

    # the Poisson transform's MFMA use, from a second, shorter timed run of
    # the same batch with the transform launches timed (algorithmic flops:
    # F*4*G*(W+H) per evaluation; the split modes issue three products each)
    transform = None
    if k != "dct" and not args.only_batched:
        tb, _ = run_batched(starts, mfma, min(args.steps, 200), args.warmup, "dct")
        t_s = tb["kernel_avg_ms"] * 1e-3
        flops = tb["kernel_tasks_per_launch"] * dct_flops(sizes, base)
        peak = {"fp32": FP32_PEAK_TFLOPS, "bf16x3": BF16_PEAK_TFLOPS, "fp16_scaled": FP16_PEAK_TFLOPS}[args.precision]
        issued = 1 if args.precision == "fp32" else 3
        transform = {"bound": "mfma" if issued == 3 else "valu", "kernel": "dct (4 stages)",
                     "algorithmic_per_launch": flops, "avg_launch_ms": tb["kernel_avg_ms"],
                     "evaluations_per_launch": tb["kernel_tasks_per_launch"],
                     "achieved": flops / t_s / 1e12 if t_s > 0 else 0.0, "unit": "TFLOP/s",
                     "issued": issued * flops / t_s / 1e12 if t_s > 0 else 0.0, "peak": peak,
                     "issued_frac": issued * flops / t_s / 1e12 / peak if t_s > 0 else 0.0}
        progress(f"transform: {transform}")

    serving = None
    if not args.no_serving and not args.only_batched:
        stream = synth.serving_stream(args.serving_jobs, rate=args.serving_rate)
        serving = {"jobs": args.serving_jobs, "arrival_rate_jobs_per_s": args.serving_rate,
                   "cells": "log-uniform 1k-84k, grid side nextpow2(sqrt(cells/4)) in 64-256, "
                            "weight exponents {2,4,6}", "clients": args.serving_clients}
        for name, precision, max_batch, at_rate in (("batched_mfma_at_rate", mfma, 256, True),
                                                    ("batched_mfma_saturated", mfma, 256, False),
                                                    ("one_at_a_time_fp32_at_rate", abi.DCT_FP32, 1, True),
                                                    ("one_at_a_time_fp32_saturated", abi.DCT_FP32, 1, False)):
            serving[name] = run_serving(stream, precision, max_batch, at_rate, args.serving_clients)
            progress(f"serving {name}: {serving[name]}")
        serving["saturated_jobs_per_s_ratio"] = (serving["batched_mfma_saturated"]["jobs_per_s"]
                                                 / serving["one_at_a_time_fp32_saturated"]["jobs_per_s"])

    eco = None
    if args.eco_requests > 0 and not args.only_batched:
        eco = run_eco(args.eco_requests, args.eco_full, args.serving_clients)

    cpu = None
    if not args.no_cpu and not args.only_batched:
        progress("cpu baseline")
        v1, dt1, parity = cpu_baseline(base, args.cpu_iterations, mfma)
        nproc = os.cpu_count() or 1
        # the reference's parallel sections on this workload: 2 density fields
        # and 4 WA sums, so at most 4 threads work at once
        threads = min(4, nproc)
        va, dta = cpu_all_cores(base, args.cpu_iterations, max(nproc, 2))
        cpu = {"value": va, "unit": "GP iterations/s", "cores": threads, "kind": "port",
               "sample": f"{args.cpu_iterations} GP iterations of the 84k-cell / 256x256 workload, one job "
                         f"(FP64 C++ restatement of the reference loop, its density fields and four WA sums "
                         f"in parallel as the reference's rayon sections are, {dta:.1f} s)",
               "single_core": {"value": v1, "cores": 1, "seconds": dt1},
               "nproc": nproc, "cpu_model": cpu_model(),
               "error_vs_oracle": parity}

    line = {
        "metric": METRIC,
        "value": batched["it_per_s"],
        "unit": "GP iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": batched["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"ecp5_85k: {base.meta['cells']} cells, {base.width}x{base.height} bins, "
                               f"{base.meta['nets']} nets, {base.meta['arcs']} timed arcs; "
                               f"{args.batch} multi-start placements in lockstep",
                   "arithmetic": "f64 coordinates, wirelength and reductions; int64 fixed-point density; "
                                 f"{args.precision} Poisson transform",
                   "batch": args.batch, "precision": args.precision,
                   "evaluations_per_iteration": batched["evaluations_per_iteration"],
                   "trials_per_iteration": batched["trials_per_iteration"]},
        "forms": {"fp32_plain_single_it_per_s": plain["it_per_s"],
                  f"{args.precision}_single_it_per_s": single["it_per_s"],
                  f"{args.precision}_batched_it_per_s": batched["it_per_s"],
                  "batched_per_job_it_per_s": batched["per_job_it_per_s"],
                  "batched_vs_fp32_plain": batched["it_per_s"] / plain["it_per_s"]},
        "roofline": roofline,
        "roofline_kernels": kernels,
        "roofline_transform": transform,
        "config2": config2,
        "cpu_baseline": cpu,
        "serving": serving,
        "eco": eco,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
