#!/usr/bin/env python3
"""Benchmark: GP iterations/s of the electrostatic global-placement hot loop.

One step = one GP iteration (one dynamic_nesterov_step of
optimize_density_once, crates/texo-pnr/src/eplace.rs:503-574) on the
84k-cell / 256x256-bin ECP5-85-scale workload, all inputs resident in HBM.
The stop rule is disabled so exactly K iterations are timed after W warm-up
iterations (the session's one-time set-up and first evaluations are inside the
warm-up).  Prints one JSON line.
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

METRIC = "GP iters/s at 84k cells, 256\u00b2 bins; jobs/s + p99 latency, MFMA vs FP32 plain"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0


def algorithmic_bytes(kernel: str, sess, problem) -> float:
    """Compulsory HBM bytes of one launch of `kernel` (each input read once,
    each output written once, at storage precision)."""
    M = problem.n_members
    Fl = sess.n_fillers
    F, G = sess.n_fields, problem.width * problem.height
    if kernel == "deposit":
        # coordinates (16 B), member offset/charge/field/unit (29 B) or filler
        # charge/field (9 B), 9 int64 read-modify-writes (144 B)
        return M * (16 + 29 + 144) + Fl * (16 + 9 + 144)
    if kernel == "wa":
        pins = problem.pin_cell.size + 2 * problem.arc_extra_weight.size
        nets = problem.n_nets + problem.arc_extra_weight.size
        # per pin: unit 4 + offset 16 + coordinates 16 + derivative 16; per net: start 4 + weight 8 + id 4
        return pins * 52 + nets * 16
    if kernel == "gather":
        DU = sess.n_density_units
        # per unit: unit id, member range, coordinates; per member 3x3 phi taps (36 B FP32) + offset/charge/field;
        # outputs coef*G per field (16 B) per unit
        return DU * (4 + 8 + 16 + 16 * F) + M * (36 + 25)
    raise ValueError(kernel)


def dct_flops(sess, problem) -> float:
    W, H, F = problem.width, problem.height, sess.n_fields
    return F * 4.0 * W * H * (W + H)


def cpu_baseline(problem, iterations: int):
    import oracle
    from texo_amd import abi
    lib = oracle.load()
    lib.oracle_set_threads(1)
    o = abi.Session(lib, "oracle", problem)
    o.iterate(1)  # set-up evaluations and the first iteration, untimed
    t0 = time.perf_counter()
    o.iterate(iterations)
    dt = time.perf_counter() - t0
    o.close()
    return iterations / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="ecp5_85k")
    ap.add_argument("--precision", default="bf16x3", choices=["fp32", "bf16x3"])
    ap.add_argument("--roofline-kernel", default="deposit", choices=["deposit", "wa", "gather", "dct"])
    ap.add_argument("--cpu-iterations", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--compare", action="store_true", help="also time the other DCT precision")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import texo_amd
    from texo_amd import abi, synth

    problem = synth.CONFIGS[args.config]()
    prec = {"fp32": abi.DCT_FP32, "bf16x3": abi.DCT_BF16X3}

    def run(precision: int, timing_kernel: str):
        sess = texo_amd.Session(problem, abi.options(dct_precision=precision))
        sess.iterate(args.warmup)
        sess.set_kernel_timing(timing_kernel)
        rep0 = sess.iterate(0)
        t0 = time.perf_counter()
        rep = sess.iterate(args.steps)  # returns after a stream synchronisation
        dt = time.perf_counter() - t0
        avg_ms, launches = sess.kernel_timing()
        evals = rep.evaluations - rep0.evaluations
        trials = rep.line_search_trials - rep0.line_search_trials
        out = {"seconds": dt, "it_per_s": args.steps / dt, "kernel_avg_ms": avg_ms, "kernel_launches": launches,
               "evaluations": evals, "trials": trials, "hpwl": rep.final_hpwl,
               "overflow": list(rep.field_overflow[:rep.n_fields])}
        return sess, out

    sess, main_run = run(prec[args.precision], args.roofline_kernel)
    other = None
    if args.compare:
        alt = "fp32" if args.precision == "bf16x3" else "bf16x3"
        s2, other = run(prec[alt], "dct")
        other["precision"] = alt
        s2.close()

    k = args.roofline_kernel
    avg_s = main_run["kernel_avg_ms"] * 1e-3
    if k == "dct":
        flops = dct_flops(sess, problem)
        achieved = flops / avg_s / 1e12 if avg_s > 0 else 0.0
        peak = BF16_PEAK_TFLOPS if args.precision == "bf16x3" else FP32_PEAK_TFLOPS
        roofline = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                    "frac": achieved / peak, "traffic": None, "kernel": "dct (4 stages)",
                    "algorithmic_per_launch": flops}
    else:
        nbytes = algorithmic_bytes(k, sess, problem)
        achieved = nbytes / avg_s / 1e9 if avg_s > 0 else 0.0
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": k,
                    "algorithmic_per_launch": nbytes, "avg_launch_ms": main_run["kernel_avg_ms"]}

    cpu = None
    if not args.no_cpu:
        v, dt = cpu_baseline(problem, args.cpu_iterations)
        cpu = {"value": v, "unit": "GP iterations/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_iterations} GP iterations of {args.config} (FP64 C++ restatement, "
                         f"{dt:.1f} s)"}

    line = {
        "metric": METRIC,
        "value": main_run["it_per_s"],
        "unit": "GP iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * main_run["seconds"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp64 coords/reductions, int64 fixed-point density, "
                 + ("bf16x3 MFMA" if args.precision == "bf16x3" else "fp32") + " DCT",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: {problem.meta['cells']} cells, {problem.width}x{problem.height} bins, "
                               f"{problem.meta['nets']} nets, {problem.meta['arcs']} timed arcs",
                   "precision": args.precision, "evaluations_per_iteration":
                       main_run["evaluations"] / args.steps, "trials_per_iteration": main_run["trials"] / args.steps},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if other:
        line["compare"] = {other["precision"]: other["it_per_s"]}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
