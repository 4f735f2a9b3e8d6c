"""Writes tests/golden/trajectories.json: the FP64 oracle's fixed-iteration
trajectories of the three single-job configurations at their full lengths.

    python -m tests.golden.make_trajectories

The oracle restates the reference loop (optimize_density_once,
crates/texo-pnr/src/eplace.rs:439-582, with the stop rules off); a
1,000-iteration run of the 84k-cell workload takes minutes on one host, so
the GPU trajectory tests compare against these recorded end states instead
of re-running the oracle.  Each entry records a fingerprint of the generated
problem so a changed generator invalidates it loudly.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "trajectories.json")

# (configuration, iterations, gamma base): the lengths and gamma the
# configurations state (1,024 cells: 200 iterations at gamma = 4 bins;
# 24k cells: 500; 84k cells: 1,000)
RUNS = [("plain_1k", 200, 4.0), ("ecp5_25k", 500, 0.0), ("ecp5_85k", 1000, 0.0)]


def fingerprint(problem) -> str:
    h = hashlib.sha256()
    for name in sorted(vars(problem)):
        v = getattr(problem, name)
        if isinstance(v, np.ndarray):
            h.update(name.encode())
            h.update(np.ascontiguousarray(v).tobytes())
    h.update(f"{problem.width}x{problem.height}".encode())
    return h.hexdigest()[:32]


def main():
    import oracle
    from texo_amd import abi, synth
    lib = oracle.load()
    lib.oracle_set_threads(8)
    out = {}
    for name, iterations, gamma in RUNS:
        p = synth.CONFIGS[name]()
        t0 = time.perf_counter()
        s = abi.Session(lib, "oracle", p, abi.options(gamma_base=gamma))
        rep = s.iterate(iterations).as_dict()
        sched = s.schedule()
        s.close()
        rep.pop("elapsed_seconds", None)
        out[name] = {"iterations": iterations, "gamma_base": gamma, "fingerprint": fingerprint(p),
                     "report": rep, "schedule": sched, "oracle_seconds": round(time.perf_counter() - t0, 1)}
        print(name, out[name], flush=True)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
