"""Builds the gfx950 HIP library libtexo_gp.so in-tree with hipcc.

Usage: python -m texo_amd.build   (or texo_amd.build.build() from Python)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtexo_gp.so")
SOURCES = ["gp_kernels.hip", "gp_dct.hip", "gp_job.hip", "gp_engine.hip", "gp_api.hip", "gp_builders.hip", "gp_cluster.hip", "gp_seed.hip", "gp_auction.hip", "gp_repair.hip"]

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # packed FP32 instructions stay disabled; ordinary float code is used
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # no implicit FMA contraction: elementwise FP64 arithmetic rounds like the
    # reference; the contraction kernels call fmaf / MFMA explicitly
    "-ffp-contract=off",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
]


def hipcc() -> str:
    for cand in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def build(verbose: bool = False) -> str:
    objs = []
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    procs = []
    for src in SOURCES:
        obj = os.path.join(build_dir, src.replace(".hip", ".o"))
        cmd = [hipcc(), *HIPCC_FLAGS, "-c", "-x", "hip", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append(src)
            sys.stderr.write(text)
        elif verbose and text.strip():
            sys.stderr.write(text)
    if failed:
        raise RuntimeError(f"hipcc failed for {failed}")
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "--offload-arch=gfx950", *objs, "-o", tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    # static archive of the same objects, for hosts that link the library
    # into their own binary (a Rust -sys crate); each object carries its own
    # gfx950 code object, so no device link step is needed
    archive = os.path.join(HERE, "libtexo_gp.a")
    if os.path.exists(archive):
        os.remove(archive)
    subprocess.run(["ar", "rcs", archive, *objs], check=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
