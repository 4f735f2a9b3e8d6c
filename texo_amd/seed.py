"""Python mirror of the coarse quadratic seed (include/texo_gp.h, "coarse
quadratic seed"): the initial coordinates of analytical placement
(analytical_place, crates/texo-pnr/src/lib.rs:4042-4066) and the per-axis
Jacobi-PCG solve (solve_analytical_axis, lib.rs:4529-4556), computed on the
GPU through libtexo_gp.so."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi, load


def declare(lib, prefix: str = "texo_gp"):
    """Declares the seed entry points; `prefix` "oracle" declares the CPU
    restatement's, which take no device argument."""
    P, u8, u32, u64, d = C.POINTER, C.c_uint8, C.c_uint32, C.c_uint64, C.c_double
    dev = [C.c_int32] if prefix == "texo_gp" else []
    f = lib[f"{prefix}_solve_axis"]
    f.argtypes = dev + [u32, P(d), P(u32), P(u32), P(d), P(d), P(u8), P(d), u32, P(d), P(u64)]
    f.restype = C.c_int
    f = lib[f"{prefix}_quadratic_seed"]
    if prefix == "texo_gp":
        f.argtypes = [P(abi.texo_gp_problem), C.c_int32, P(d), P(abi.texo_gp_report)]
    else:
        f.argtypes = [P(abi.texo_gp_problem), P(d), P(abi.texo_gp_report)]
    f.restype = C.c_int


class _Named:
    def __init__(self, lib):
        self.lib = lib

    def __getitem__(self, name):
        return getattr(self.lib, name)


def _lib():
    lib = load()
    if not getattr(lib, "_seed_declared", False):
        declare(_Named(lib))
        lib._seed_declared = True
    return lib


def quadratic_seed(problem, device: int = 0) -> tuple[np.ndarray, int]:
    """(coordinates [n_units, 2], PCG iterations of both axes) of the coarse
    seed; raises PlacementError (DidNotConverge) where the reference's
    solve_quadratic fails."""
    lib = _lib()
    out = np.zeros(2 * problem.n_units, np.float64)
    rep = abi.texo_gp_report()
    s = problem.struct()
    rc = lib.texo_gp_quadratic_seed(C.byref(s), device, abi.ptr(out, C.c_double), C.byref(rep))
    if rc != 0:
        raise abi.PlacementError(rep)
    return out.reshape(-1, 2), int(rep.iterations)


def csr(adjacency) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(start, other, weight) of per-row lists of (other, weight) in insertion order."""
    start = np.zeros(len(adjacency) + 1, np.uint32)
    start[1:] = np.cumsum([len(a) for a in adjacency])
    other = np.array([o for a in adjacency for o, _ in a], np.uint32)
    weight = np.array([w for a in adjacency for _, w in a], np.float64)
    return start, other, weight


def solve_axis(diagonal, adjacency, rhs, anchored, initial, center: int, device: int = 0) -> tuple[np.ndarray, int]:
    """solve_analytical_axis of explicit equations; `adjacency` is per-row
    lists of (other, weight) or a (start, other, weight) CSR."""
    lib = _lib()
    diagonal = np.ascontiguousarray(diagonal, np.float64)
    n = diagonal.size
    start, other, weight = adjacency if isinstance(adjacency, tuple) else csr(adjacency)
    start = np.ascontiguousarray(start, np.uint32)
    other = np.ascontiguousarray(other if len(other) else np.zeros(1), np.uint32)
    weight = np.ascontiguousarray(weight if len(weight) else np.zeros(1), np.float64)
    rhs = np.ascontiguousarray(rhs, np.float64)
    anchored = np.ascontiguousarray(anchored, np.uint8)
    initial = np.ascontiguousarray(initial, np.float64)
    if not (start.size == n + 1 and rhs.size == anchored.size == initial.size == n):
        raise ValueError("solve_axis: diagonal, rhs, anchored and initial need one entry per row, start n+1")
    out = np.zeros(max(n, 1), np.float64)
    it = C.c_uint64()
    rc = lib.texo_gp_solve_axis(device, n, abi.ptr(diagonal, C.c_double), abi.ptr(start, C.c_uint32),
                                abi.ptr(other, C.c_uint32), abi.ptr(weight, C.c_double), abi.ptr(rhs, C.c_double),
                                abi.ptr(anchored, C.c_uint8), abi.ptr(initial, C.c_double), center,
                                abi.ptr(out, C.c_double), C.byref(it))
    if rc != 0:
        err = abi.PlacementError.from_status(rc)
        err.index = it.value
        raise err
    return out[:n], int(it.value)
