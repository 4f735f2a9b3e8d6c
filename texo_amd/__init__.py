"""MI355X-native electrostatic global placement (texo-pnr eplace hot path).

`place(problem, ...)` mirrors texo-pnr's `eplace::place`
(crates/texo-pnr/src/eplace.rs:242-253): it returns one integer tile target
per placement unit or raises `PlacementError` carrying the EplaceError
variant.  All computation runs in libtexo_gp.so (hand-written gfx950 HIP
kernels behind the C ABI of include/texo_gp.h); there is no CPU fallback, and
loading fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import abi
from .abi import PlacementError, Problem, options  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtexo_gp.so")

_lib = None


def load():
    """Loads libtexo_gp.so (built by texo_amd.build).  Raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m texo_amd.build` "
                           "(the placement path has no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    abi.declare_session_api(lib, "texo_gp")
    abi.declare_server_api(lib, "texo_gp")
    abi.declare_batch_api(lib, "texo_gp")
    P, d = C.POINTER, C.c_double
    lib.texo_gp_last_error.restype = C.c_char_p
    lib.texo_gp_last_error.argtypes = []
    lib.texo_gp_session_poisson.argtypes = [C.c_void_p, P(d), P(d), P(d)]
    lib.texo_gp_session_poisson.restype = C.c_int
    lib.texo_gp_session_set_kernel_timing.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
    lib.texo_gp_session_set_kernel_timing.restype = C.c_int
    lib.texo_gp_session_kernel_timing.argtypes = [C.c_void_p, P(d), P(C.c_uint64), P(d)]
    lib.texo_gp_session_kernel_timing.restype = C.c_int
    lib.texo_gp_session_round_targets.argtypes = [C.c_void_p, P(d), P(C.c_uint32)]
    lib.texo_gp_session_round_targets.restype = C.c_int
    _lib = lib
    return lib


def last_error() -> str:
    return load().texo_gp_last_error().decode(errors="replace")


class Session(abi.Session):
    """Device-resident placement session (one HIP stream, one HBM arena)."""

    def __init__(self, problem: Problem, opts: Optional[abi.texo_gp_options] = None):
        super().__init__(load(), "texo_gp", problem, opts)

    def poisson(self, density: np.ndarray):
        F, G = self.n_fields, self.problem.width * self.problem.height
        rho = np.ascontiguousarray(density, dtype=np.float64).reshape(-1)
        assert rho.size == F * G
        phi = np.zeros(F * G)
        energy = np.zeros(F)
        rc = self.lib.texo_gp_session_poisson(self.handle, abi.ptr(rho, C.c_double), abi.ptr(phi, C.c_double),
                                              abi.ptr(energy, C.c_double))
        if rc != 0:
            raise PlacementError.from_status(rc)
        return phi.reshape(F, self.problem.height, self.problem.width), energy

    def round_targets(self, coordinates: np.ndarray) -> np.ndarray:
        """Device-rounded unit targets [n_units, 2] of interleaved coordinates."""
        x = np.ascontiguousarray(coordinates, dtype=np.float64)
        assert x.size >= 2 * self.problem.n_units
        t = np.zeros(2 * self.problem.n_units, np.uint32)
        rc = self.lib.texo_gp_session_round_targets(self.handle, abi.ptr(x, C.c_double), abi.ptr(t, C.c_uint32))
        if rc != 0:
            raise PlacementError.from_status(rc)
        return t.reshape(-1, 2)

    def set_kernel_timing(self, kernel: Optional[str]):
        self.lib.texo_gp_session_set_kernel_timing(self.handle, (kernel or "").encode(), 1 if kernel else 0)

    def kernel_timing(self):
        """(average ms per launch, launches, average evaluations per launch)."""
        avg, count, tasks = C.c_double(), C.c_uint64(), C.c_double()
        rc = self.lib.texo_gp_session_kernel_timing(self.handle, C.byref(avg), C.byref(count), C.byref(tasks))
        if rc != 0:
            raise PlacementError.from_status(rc)
        return avg.value, count.value, tasks.value


def place(problem: Problem, dct_precision: int = abi.DCT_FP32, fixed_iterations: int = 0, trace: bool = False,
          gamma_base: float = 0.0, device: int = 0):
    """eplace::place: returns (targets[n_units, 2] uint32, report dict)."""
    lib = load()
    targets = np.zeros(2 * problem.n_units, np.uint32)
    rep = abi.texo_gp_report()
    opts = options(dct_precision=dct_precision, fixed_iterations=fixed_iterations, trace=trace,
                   gamma_base=gamma_base, device=device)
    rc = lib.texo_gp_place(C.byref(problem.struct()), C.byref(opts), abi.ptr(targets, C.c_uint32), C.byref(rep))
    if rc != 0:
        raise PlacementError(rep)
    return targets.reshape(-1, 2), rep.as_dict()


class Server:
    """Continuous-batching placement server on one GPU.

    Jobs submitted here run in lockstep rounds with every other admitted job:
    each round issues one launch per kernel family over the whole batch, and a
    job's result is bit-identical to the same job placed alone.  `submit`
    deep-copies and validates the problem on the calling thread; `wait` and
    `poll` return `(targets[n_units, 2], report)` or raise `PlacementError`.
    """

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0, max_batch: int = 0):
        self.lib = load()
        self.handle = self.lib.texo_gp_server_create(device, hbm_budget_bytes, max_batch)
        if not self.handle:
            raise RuntimeError(f"texo_gp_server_create failed: {last_error()}")
        self._units = {}

    def submit(self, problem: Problem, opts: Optional[abi.texo_gp_options] = None) -> int:
        opts = opts if opts is not None else options()
        job = self.lib.texo_gp_server_submit(self.handle, C.byref(problem.struct()), C.byref(opts))
        if job <= 0:
            raise PlacementError.from_status(-job)
        self._units[job] = problem.n_units
        return job

    def _result(self, fn, job: int, raise_on_error: bool):
        targets = np.zeros(2 * self._units[job], np.uint32)
        rep = abi.texo_gp_report()
        rc = fn(self.handle, job, abi.ptr(targets, C.c_uint32), C.byref(rep))
        if rc == abi.STATUS_CODE["NotReady"]:
            return None
        del self._units[job]
        if rc != 0:
            if raise_on_error:
                raise PlacementError(rep)
            return None, rep.as_dict()
        return targets.reshape(-1, 2), rep.as_dict()

    def poll(self, job: int, raise_on_error: bool = True):
        """None while the job runs, else its result (released by this call).
        With raise_on_error=False a failed job gives (None, report)."""
        return self._result(self.lib.texo_gp_server_poll, job, raise_on_error)

    def wait(self, job: int, raise_on_error: bool = True):
        return self._result(self.lib.texo_gp_server_wait, job, raise_on_error)

    def cancel(self, job: int) -> int:
        """Cancels a pending or running job (it then completes with status
        Cancelled); returns the library's status code."""
        return self.lib.texo_gp_server_cancel(self.handle, job)

    def stats(self) -> dict:
        st = abi.texo_gp_server_stats()
        rc = self.lib.texo_gp_server_stats_get(self.handle, C.byref(st))
        if rc != 0:
            raise PlacementError.from_status(rc)
        return st.as_dict()

    def close(self):
        if getattr(self, "handle", None):
            self.lib.texo_gp_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Lockstep batch on one GPU: every job advances in the same rounds, one
    launch per kernel family covering all of them (multi-start seeds,
    timing-weight variants, unrelated designs).  Each job's trajectory is
    bit-identical to the same job run alone."""

    def __init__(self, device: int = 0):
        self.lib = load()
        self.device = device
        self.handle = self.lib.texo_gp_batch_create(device)
        if not self.handle:
            raise RuntimeError(f"texo_gp_batch_create failed: {last_error()}")
        self.problems = []

    def add(self, problem: Problem, opts: Optional[abi.texo_gp_options] = None) -> int:
        opts = opts if opts is not None else options(device=self.device)
        rep = abi.texo_gp_report()
        idx = self.lib.texo_gp_batch_add(self.handle, C.byref(problem.struct()), C.byref(opts), C.byref(rep))
        if idx < 0:
            raise PlacementError(rep)
        self.problems.append(problem)
        return idx

    def run(self):
        rc = self.lib.texo_gp_batch_run(self.handle)
        if rc != 0:
            raise PlacementError.from_status(rc)

    def iterate(self, iterations: int):
        rc = self.lib.texo_gp_batch_iterate(self.handle, iterations)
        if rc != 0:
            raise PlacementError.from_status(rc)

    def result(self, job: int, targets: bool = True):
        """(targets[n_units, 2] or None, report dict); raises PlacementError for a failed job."""
        t = np.zeros(2 * self.problems[job].n_units, np.uint32) if targets else None
        rep = abi.texo_gp_report()
        rc = self.lib.texo_gp_batch_result(self.handle, job, abi.ptr(t, C.c_uint32), C.byref(rep))
        if rc != 0:
            raise PlacementError(rep)
        return (t.reshape(-1, 2) if t is not None else None), rep.as_dict()

    def sizes(self, job: int) -> dict:
        n, f, fl, du = C.c_uint64(), C.c_uint32(), C.c_uint64(), C.c_uint64()
        rc = self.lib.texo_gp_batch_sizes(self.handle, job, C.byref(n), C.byref(f), C.byref(fl), C.byref(du))
        if rc != 0:
            raise PlacementError.from_status(rc)
        return {"coordinates": n.value, "fields": f.value, "fillers": fl.value, "density_units": du.value}

    def coordinates(self, job: int) -> np.ndarray:
        x = np.zeros(self.sizes(job)["coordinates"])
        rc = self.lib.texo_gp_batch_coordinates(self.handle, job, abi.ptr(x, C.c_double))
        if rc != 0:
            raise PlacementError.from_status(rc)
        return x

    def set_kernel_timing(self, kernel: Optional[str]):
        self.lib.texo_gp_batch_set_kernel_timing(self.handle, (kernel or "").encode(), 1 if kernel else 0)

    def kernel_timing(self):
        """(average ms per launch, launches, average evaluations per launch)."""
        avg, count, tasks = C.c_double(), C.c_uint64(), C.c_double()
        rc = self.lib.texo_gp_batch_kernel_timing(self.handle, C.byref(avg), C.byref(count), C.byref(tasks))
        if rc != 0:
            raise PlacementError.from_status(rc)
        return avg.value, count.value, tasks.value

    def close(self):
        if getattr(self, "handle", None):
            self.lib.texo_gp_batch_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
