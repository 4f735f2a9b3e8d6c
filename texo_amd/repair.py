"""Python mirror of the legalization repair (include/texo_gp.h,
"legalization repair"): the transactional repair of a relaxed
unit-to-object projection (repair_relaxed,
crates/texo-pnr/src/legalization/repair.rs:60-155) over explicit candidate
footprints, through libtexo_gp.so (on the host: the search is serial).

A problem is described unit by unit: each unit has its cells (count and
first cell id), its GP target, whether it is fixed, its preferred candidate
(the relaxed projection's, or None) and its candidates; a candidate lists
the objects it occupies (one per cell; the first one's point orders the
candidates), the (shared resource, value) pairs it requires and the
(pin wire, net) pairs it puts on wires."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import abi, load


class texo_gp_repair_problem(C.Structure):
    _fields_ = [
        ("width", C.c_uint32), ("height", C.c_uint32),
        ("n_objects", C.c_uint64), ("object_xy", C.POINTER(C.c_uint32)),
        ("n_wires", C.c_uint64), ("wire_capacity", C.POINTER(C.c_uint32)),
        ("n_units", C.c_uint64), ("unit_cells", C.POINTER(C.c_uint32)),
        ("unit_first_cell", C.POINTER(C.c_uint64)), ("unit_target_xy", C.POINTER(C.c_uint32)),
        ("unit_fixed", C.POINTER(C.c_uint8)), ("preferred", C.POINTER(C.c_int64)),
        ("cand_start", C.POINTER(C.c_uint64)),
        ("cand_object_start", C.POINTER(C.c_uint64)), ("cand_object", C.POINTER(C.c_uint32)),
        ("cand_shared_start", C.POINTER(C.c_uint64)), ("cand_shared", C.POINTER(C.c_uint64)),
        ("cand_wire_start", C.POINTER(C.c_uint64)), ("cand_wire", C.POINTER(C.c_uint64)),
    ]


STATS = ("initial_pending", "direct_fast_choices", "direct_fast_moves", "direct_fast_fallbacks",
         "augmenting_roots", "augmenting_fallbacks", "states_visited", "failed_states", "max_unplaced",
         "max_blocker_branches", "units_evicted", "choices_examined")


@dataclass
class Candidate:
    objects: Sequence[int]
    shared: Sequence[tuple] = ()
    wires: Sequence[tuple] = ()


@dataclass
class RepairUnit:
    candidates: Sequence[Candidate]
    target: tuple = (0, 0)
    preferred: Optional[int] = None
    fixed: bool = False
    cells: int = 1
    first_cell: int = 0


@dataclass
class RepairProblem:
    width: int
    height: int
    object_xy: Sequence[tuple]
    units: Sequence[RepairUnit]
    wire_capacity: Sequence[int] = field(default_factory=list)

    def struct(self) -> texo_gp_repair_problem:
        """The C view of the problem (the arrays it points to live on self)."""
        u = self.units
        cands = [c for unit in u for c in unit.candidates]
        a = {}
        a["object_xy"] = np.array([v for xy in self.object_xy for v in xy] or [0], np.uint32)
        a["wire_capacity"] = np.array(list(self.wire_capacity) or [0], np.uint32)
        a["unit_cells"] = np.array([x.cells for x in u] or [1], np.uint32)
        a["unit_first_cell"] = np.array([x.first_cell for x in u] or [0], np.uint64)
        a["unit_target_xy"] = np.array([v for x in u for v in x.target] or [0], np.uint32)
        a["unit_fixed"] = np.array([1 if x.fixed else 0 for x in u] or [0], np.uint8)
        a["preferred"] = np.array([-1 if x.preferred is None else x.preferred for x in u] or [0], np.int64)
        a["cand_start"] = np.concatenate([[0], np.cumsum([len(x.candidates) for x in u])]).astype(np.uint64)

        def csr(rows, width):
            start = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.uint64)
            flat = [v for r in rows for item in r for v in ((item,) if width == 1 else item)]
            return start, flat
        s, f = csr([c.objects for c in cands], 1)
        a["cand_object_start"], a["cand_object"] = s, np.array(f or [0], np.uint32)
        s, f = csr([c.shared for c in cands], 2)
        a["cand_shared_start"], a["cand_shared"] = s, np.array(f or [0], np.uint64)
        s, f = csr([c.wires for c in cands], 2)
        a["cand_wire_start"], a["cand_wire"] = s, np.array(f or [0], np.uint64)
        self._arrays = a
        ct = {"object_xy": C.c_uint32, "wire_capacity": C.c_uint32, "unit_cells": C.c_uint32,
              "unit_first_cell": C.c_uint64, "unit_target_xy": C.c_uint32, "unit_fixed": C.c_uint8,
              "preferred": C.c_int64, "cand_start": C.c_uint64, "cand_object_start": C.c_uint64,
              "cand_object": C.c_uint32, "cand_shared_start": C.c_uint64, "cand_shared": C.c_uint64,
              "cand_wire_start": C.c_uint64, "cand_wire": C.c_uint64}
        st = texo_gp_repair_problem()
        st.width, st.height = self.width, self.height
        st.n_objects, st.n_wires, st.n_units = len(self.object_xy), len(self.wire_capacity), len(u)
        for k, t in ct.items():
            setattr(st, k, abi.ptr(a[k], t))
        return st


class RepairError(abi.PlacementError):
    """InvalidPlacement (the repair's projection_error) or NoBel, with the
    first pending unit of a NoBel in `unit`."""

    def __init__(self, status: int, unit: int):
        r = abi.texo_gp_report()
        r.status = status
        super().__init__(r)
        self.unit = unit


def declare(lib, prefix: str = "texo_gp"):
    P = C.POINTER
    f = getattr(lib, f"{prefix}_repair")
    f.argtypes = [P(texo_gp_repair_problem), P(C.c_int64), P(C.c_uint64), P(C.c_int64)]
    f.restype = C.c_int


def _lib():
    lib = load()
    if not getattr(lib, "_repair_declared", False):
        declare(lib)
        lib._repair_declared = True
    return lib


def repair(problem: RepairProblem, lib=None, prefix: str = "texo_gp"):
    """(placed: int64[n_units], the candidate of each unit; stats dict).
    Raises RepairError (InvalidPlacement, NoBel) or PlacementError
    (InvalidArgument)."""
    if lib is None:
        lib = _lib()
    elif not getattr(lib, f"_repair_declared_{prefix}", False):
        declare(lib, prefix)
        setattr(lib, f"_repair_declared_{prefix}", True)
    st = problem.struct()
    n = len(problem.units)
    placed = np.full(max(n, 1), -1, np.int64)
    stats = np.zeros(12, np.uint64)
    err = C.c_int64(-1)
    rc = getattr(lib, f"{prefix}_repair")(C.byref(st), abi.ptr(placed, C.c_int64), abi.ptr(stats, C.c_uint64),
                                          C.byref(err))
    if rc in (abi.STATUS_CODE["InvalidPlacement"], abi.STATUS_CODE["NoBel"]):
        raise RepairError(rc, err.value)
    if rc != 0:
        raise abi.PlacementError.from_status(rc)
    return placed[:n], dict(zip(STATS, (int(v) for v in stats)))
