"""ctypes mirror of include/texo_gp.h.

The structs here are byte-for-byte the C structs of the drop-in boundary.  A
`Problem` owns the numpy arrays a `texo_gp_problem` borrows, so the struct
stays valid for as long as the Problem object lives.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Dict, Optional

import numpy as np

ABI_VERSION = 1
MAX_FIELDS = 4

KIND_LOGIC = 0x000
KIND_REGISTER = 0x200
KIND_MEMORY = 0x300
KIND_CLOCK = 0x400
KIND_IO = 0x500
KIND_CONSTANT = 0x600
KIND_DSP = 0x700


def KIND_LUT(n: int) -> int:
    return 0x100 + n


DCT_FP32 = 0
DCT_BF16X3 = 1
DCT_FP16_SCALED = 2
DCT_NAMES = {DCT_FP32: "fp32", DCT_BF16X3: "bf16x3", DCT_FP16_SCALED: "fp16_scaled"}

STATUS = {
    0: "OK",
    10: "DensityFieldChanged", 11: "InvalidNormalization", 12: "DidNotConverge",
    100: "EmptyGrid", 101: "GridTooLarge", 102: "DensityLength", 103: "NonFiniteDensity",
    104: "EmptyUnit", 105: "NonFiniteUnitCoordinate", 106: "InvalidMemberCharge",
    107: "PointOutsideGrid", 108: "FixedCapacityExhausted", 109: "InsufficientCapacity",
    110: "InsufficientArea", 111: "NonFiniteFiller", 112: "NegativeFillerCharge",
    113: "UnexpectedFillerKind", 114: "FillerChargeMismatch",
    200: "InvalidBounds", 201: "InvalidMetricLength", 202: "NonPositiveMetric",
    203: "InvalidInitialCoordinate", 204: "InvalidMomentumParameter",
    205: "InvalidLipschitzEstimate", 206: "InvalidObjective", 207: "InvalidGradientLength",
    208: "NonFiniteGradient", 209: "LineSearchOverflow", 210: "IterationOverflow",
    211: "Evaluation",
    300: "InvalidMemberArea", 301: "InvalidFillerArea", 302: "PostconditionViolation",
    303: "ArithmeticOverflow",
    400: "AuctionArithmeticOverflow", 401: "ObjectOutOfRange", 402: "BaseCostOutOfRange",
    403: "DuplicateBestObject",
    410: "InvalidPlacement", 411: "NoBel",
    900: "InvalidArgument", 901: "Device", 902: "OutOfMemory", 903: "Unsupported",
    904: "Cancelled", 905: "NotReady", 906: "Injected",
}
STATUS_CODE = {name: code for code, name in STATUS.items()}

_p = C.POINTER


class texo_gp_problem(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32),
        ("n_cells", C.c_uint32), ("cell_unit", _p(C.c_uint32)), ("cell_fixed", _p(C.c_uint8)),
        ("cell_fixed_xy", _p(C.c_double)), ("cell_offset_xy", _p(C.c_double)),
        ("n_units", C.c_uint32), ("unit_fixed", _p(C.c_uint8)),
        ("unit_initial_xy", _p(C.c_double)), ("unit_bounds", _p(C.c_double)),
        ("unit_member_start", _p(C.c_uint32)), ("member_kind", _p(C.c_uint16)),
        ("member_offset_xy", _p(C.c_double)), ("member_cell", _p(C.c_uint32)),
        ("n_capacity_kinds", C.c_uint32), ("capacity_kind", _p(C.c_uint16)),
        ("capacity", _p(C.c_uint32)),
        ("n_fixed_occupancy", C.c_uint32), ("fixed_occupancy_kind", _p(C.c_uint16)),
        ("fixed_occupancy_xy", _p(C.c_uint32)),
        ("n_nets", C.c_uint32), ("net_pin_start", _p(C.c_uint32)), ("pin_cell", _p(C.c_uint32)),
        ("n_arcs", C.c_uint32), ("arc_cells", _p(C.c_uint32)), ("arc_extra_weight", _p(C.c_uint64)),
        ("routing_horizontal", _p(C.c_uint32)), ("routing_vertical", _p(C.c_uint32)),
        ("n_register_controls", C.c_uint32), ("register_cell", _p(C.c_uint32)),
        ("register_clock_lsr", _p(C.c_uint64)), ("register_ce", _p(C.c_uint64)),
        ("best_effort", C.c_uint8),
    ]


class texo_gp_options(C.Structure):
    _fields_ = [
        ("dct_precision", C.c_uint32), ("fixed_iterations", C.c_uint32),
        ("max_iterations", C.c_uint32), ("trace", C.c_uint32), ("gamma_base", C.c_double),
        ("device", C.c_int32), ("inject_failure_at_evaluation", C.c_uint32),
    ]


class texo_gp_report(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("err_index", C.c_int64), ("err_index2", C.c_int64),
        ("err_kind", C.c_uint32), ("err_x", C.c_uint32), ("err_y", C.c_uint32),
        ("err_value", C.c_double * 2),
        ("iterations", C.c_uint64), ("evaluations", C.c_uint64),
        ("line_search_trials", C.c_uint64), ("area_rounds", C.c_uint64),
        ("dct_precision", C.c_uint32), ("fixed_iterations", C.c_uint32),
        ("gamma_base", C.c_double), ("final_wirelength", C.c_double), ("final_hpwl", C.c_double),
        ("n_fields", C.c_uint32), ("field_kind", C.c_uint32 * MAX_FIELDS),
        ("field_overflow", C.c_double * MAX_FIELDS), ("field_energy", C.c_double * MAX_FIELDS),
        ("elapsed_seconds", C.c_double),
    ]

    def as_dict(self) -> dict:
        n = self.n_fields
        return {
            "status": STATUS.get(self.status, str(self.status)),
            "iterations": self.iterations, "evaluations": self.evaluations,
            "line_search_trials": self.line_search_trials, "area_rounds": self.area_rounds,
            "final_wirelength": self.final_wirelength, "final_hpwl": self.final_hpwl,
            "field_kind": list(self.field_kind[:n]),
            "field_overflow": list(self.field_overflow[:n]),
            "field_energy": list(self.field_energy[:n]),
            "elapsed_seconds": self.elapsed_seconds,
        }


class texo_gp_field_eval(C.Structure):
    _fields_ = [
        ("energy", C.c_double), ("normalized_overflow", C.c_double), ("net_charge", C.c_double),
        ("force_l1", C.c_double), ("real_area", C.c_double), ("filler_charge", C.c_double),
    ]


class texo_gp_schedule_state(C.Structure):
    _fields_ = [
        ("n_fields", C.c_uint32), ("multiplier", C.c_double * MAX_FIELDS), ("multiplier_step", C.c_double),
        ("initial_energy", C.c_double * MAX_FIELDS), ("gamma", C.c_double),
        ("gamma_weight", C.c_double * MAX_FIELDS), ("momentum", C.c_double), ("lipschitz", C.c_double),
    ]

    def as_dict(self) -> dict:
        n = self.n_fields
        return {"multiplier": list(self.multiplier[:n]), "multiplier_step": self.multiplier_step,
                "initial_energy": list(self.initial_energy[:n]), "gamma": self.gamma,
                "gamma_weight": list(self.gamma_weight[:n]), "momentum": self.momentum,
                "lipschitz": self.lipschitz}


def ptr(a: Optional[np.ndarray], ctype):
    if a is None:
        return C.POINTER(ctype)()
    assert a.flags["C_CONTIGUOUS"], "arrays crossing the ABI must be C-contiguous"
    return a.ctypes.data_as(C.POINTER(ctype))


@dataclass
class Problem:
    """Flattened placement problem (numpy SoA) matching texo_gp_problem."""

    width: int
    height: int
    cell_unit: np.ndarray
    cell_fixed: np.ndarray
    cell_fixed_xy: np.ndarray
    cell_offset_xy: np.ndarray
    unit_fixed: np.ndarray
    unit_initial_xy: np.ndarray
    unit_bounds: np.ndarray
    unit_member_start: np.ndarray
    member_kind: np.ndarray
    member_offset_xy: np.ndarray
    member_cell: np.ndarray
    capacity_kind: np.ndarray
    capacity: np.ndarray
    net_pin_start: np.ndarray
    pin_cell: np.ndarray
    fixed_occupancy_kind: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint16))
    fixed_occupancy_xy: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    arc_cells: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    arc_extra_weight: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint64))
    routing_horizontal: Optional[np.ndarray] = None
    routing_vertical: Optional[np.ndarray] = None
    register_cell: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    register_clock_lsr: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint64))
    register_ce: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint64))
    best_effort: bool = False
    meta: Dict = field(default_factory=dict)

    def __post_init__(self):
        casts = {
            "cell_unit": np.uint32, "cell_fixed": np.uint8, "cell_fixed_xy": np.float64,
            "cell_offset_xy": np.float64, "unit_fixed": np.uint8, "unit_initial_xy": np.float64,
            "unit_bounds": np.float64, "unit_member_start": np.uint32, "member_kind": np.uint16,
            "member_offset_xy": np.float64, "member_cell": np.uint32, "capacity_kind": np.uint16,
            "capacity": np.uint32, "net_pin_start": np.uint32, "pin_cell": np.uint32,
            "fixed_occupancy_kind": np.uint16, "fixed_occupancy_xy": np.uint32,
            "arc_cells": np.uint32, "arc_extra_weight": np.uint64,
            "register_cell": np.uint32, "register_clock_lsr": np.uint64, "register_ce": np.uint64,
        }
        for name, dt in casts.items():
            setattr(self, name, np.ascontiguousarray(np.asarray(getattr(self, name)).reshape(-1), dtype=dt))
        for name in ("routing_horizontal", "routing_vertical"):
            v = getattr(self, name)
            if v is not None:
                setattr(self, name, np.ascontiguousarray(np.asarray(v).reshape(-1), dtype=np.uint32))
        self._struct = None

    @property
    def n_units(self) -> int:
        return int(self.unit_fixed.size)

    @property
    def n_cells(self) -> int:
        return int(self.cell_unit.size)

    @property
    def n_nets(self) -> int:
        return int(self.net_pin_start.size - 1)

    @property
    def n_members(self) -> int:
        return int(self.member_kind.size)

    def struct(self) -> texo_gp_problem:
        s = texo_gp_problem()
        s.abi_version = ABI_VERSION
        s.width, s.height = self.width, self.height
        s.n_cells = self.n_cells
        s.cell_unit = ptr(self.cell_unit, C.c_uint32)
        s.cell_fixed = ptr(self.cell_fixed, C.c_uint8)
        s.cell_fixed_xy = ptr(self.cell_fixed_xy, C.c_double)
        s.cell_offset_xy = ptr(self.cell_offset_xy, C.c_double)
        s.n_units = self.n_units
        s.unit_fixed = ptr(self.unit_fixed, C.c_uint8)
        s.unit_initial_xy = ptr(self.unit_initial_xy, C.c_double)
        s.unit_bounds = ptr(self.unit_bounds, C.c_double)
        s.unit_member_start = ptr(self.unit_member_start, C.c_uint32)
        s.member_kind = ptr(self.member_kind, C.c_uint16)
        s.member_offset_xy = ptr(self.member_offset_xy, C.c_double)
        s.member_cell = ptr(self.member_cell, C.c_uint32)
        s.n_capacity_kinds = self.capacity_kind.size
        s.capacity_kind = ptr(self.capacity_kind, C.c_uint16)
        s.capacity = ptr(self.capacity, C.c_uint32)
        s.n_fixed_occupancy = self.fixed_occupancy_kind.size
        s.fixed_occupancy_kind = ptr(self.fixed_occupancy_kind, C.c_uint16)
        s.fixed_occupancy_xy = ptr(self.fixed_occupancy_xy, C.c_uint32)
        s.n_nets = self.n_nets
        s.net_pin_start = ptr(self.net_pin_start, C.c_uint32)
        s.pin_cell = ptr(self.pin_cell, C.c_uint32)
        s.n_arcs = self.arc_extra_weight.size
        s.arc_cells = ptr(self.arc_cells, C.c_uint32)
        s.arc_extra_weight = ptr(self.arc_extra_weight, C.c_uint64)
        s.routing_horizontal = ptr(self.routing_horizontal, C.c_uint32)
        s.routing_vertical = ptr(self.routing_vertical, C.c_uint32)
        s.n_register_controls = self.register_cell.size
        s.register_cell = ptr(self.register_cell, C.c_uint32)
        s.register_clock_lsr = ptr(self.register_clock_lsr, C.c_uint64)
        s.register_ce = ptr(self.register_ce, C.c_uint64)
        s.best_effort = 1 if self.best_effort else 0
        self._struct = s
        return s


def options(dct_precision: int = DCT_FP32, fixed_iterations: int = 0, max_iterations: int = 0,
            trace: bool = False, gamma_base: float = 0.0, device: int = 0,
            inject_failure_at_evaluation: int = 0) -> texo_gp_options:
    o = texo_gp_options()
    o.dct_precision = dct_precision
    o.fixed_iterations = fixed_iterations
    o.max_iterations = max_iterations
    o.trace = 1 if trace else 0
    o.gamma_base = gamma_base
    o.device = device
    o.inject_failure_at_evaluation = inject_failure_at_evaluation
    return o


def declare_session_api(lib, prefix: str):
    """Declares the shared session-level signatures for `prefix` in
    {"texo_gp", "oracle"} (the oracle mirrors the product one-for-one)."""
    P = C.POINTER
    d = C.c_double
    sess = C.c_void_p

    def f(name, restype, *argtypes):
        fn = getattr(lib, f"{prefix}_{name}")
        fn.restype = restype
        fn.argtypes = list(argtypes)
        return fn

    f("place", C.c_int, P(texo_gp_problem), P(texo_gp_options), P(C.c_uint32), P(texo_gp_report))
    f("session_create", sess, P(texo_gp_problem), P(texo_gp_options), P(texo_gp_report))
    f("session_destroy", None, sess)
    f("session_sizes", C.c_int, sess, P(C.c_uint64), P(C.c_uint32), P(C.c_uint64), P(C.c_uint64))
    f("session_initial", C.c_int, sess, P(d), P(d), P(d), P(C.c_uint32), P(C.c_uint32), P(d))
    f("session_evaluate", C.c_int, sess, P(d), d, P(d), P(d), P(texo_gp_field_eval), P(d), P(d), P(d), P(d))
    f("session_run", C.c_int, sess, P(C.c_uint32), P(texo_gp_report))
    f("session_iterate", C.c_int, sess, C.c_uint32, P(texo_gp_report))
    f("session_coordinates", C.c_int, sess, P(d))
    f("session_rudy", C.c_int, sess, P(d), P(d), P(d))
    f("session_combined", C.c_int, sess, P(d), d, P(d), P(d), P(d), P(d))
    f("session_preconditioner", C.c_int, sess, P(d), P(d))
    f("session_schedule", C.c_int, sess, P(texo_gp_schedule_state))
    f("session_area_proposal", C.c_int, sess, P(d), P(d), P(d), P(C.c_int32), P(d))


class Session:
    """Evaluate-level handle over either library (same contract)."""

    def __init__(self, lib, prefix: str, problem: Problem, opts: Optional[texo_gp_options] = None):
        self.lib, self.prefix, self.problem = lib, prefix, problem
        self.opts = opts if opts is not None else options()
        self.report = texo_gp_report()
        self.handle = getattr(lib, f"{prefix}_session_create")(
            C.byref(problem.struct()), C.byref(self.opts), C.byref(self.report))
        if not self.handle:
            raise PlacementError(self.report)
        nc, nf, nfl, ndu = C.c_uint64(), C.c_uint32(), C.c_uint64(), C.c_uint64()
        getattr(lib, f"{prefix}_session_sizes")(self.handle, C.byref(nc), C.byref(nf), C.byref(nfl),
                                                C.byref(ndu))
        self.n_coordinates, self.n_fields = nc.value, nf.value
        self.n_fillers, self.n_density_units = nfl.value, ndu.value
        self.initial = np.zeros(self.n_coordinates)
        self.lower = np.zeros(self.n_coordinates)
        self.upper = np.zeros(self.n_coordinates)
        self.field_kind = np.zeros(max(self.n_fields, 1), np.uint32)
        self.filler_field = np.zeros(max(self.n_fillers, 1), np.uint32)
        self.filler_charge = np.zeros(max(self.n_fillers, 1))
        getattr(lib, f"{prefix}_session_initial")(
            self.handle, ptr(self.initial, C.c_double), ptr(self.lower, C.c_double),
            ptr(self.upper, C.c_double), ptr(self.field_kind, C.c_uint32),
            ptr(self.filler_field, C.c_uint32), ptr(self.filler_charge, C.c_double))

    def close(self):
        if self.handle:
            getattr(self.lib, f"{self.prefix}_session_destroy")(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def evaluate(self, coordinates: np.ndarray, gamma: float, want_grids: bool = False) -> dict:
        x = np.ascontiguousarray(coordinates, dtype=np.float64)
        assert x.size == self.n_coordinates
        F, DU, Fl = self.n_fields, self.n_density_units, self.n_fillers
        G = self.problem.width * self.problem.height
        wa = C.c_double()
        wag = np.zeros(2 * self.problem.n_units)
        fields = (texo_gp_field_eval * max(F, 1))()
        ug = np.zeros(max(F * DU * 2, 1))
        fg = np.zeros(max(2 * Fl, 1))
        dens = np.zeros(F * G) if want_grids else None
        pot = np.zeros(F * G) if want_grids else None
        rc = getattr(self.lib, f"{self.prefix}_session_evaluate")(
            self.handle, ptr(x, C.c_double), gamma, C.byref(wa), ptr(wag, C.c_double), fields,
            ptr(ug, C.c_double), ptr(fg, C.c_double), ptr(dens, C.c_double), ptr(pot, C.c_double))
        if rc != 0:
            raise PlacementError.from_status(rc)
        out = {
            "wa_value": wa.value, "wa_gradient": wag,
            "energy": np.array([fields[i].energy for i in range(F)]),
            "overflow": np.array([fields[i].normalized_overflow for i in range(F)]),
            "net_charge": np.array([fields[i].net_charge for i in range(F)]),
            "force_l1": np.array([fields[i].force_l1 for i in range(F)]),
            "real_area": np.array([fields[i].real_area for i in range(F)]),
            "filler_charge": np.array([fields[i].filler_charge for i in range(F)]),
            "unit_gradient": ug[: F * DU * 2].reshape(F, DU, 2),
            "filler_gradient": fg[: 2 * Fl].reshape(Fl, 2),
        }
        if want_grids:
            H, W = self.problem.height, self.problem.width
            out["density"] = dens.reshape(F, H, W)
            out["potential"] = pot.reshape(F, H, W)
        return out

    def run(self):
        t = np.zeros(2 * self.problem.n_units, np.uint32)
        rep = texo_gp_report()
        rc = getattr(self.lib, f"{self.prefix}_session_run")(self.handle, ptr(t, C.c_uint32), C.byref(rep))
        if rc != 0:
            raise PlacementError(rep)
        return t.reshape(-1, 2), rep

    def iterate(self, iterations: int):
        rep = texo_gp_report()
        rc = getattr(self.lib, f"{self.prefix}_session_iterate")(self.handle, iterations, C.byref(rep))
        if rc != 0:
            raise PlacementError(rep)
        return rep

    def coordinates(self) -> np.ndarray:
        x = np.zeros(self.n_coordinates)
        getattr(self.lib, f"{self.prefix}_session_coordinates")(self.handle, ptr(x, C.c_double))
        return x

    def rudy(self, coordinates: np.ndarray):
        G = self.problem.width * self.problem.height
        h, v = np.zeros(G), np.zeros(G)
        x = np.ascontiguousarray(coordinates, dtype=np.float64)
        rc = getattr(self.lib, f"{self.prefix}_session_rudy")(
            self.handle, ptr(x, C.c_double), ptr(h, C.c_double), ptr(v, C.c_double))
        if rc != 0:
            raise PlacementError.from_status(rc)
        shape = (self.problem.height, self.problem.width)
        return h.reshape(shape), v.reshape(shape)


    def combined(self, coordinates: np.ndarray, gamma: float, multipliers, scales):
        """(value, gradient[C]) of the combined objective with explicit
        multipliers and initial-energy scales."""
        x = np.ascontiguousarray(coordinates, dtype=np.float64)
        m = np.ascontiguousarray(multipliers, dtype=np.float64)
        sc = np.ascontiguousarray(scales, dtype=np.float64)
        g = np.zeros(self.n_coordinates)
        v = C.c_double()
        rc = getattr(self.lib, f"{self.prefix}_session_combined")(
            self.handle, ptr(x, C.c_double), gamma, ptr(m, C.c_double), ptr(sc, C.c_double), C.byref(v),
            ptr(g, C.c_double))
        if rc != 0:
            raise PlacementError.from_status(rc)
        return v.value, g

    def preconditioner(self, multipliers) -> np.ndarray:
        m = np.ascontiguousarray(multipliers, dtype=np.float64)
        d = np.zeros(self.n_coordinates)
        rc = getattr(self.lib, f"{self.prefix}_session_preconditioner")(self.handle, ptr(m, C.c_double),
                                                                       ptr(d, C.c_double))
        if rc != 0:
            raise PlacementError.from_status(rc)
        return d

    def area_proposal(self, coordinates: np.ndarray) -> dict:
        """propose_area_adjustment at `coordinates` (nothing is applied)."""
        x = np.ascontiguousarray(coordinates, dtype=np.float64)
        ma = np.zeros(max(self.problem.n_members, 1))
        fa = np.zeros(max(self.n_fillers, 1))
        reset, inc = C.c_int32(), C.c_double()
        rc = getattr(self.lib, f"{self.prefix}_session_area_proposal")(
            self.handle, ptr(x, C.c_double), ptr(ma, C.c_double), ptr(fa, C.c_double), C.byref(reset),
            C.byref(inc))
        if rc != 0:
            raise PlacementError.from_status(rc)
        return {"member_areas": ma[:self.problem.n_members], "filler_areas": fa[:self.n_fillers],
                "requires_reset": bool(reset.value), "total_relative_increase": inc.value}

    def schedule(self) -> dict:
        st = texo_gp_schedule_state()
        rc = getattr(self.lib, f"{self.prefix}_session_schedule")(self.handle, C.byref(st))
        if rc != 0:
            raise PlacementError.from_status(rc)
        return st.as_dict()


class PlacementError(RuntimeError):
    """A typed failure of the placement call (mirrors EplaceError)."""

    def __init__(self, report: texo_gp_report):
        self.status = report.status
        self.name = STATUS.get(report.status, str(report.status))
        self.index = report.err_index
        self.index2 = report.err_index2
        self.kind = report.err_kind
        self.point = (report.err_x, report.err_y)
        self.values = (report.err_value[0], report.err_value[1])
        super().__init__(f"{self.name} (status {self.status}, index {self.index}, index2 {self.index2}, "
                         f"kind {self.kind:#x}, point {self.point}, values {self.values})")

    @classmethod
    def from_status(cls, status: int):
        r = texo_gp_report()
        r.status = status
        return cls(r)


class texo_gp_server_stats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64), ("admitted", C.c_uint64), ("completed", C.c_uint64),
        ("failed", C.c_uint64), ("rounds", C.c_uint64), ("active", C.c_uint64),
        ("pending", C.c_uint64), ("hbm_in_use", C.c_uint64), ("busy_seconds", C.c_double),
        ("admit_seconds", C.c_double), ("retire_seconds", C.c_double),
        ("wait_seconds", C.c_double), ("launch_seconds", C.c_double), ("graph_rounds", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


def declare_server_api(lib, prefix: str):
    """Declares the continuous-batching server entry points of the product library."""
    P = C.POINTER

    def f(name, restype, *argtypes):
        fn = getattr(lib, f"{prefix}_{name}")
        fn.restype = restype
        fn.argtypes = list(argtypes)

    f("server_create", C.c_void_p, C.c_int32, C.c_uint64, C.c_uint32)
    f("server_submit", C.c_int64, C.c_void_p, P(texo_gp_problem), P(texo_gp_options))
    f("server_poll", C.c_int, C.c_void_p, C.c_int64, P(C.c_uint32), P(texo_gp_report))
    f("server_wait", C.c_int, C.c_void_p, C.c_int64, P(C.c_uint32), P(texo_gp_report))
    f("server_cancel", C.c_int, C.c_void_p, C.c_int64)
    f("server_stats_get", C.c_int, C.c_void_p, P(texo_gp_server_stats))
    f("server_destroy", None, C.c_void_p)


def declare_batch_api(lib, prefix: str):
    """Declares the lockstep batch entry points of the product library."""
    P = C.POINTER

    def f(name, restype, *argtypes):
        fn = getattr(lib, f"{prefix}_{name}")
        fn.restype = restype
        fn.argtypes = list(argtypes)

    f("batch_create", C.c_void_p, C.c_int32)
    f("batch_add", C.c_int64, C.c_void_p, P(texo_gp_problem), P(texo_gp_options), P(texo_gp_report))
    f("batch_run", C.c_int, C.c_void_p)
    f("batch_iterate", C.c_int, C.c_void_p, C.c_uint32)
    f("batch_result", C.c_int, C.c_void_p, C.c_uint64, P(C.c_uint32), P(texo_gp_report))
    f("batch_sizes", C.c_int, C.c_void_p, C.c_uint64, P(C.c_uint64), P(C.c_uint32), P(C.c_uint64), P(C.c_uint64))
    f("batch_coordinates", C.c_int, C.c_void_p, C.c_uint64, P(C.c_double))
    f("batch_set_kernel_timing", C.c_int, C.c_void_p, C.c_char_p, C.c_int)
    f("batch_kernel_timing", C.c_int, C.c_void_p, P(C.c_double), P(C.c_uint64), P(C.c_double))
    f("batch_destroy", None, C.c_void_p)
