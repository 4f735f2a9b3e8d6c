"""FP64 CPU oracle (TEST INFRASTRUCTURE ONLY).

Loaded by the tests, `__graft_entry__.smoke()` and the benchmark's CPU-baseline
leg as the checker.  The product (`texo_amd`) never imports this package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_gp.so")


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def load():
    """Loads the oracle library, building it first if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    from texo_amd import abi
    lib = C.CDLL(LIB_PATH)
    abi.declare_session_api(lib, "oracle")
    d, P, u32, u64 = C.c_double, C.POINTER, C.c_uint32, C.c_uint64
    lib.oracle_set_threads.argtypes = [C.c_int]
    lib.oracle_poisson_solve.argtypes = [u32, u32, P(d), P(d), P(d)]
    lib.oracle_poisson_solve_matrix.argtypes = [u32, u32, P(d), P(d), P(d)]
    lib.oracle_bspline_axis.argtypes = [u32, d, P(u32), P(d), P(d)]
    lib.oracle_bspline_axis.restype = None
    lib.oracle_deposit_bspline.argtypes = [P(d), u32, u32, d, d, d]
    lib.oracle_deposit_bspline.restype = None
    lib.oracle_sample_bspline.argtypes = [P(d), u32, u32, d, d, P(d)]
    lib.oracle_sample_bspline.restype = None
    lib.oracle_weighted_average.argtypes = [u32, u32, P(d), P(C.c_uint8), P(d), P(d), u32, P(u32), P(u32),
                                            P(u32), u32, P(u32), P(u32), P(u64), d, P(d), P(d), P(d)]
    lib.oracle_wirelength_preconditioner.argtypes = [u32, u32, P(u32), P(u32), u32, P(u32), P(u64), P(d), P(d)]
    lib.oracle_box_metric_kkt_residual.argtypes = [u64, P(d), P(d), P(d), P(d), P(d)]
    lib.oracle_box_metric_kkt_residual.restype = d
    for name, args in [("field_wirelength_gamma", [d, d, d]), ("multiplier_growth", [d]),
                       ("density_overflow_target", [u32]), ("routability_optimized_area", [d, d, d]),
                       ("soft_division_ceiling", [d, d])]:
        fn = getattr(lib, "oracle_" + name)
        fn.argtypes = args
        fn.restype = d
    lib.oracle_weighted_gamma_mean.argtypes = [u32, P(d), P(d), d]
    lib.oracle_weighted_gamma_mean.restype = d
    lib.oracle_augmented_density.argtypes = [d, d, d, P(d), P(d)]
    lib.oracle_augmented_density.restype = None
    lib.oracle_rounded_coordinate.argtypes = [d, u32]
    lib.oracle_rounded_coordinate.restype = u32
    lib.oracle_continuous_routing_demand.argtypes = [u32, u32, u64, P(d), P(d), P(d)]
    lib.oracle_session_hpwl.argtypes = [C.c_void_p, P(d)]
    lib.oracle_session_hpwl.restype = d
    lib.oracle_session_combined.argtypes = [C.c_void_p, P(d), d, P(d), P(d), P(d), P(d)]
    lib.oracle_session_preconditioner.argtypes = [C.c_void_p, P(d), P(d)]
    lib.oracle_clustering_areas.argtypes = [u64, P(u32), P(d), u64, P(u32), P(u64), P(u64), u64, P(d)]
    lib.oracle_adjust_instance_areas.argtypes = [u64, C.c_void_p, u64, P(u32), P(d), P(d), P(d),
                                                 P(C.c_int)]
    i64, u8, i = C.c_int64, C.c_uint8, C.c_int
    lib.oracle_density_fillers.argtypes = [u32, u32, u32, P(u32), P(u32), u32, P(u32), C.c_void_p, P(d), u64, P(u32),
                                           P(d), P(d)]
    lib.oracle_density_fillers.restype = i64
    lib.oracle_density_evaluate.argtypes = [u32, u32, u32, P(u32), P(u32), u32, P(u32), C.c_void_p, P(d), u64, P(u32),
                                            P(d), P(d), u32, C.c_void_p, P(d), P(d)]
    lib.oracle_density_evaluate.restype = i64
    for name in ("density_is_converged", "area_adjustment_is_ready"):
        getattr(lib, "oracle_" + name).argtypes = [u32, P(u32), P(d)]
    lib.oracle_density_kind_is_exchangeable.argtypes = [u32]
    lib.oracle_routability_adjusted_member_area.argtypes = [u32, d, d, d]
    lib.oracle_routability_adjusted_member_area.restype = d
    lib.oracle_area_adjusted_multiplier_step.argtypes = [u32, P(d)]
    lib.oracle_area_adjusted_multiplier_step.restype = d
    lib.oracle_area_adjusted_multipliers.argtypes = [u32, P(d), P(d), P(d), u64, P(d), P(d), P(d)]
    lib.oracle_stationary_target_is_fixed.argtypes = [P(i), P(u32), i, P(u32), u64]
    lib.oracle_stop_without_convergence.argtypes = [i, u64, u32, P(u32), P(d), C.c_void_p]
    lib.oracle_iteration_limit_reached.argtypes = [u64]
    lib.oracle_checkpoint_choice.argtypes = [u32, u32, P(u32), P(d)]
    lib.oracle_checkpoint_choice.restype = u32
    lib.oracle_session_gamma_weights.argtypes = [C.c_void_p, P(d)]
    lib.oracle_net_bounding_boxes.argtypes = [u32, u32, P(d), P(u8), P(d), P(d), u32, P(u32), P(u32), P(u32), P(d),
                                              P(u64)]
    lib.oracle_nesterov_step_lengths.argtypes = [C.POINTER(NesterovState), u64, u64, P(d), u64, P(d), u64, P(d),
                                                 EVALUATE_FN, C.c_void_p, C.POINTER(NesterovResult)]
    lib.oracle_nesterov_step.argtypes = [C.POINTER(NesterovState), u64, P(d), P(d), P(d), EVALUATE_FN, C.c_void_p,
                                         C.POINTER(NesterovResult)]
    from texo_amd import builders
    builders.declare(_Named(lib), "oracle")
    lib.oracle_criticality_weight.argtypes = [u64, u64, u64, u64, u32]
    lib.oracle_criticality_weight.restype = u64
    _lib = lib
    return lib


class _Named:
    """Item access to a CDLL's functions by name (for shared declarations)."""

    def __init__(self, lib):
        self.lib = lib

    def __getitem__(self, name):
        return getattr(self.lib, name)


class OracleAreaMember(C.Structure):
    _fields_ = [("unit", C.c_uint32), ("kind", C.c_uint32), ("current_area", C.c_double),
                ("routability_area", C.c_double), ("pin_area", C.c_double),
                ("clustering_area", C.c_double)]


class NesterovState(C.Structure):
    _fields_ = [("coordinates", C.POINTER(C.c_double)), ("previous_coordinates", C.POINTER(C.c_double)),
                ("momentum", C.c_double), ("lipschitz", C.c_double), ("has_lipschitz", C.c_int)]


class NesterovResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("accepted", C.c_int32), ("objective", C.c_double),
                ("stationarity", C.c_double), ("coordinate_change", C.c_double),
                ("line_search_trials", C.c_uint64), ("evaluations", C.c_uint64),
                ("payload_evaluation", C.c_int64), ("error_index", C.c_int64)]


EVALUATE_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double), C.c_uint64, C.POINTER(C.c_double),
                          C.POINTER(C.c_double))


class DensityMember(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("offset_x", C.c_double), ("offset_y", C.c_double), ("charge", C.c_double)]


class FieldResult(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("available", C.c_uint64), ("real_charge", C.c_uint64),
                ("real_area", C.c_double), ("filler_charge", C.c_double), ("energy", C.c_double),
                ("overflow", C.c_double), ("net_charge", C.c_double), ("force_l1", C.c_double)]
