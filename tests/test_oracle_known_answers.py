"""Pins the oracle's schedule, stop, checkpoint, density-model and Nesterov
helpers to the reference's own known-answer tests (CPU only).

Each test names the reference test it restates (file:line).  Together with
tests/test_oracle.py these give every oracle function a check that a dropped
term, a wrong sign or a swapped operand would fail.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from texo_amd import abi
from tests import problems

D = C.c_double
LUT4, FF, MEM = problems.LUT4, problems.FF, abi.KIND_MEMORY


def arr(ctype, values):
    return (ctype * max(len(values), 1))(*values)


def close(actual, expected, tol):
    # assert_close of the reference tests: |a - e| <= tol * max(|e|, 1)
    return abs(actual - expected) <= tol * max(abs(expected), 1.0)


# --------------------------------------------------------------- eplace.rs --
def test_area_adjusted_multiplier_step(oracle_lib):
    # multiplier_growth_matches_elfplace_endpoints_and_initial_step (eplace.rs:1698-1723)
    step = oracle_lib.oracle_area_adjusted_multiplier_step(2, arr(D, [3.0, 4.0]))
    assert close(step, 0.30, 1e-15)
    assert np.isnan(oracle_lib.oracle_area_adjusted_multiplier_step(2, arr(D, [0.0, 0.0])))


def fields(kinds, overflows):
    return len(kinds), arr(C.c_uint32, kinds), arr(D, overflows)


def test_scarce_memory_overflow_has_an_independent_stop(oracle_lib):
    # scarce_memory_overflow_has_an_independent_stop_condition (eplace.rs:1751-1759)
    assert oracle_lib.oracle_density_is_converged(*fields([LUT4, MEM], [0.09, 0.21])) == 0
    assert oracle_lib.oracle_density_is_converged(*fields([LUT4, MEM], [0.09, 0.20])) == 1
    # a LUT field above its 0.10 target alone blocks convergence
    assert oracle_lib.oracle_density_is_converged(*fields([LUT4, MEM], [0.1000001, 0.0])) == 0


def test_area_adjustment_starts_at_fifteen_percent(oracle_lib):
    # area_adjustment_starts_at_lut_and_register_fifteen_percent (eplace.rs:1761-1776)
    assert oracle_lib.oracle_area_adjustment_is_ready(*fields([LUT4, FF, MEM], [0.15, 0.149, 0.99])) == 1
    assert oracle_lib.oracle_area_adjustment_is_ready(*fields([LUT4, FF], [0.151, 0.149])) == 0
    assert oracle_lib.oracle_area_adjustment_is_ready(*fields([LUT4, FF], [0.15, 0.1500001])) == 0


def test_rudy_inflates_slice_area_but_not_memory(oracle_lib):
    # rudy_inflates_exchangeable_slice_area_but_not_hard_memory_charge (eplace.rs:1778-1797)
    s2 = np.sqrt(2.0)
    f = oracle_lib.oracle_routability_adjusted_member_area
    assert close(f(LUT4, 1.0, s2, 0.0), 2.0, np.finfo(float).eps)
    assert close(f(FF, 1.0, 0.0, s2), 2.0, np.finfo(float).eps)
    assert f(MEM, 1.0, s2, s2) == 1.0
    # below the knee the area scales with the squared larger utilisation
    assert f(LUT4, 2.0, 0.5, 0.25) == 2.0 * 0.25


def test_area_adjusted_multiplier_matches_force_balance(oracle_lib):
    # area_adjusted_multiplier_matches_elfplace_force_balance (eplace.rs:1799-1822)
    m = (D * 2)()
    rc = oracle_lib.oracle_area_adjusted_multipliers(2, arr(D, [5.0, 10.0]), arr(D, [2.0, 3.0]), arr(D, [10.0, 20.0]),
                                                     2, arr(D, [3.0, -1.0]), arr(D, [4.0, 0.0]), m)
    assert rc == 0
    assert close(m[0], m[1], 1e-15)
    assert close(2.0 * m[0] + 3.0 * m[1], 0.8, 1e-14)
    # a zero wirelength force cannot be balanced
    rc = oracle_lib.oracle_area_adjusted_multipliers(2, arr(D, [5.0, 10.0]), arr(D, [2.0, 3.0]), arr(D, [10.0, 20.0]),
                                                     2, arr(D, [0.0, 0.0]), arr(D, [0.0, 0.0]), m)
    assert abi.STATUS[rc] == "InvalidNormalization"


def test_stationary_target_must_survive_a_multiplier_update(oracle_lib):
    # stationary_rounded_target_must_survive_a_multiplier_update (eplace.rs:1824-1844)
    has = C.c_int(0)
    prev = (C.c_uint32 * 2)()
    target = arr(C.c_uint32, [3, 5])
    f = oracle_lib.oracle_stationary_target_is_fixed
    assert f(C.byref(has), prev, 0, target, 1) == 0
    assert f(C.byref(has), prev, 0, target, 1) == 1
    assert f(C.byref(has), prev, 1, target, 1) == 0
    assert has.value == 0
    # a different stationary target restarts the comparison
    assert f(C.byref(has), prev, 0, target, 1) == 0
    assert f(C.byref(has), prev, 0, arr(C.c_uint32, [3, 6]), 1) == 0


def test_only_best_effort_keeps_a_nonconverged_checkpoint(oracle_lib):
    # only_best_effort_solves_keep_a_nonconverged_checkpoint (eplace.rs:1846-1857)
    rep = abi.texo_gp_report()
    rc = oracle_lib.oracle_stop_without_convergence(0, 3000, *fields([LUT4], [0.30]), C.byref(rep))
    assert abi.STATUS[rc] == "DidNotConverge" and rep.err_index == 3000 and rep.err_value[0] == 0.30
    assert oracle_lib.oracle_stop_without_convergence(1, 3000, *fields([LUT4], [0.30]), None) == 0


def test_iteration_cap_is_a_hard_boundary(oracle_lib):
    # paper_iteration_cap_is_a_hard_nonconvergence_boundary (eplace.rs:1859-1864)
    f = oracle_lib.oracle_iteration_limit_reached
    assert f(2999) == 0 and f(3000) == 1 and f(2 ** 64 - 1) == 1


def test_checkpoint_retains_the_lowest_overflow_excess(oracle_lib):
    # checkpoint_retains_the_lowest_field_overflow_excess (eplace.rs:1866-1875)
    choice = oracle_lib.oracle_checkpoint_choice
    assert choice(3, 1, arr(C.c_uint32, [MEM]), arr(D, [0.40, 0.25, 0.30])) == 1
    # ties keep the earlier checkpoint; the total excess breaks equal maxima
    assert choice(2, 1, arr(C.c_uint32, [MEM]), arr(D, [0.25, 0.25])) == 0
    assert choice(2, 2, arr(C.c_uint32, [LUT4, FF]), arr(D, [0.30, 0.20, 0.30, 0.15])) == 1
    # below-target overflows carry no excess at all
    assert choice(2, 1, arr(C.c_uint32, [LUT4]), arr(D, [0.05, 0.01])) == 0


def test_rigid_unit_keeps_resource_specific_pin_weights(oracle_lib):
    # rigid_lut_register_unit_keeps_resource_specific_cell_pin_weights (eplace.rs:1877-1903):
    # one LUT+FF unit whose LUT cell has baseline net incidence 10 and whose
    # FF cell has 1 (ten and one 2-pin nets to a fixed IO cell).
    W, H = 2, 1
    cap = {LUT4: np.ones((H, W), np.uint32), FF: np.ones((H, W), np.uint32)}
    nets = [[0, 2]] * 10 + [[1, 2]]
    p = problems.density_problem(W, H, cap, [(0.0, 0.0, [(LUT4, 0.0, 0.0), (FF, 0.0, 0.0)])], nets=nets,
                                 fixed_units=[(1.0, 0.0)])
    s = abi.Session(oracle_lib, "oracle", p)
    w = (D * 2)()
    assert oracle_lib.oracle_session_gamma_weights(C.c_void_p(s.handle), w) == 0
    assert list(s.field_kind) == [LUT4, FF]
    assert w[0] == 10.0 and w[1] == 1.0
    s.close()


def test_only_exchangeable_kinds_get_density_fields(oracle_lib):
    # only_interchangeable_ecp5_resource_classes_get_density_fields (eplace.rs:1905-1914)
    f = oracle_lib.oracle_density_kind_is_exchangeable
    assert f(LUT4) and f(abi.KIND_LUT(6)) and f(FF) and f(MEM)
    for k in (abi.KIND_CLOCK, abi.KIND_LOGIC, abi.KIND_IO, abi.KIND_CONSTANT):
        assert not f(k)


# ------------------------------------------------- analytical_placement.rs --
def test_weighted_average_uses_fixed_coordinates_and_macro_offsets(oracle_lib):
    # weighted_average_uses_fixed_coordinates_and_macro_offsets (analytical_placement.rs:990-1062)
    # cells: root(u0), member(u0), fixed(u1), moving(u2); net member -> fixed, moving
    cells = 4
    positions = [(2.0, 4.0), (99.0, 99.0), (7.0, 8.0)]
    fixed = [0, 0, 1, 0]
    fixed_xy = [(0, 0), (0, 0), (10.0, 5.0), (0, 0)]
    offsets = [(0.0, 0.0), (3.0, -2.0), (0.0, 0.0), (-1.0, 1.0)]

    def run(pos, units, pin_unit, pin_cell, fmask, fxy, offs):
        n = len(pin_cell)
        gx, gy, v = np.zeros(units), np.zeros(units), D()
        u32 = lambda a: arr(C.c_uint32, a)
        rc = oracle_lib.oracle_weighted_average(units, len(fmask), arr(D, [c for p in pos for c in p]),
                                                arr(C.c_uint8, fmask), arr(D, [c for p in fxy for c in p]),
                                                arr(D, [c for p in offs for c in p]), 1, u32([0, n]), u32(pin_unit),
                                                u32(pin_cell), 0, u32([]), u32([]), arr(C.c_uint64, []), 0.25,
                                                C.byref(v), abi.ptr(gx, D), abi.ptr(gy, D))
        assert rc == 0
        return v.value, gx, gy

    obj = run(positions, 3, [0, 1, 2], [1, 2, 3], fixed, fixed_xy, offsets)
    # wide_net(2): source at (5, 2), fixed sink (10, 5), moving sink (6, 9)
    ref = run([(5.0, 2.0), (0.0, 0.0), (6.0, 9.0)], 3, [0, 1, 2], [0, 1, 2], [0, 1, 0],
              [(0, 0), (10.0, 5.0), (0, 0)], [(0, 0)] * 3)
    assert close(obj[0], ref[0], 1e-14)
    for u in (0, 2):
        assert close(obj[1][u], ref[1][u], 1e-14) and close(obj[2][u], ref[2][u], 1e-14)
    assert obj[1][1] == 0.0 and obj[2][1] == 0.0
    moved = list(positions)
    moved[1] = (-1000.0, -1000.0)
    again = run(moved, 3, [0, 1, 2], [1, 2, 3], fixed, fixed_xy, offsets)
    assert again[0] == obj[0] and np.array_equal(again[1], obj[1]) and np.array_equal(again[2], obj[2])


def test_rudy_boxes_skip_internal_carry_nets(oracle_lib):
    # rudy_boxes_exclude_internal_carry_nets_but_keep_every_external_fanout (analytical_placement.rs:1064-1104)
    # cells carry(u0), successor(u0), packed-ff(u0), external-a(u1), external-b(u2)
    units, cells = 3, 5
    positions = arr(D, [4.0, 7.0, 12.0, 2.0, 1.0, 11.0])
    offsets = arr(D, [0.0, 0.0, 3.0, 1.0, 3.0, 1.0, 0.0, 0.0, 0.0, 0.0])
    net_start = arr(C.c_uint32, [0, 2, 6])
    pin_cell = [0, 1, 1, 2, 3, 4]
    cell_unit = [0, 0, 0, 1, 2]
    boxes = (D * 8)()
    n = C.c_uint64()
    oracle_lib.oracle_net_bounding_boxes(units, cells, positions, arr(C.c_uint8, [0] * cells), arr(D, [0.0] * 10),
                                         offsets, 2, net_start, arr(C.c_uint32, [cell_unit[c] for c in pin_cell]),
                                         arr(C.c_uint32, pin_cell), boxes, C.byref(n))
    assert n.value == 1 and list(boxes[:4]) == [1.0, 12.0, 2.0, 11.0]


# ---------------------------------------------- electrostatic_placement.rs --
def fixture(W, H, capacity, units):
    """(kinds, capacity, unit_member_start, members, unit_xy) for the
    explicit-charge density fixture API; units: [(x, y, [(kind, ox, oy, q)])]."""
    kinds = sorted(capacity)
    cap = np.concatenate([np.asarray(capacity[k], np.uint32).reshape(-1) for k in kinds])
    start = np.cumsum([0] + [len(u[2]) for u in units]).astype(np.uint32)
    mem = (oracle.DensityMember * max(int(start[-1]), 1))(
        *[oracle.DensityMember(k, ox, oy, q) for _, _, ms in units for k, ox, oy, q in ms])
    xy = np.array([c for u in units for c in u[:2]], float)
    return np.array(kinds, np.uint32), cap, start, mem, xy


def fixture_fillers(lib, W, H, fx):
    kinds, cap, start, mem, xy = fx
    fk, fxy, fq = np.zeros(64, np.uint32), np.zeros(128), np.zeros(64)
    n = lib.oracle_density_fillers(W, H, len(kinds), abi.ptr(kinds, C.c_uint32), abi.ptr(cap, C.c_uint32),
                                   len(start) - 1, abi.ptr(start, C.c_uint32), mem, abi.ptr(xy, D), 64,
                                   abi.ptr(fk, C.c_uint32), abi.ptr(fxy, D), abi.ptr(fq, D))
    assert n >= 0, abi.STATUS[-n]
    return fk[:n].copy(), fxy[:2 * n].copy(), fq[:n].copy()


def fixture_evaluate(lib, W, H, fx, unit_xy, fillers):
    kinds, cap, start, mem, _ = fx
    fk, fxy, fq = fillers
    U = len(start) - 1
    res = (oracle.FieldResult * 4)()
    ug, fg = np.zeros(4 * U * 2), np.zeros(max(2 * len(fk), 1))
    xy = np.ascontiguousarray(unit_xy, float)
    n = lib.oracle_density_evaluate(W, H, len(kinds), abi.ptr(kinds, C.c_uint32), abi.ptr(cap, C.c_uint32), U,
                                    abi.ptr(start, C.c_uint32), mem, abi.ptr(xy, D), len(fk),
                                    abi.ptr(fk, C.c_uint32), abi.ptr(np.ascontiguousarray(fxy), D),
                                    abi.ptr(np.ascontiguousarray(fq), D), 4, res, abi.ptr(ug, D), abi.ptr(fg, D))
    assert n >= 0, abi.STATUS[-n]
    return list(res[:n]), ug[:n * U * 2].reshape(n, U, 2), fg


def test_carry_macro_and_register_share_one_origin(oracle_lib):
    # adjusted_carry_macro_and_register_share_one_origin_across_density_fields
    # (electrostatic_placement.rs:1534-1601)
    W, H = 8, 1
    cap = {LUT4: np.ones((H, W), np.uint32), FF: np.ones((H, W), np.uint32)}

    def make(x):
        members = [(LUT4, float(o), 0.0, 1.0 + 0.1 * o) for o in range(4)] + [(FF, 2.0, 0.0, 1.5)]
        return fixture(W, H, cap, [(x, 0.0, members)])

    fx = make(2.0)
    fl = fixture_fillers(oracle_lib, W, H, fx)
    res, ug, _ = fixture_evaluate(oracle_lib, W, H, fx, [2.0, 0.0], fl)
    assert [(r.kind, r.real_charge) for r in res] == [(LUT4, 4), (FF, 1)]
    assert close(res[0].filler_charge, 3.4, 1e-14) and close(res[1].filler_charge, 6.5, 1e-14)

    def total(x):
        r, _, _ = fixture_evaluate(oracle_lib, W, H, make(x), [x, 0.0], fixture_fillers(oracle_lib, W, H, make(x)))
        return sum(f.energy for f in r)

    eps = 1e-6
    fd = (total(2.0 + eps) - total(2.0 - eps)) / (2 * eps)
    assert close(ug[:, 0, 0].sum(), fd, 2e-8)


def test_external_positions_match_materialised_macros(oracle_lib):
    # external_optimizer_positions_exactly_match_materialized_rigid_macros
    # (electrostatic_placement.rs:1603-1662): the session path (optimizer
    # coordinates, member offsets) equals the fixture path with the unit
    # origins and filler positions materialised, bit for bit.
    W, H, cap = problems.mixed_device_capacity()
    p = problems.density_problem(W, H, cap, [(0.25, 0.5, [(LUT4, 0.0, 0.0), (FF, 0.5, 0.25)])])
    s = abi.Session(oracle_lib, "oracle", p)
    x = s.initial.copy()
    x[0:2] = (1.25, 0.75)
    for j in range(s.n_fillers):
        x[2 + 2 * j] += 0.01 * (j % 3)
        x[3 + 2 * j] += 0.02 * (j % 2)
    e = s.evaluate(x, 1.0)
    fx = fixture(W, H, cap, [(0.25, 0.5, [(LUT4, 0.0, 0.0, 1.0), (FF, 0.5, 0.25, 1.0)])])
    fk, _, fq = fixture_fillers(oracle_lib, W, H, fx)
    res, ug, fg = fixture_evaluate(oracle_lib, W, H, fx, x[:2], (fk, x[2:].copy(), fq))
    assert [r.energy for r in res] == list(e["energy"])
    assert [r.overflow for r in res] == list(e["overflow"])
    assert np.array_equal(ug[:, 0, :], e["unit_gradient"][:, 0, :])
    assert np.array_equal(fg.reshape(-1, 2), e["filler_gradient"])
    s.close()


def test_adjusted_area_consumes_only_filler_charge(oracle_lib):
    # adjusted_physical_area_consumes_only_filler_charge (electrostatic_placement.rs:1726-1749)
    W, H, cap = problems.mixed_device_capacity()
    cap = {LUT4: cap[LUT4]}
    fx = fixture(W, H, cap, [(1.25, 0.0, [(LUT4, 0.0, 0.0, 1.75)])])
    fl = fixture_fillers(oracle_lib, W, H, fx)
    assert close(fl[2].sum(), 2.25, 1e-15)
    res, _, _ = fixture_evaluate(oracle_lib, W, H, fx, [1.25, 0.0], fl)
    assert res[0].real_area == 1.75 and res[0].filler_charge == 2.25
    assert close(res[0].net_charge, 0.0, 1e-14)


def test_filler_totals_are_checked(oracle_lib):
    # validate_filler_totals (electrostatic_placement.rs:506-539): fillers that
    # do not add up to the free capacity are a typed error
    W, H, cap = problems.mixed_device_capacity()
    cap = {LUT4: cap[LUT4]}
    fx = fixture(W, H, cap, [(1.25, 0.0, [(LUT4, 0.0, 0.0, 1.75)])])
    fk, fxy, fq = fixture_fillers(oracle_lib, W, H, fx)
    fq = fq.copy()
    fq[0] += 0.5
    with pytest.raises(AssertionError, match="FillerChargeMismatch"):
        fixture_evaluate(oracle_lib, W, H, fx, [1.25, 0.0], (fk, fxy, fq))


# ---------------------------------------------------- global_placement.rs --
def step(lib, x, xp, momentum, lipschitz, fn, lower=(-10.0,), upper=(10.0,), metric=(1.0,)):
    n = len(x)
    xa, xpa = (D * n)(*x), (D * n)(*xp)
    st = oracle.NesterovState(xa, xpa, momentum, lipschitz or 0.0, 1 if lipschitz else 0)
    res = oracle.NesterovResult()
    cb = oracle.EVALUATE_FN(fn)
    rc = lib.oracle_nesterov_step_lengths(C.byref(st), n, len(lower), arr(D, lower), len(upper), arr(D, upper),
                                          len(metric), arr(D, metric), cb, None, C.byref(res))
    return rc, st, res, list(xa), list(xpa)


def quadratic(target, log=None):
    def f(_c, x, n, v, g):
        if log is not None:
            log.append(x[0])
        d = x[0] - target
        v[0] = 0.5 * d * d
        g[0] = d
        return 0
    return f


def test_changed_objective_keeps_momentum(oracle_lib):
    # dynamic_step_re_evaluates_a_changed_objective_and_keeps_momentum (global_placement.rs:849-873)
    rc, st, res, x, xp = step(oracle_lib, [2.0], [2.0], 1.0, None, quadratic(0.0))
    assert rc == 0
    m1 = st.momentum
    old = x[0]
    old_obj = 0.5 * (old - 3.0) ** 2
    seen = []
    rc, st2, res2, x2, _ = step(oracle_lib, x, xp, st.momentum, st.lipschitz, quadratic(3.0, seen))
    assert rc == 0 and seen[0] == old and res2.accepted == 1
    assert res2.objective <= old_obj and st2.momentum > m1


def counting(calls):
    def f(_c, x, n, v, g):
        calls.append(1)
        v[0] = 0.0
        g[0] = 0.0
        return 0
    return f


def test_malformed_bounds_and_nonfinite_coordinates_are_typed(oracle_lib):
    # malformed_bounds_and_nonfinite_initial_coordinates_are_typed_errors (global_placement.rs:970-1011)
    calls = []
    rc, st, _, x, _ = step(oracle_lib, [0.5], [0.5], 1.0, None, counting(calls), lower=(), upper=(1.0,))
    assert abi.STATUS[rc] == "InvalidBounds" and not calls and x == [0.5] and st.momentum == 1.0
    rc, _, res, x, _ = step(oracle_lib, [float("nan")], [float("nan")], 1.0, None, counting(calls), lower=(0.0,),
                            upper=(1.0,))
    assert abi.STATUS[rc] == "InvalidInitialCoordinate" and res.error_index == 0 and np.isnan(x[0])


def test_invalid_metrics_are_rejected_before_evaluation(oracle_lib):
    # invalid_diagonal_metrics_are_rejected_before_evaluation (global_placement.rs:1013-1054)
    calls = []
    rc, st, res, x, _ = step(oracle_lib, [1.0], [1.0], 1.0, None, counting(calls), metric=())
    assert abi.STATUS[rc] == "InvalidMetricLength" and res.error_index == 1 and not calls and x == [1.0]
    for m in (0.0, -1.0, float("nan"), float("inf")):
        rc, _, res, x, _ = step(oracle_lib, [1.0], [1.0], 1.0, None, counting(calls), metric=(m,))
        assert abi.STATUS[rc] == "NonPositiveMetric" and res.error_index == 0 and not calls and x == [1.0]


def test_invalid_evaluation_is_transactional(oracle_lib):
    # invalid_dynamic_evaluation_is_transactional (global_placement.rs:1056-1072)
    def nan(_c, x, n, v, g):
        v[0] = float("nan")
        g[0] = 0.0
        return 0
    rc, st, _, x, xp = step(oracle_lib, [2.0], [2.0], 1.0, None, nan)
    assert abi.STATUS[rc] == "InvalidObjective" and x == [2.0] and xp == [2.0] and st.momentum == 1.0

    def inf_grad(_c, x, n, v, g):
        v[0] = 0.0
        g[0] = float("inf")
        return 0
    rc, _, res, x, _ = step(oracle_lib, [2.0], [2.0], 1.0, None, inf_grad)
    assert abi.STATUS[rc] == "NonFiniteGradient" and res.error_index == 0 and x == [2.0]


def test_invalid_parameters_are_rejected_before_evaluation(oracle_lib):
    # invalid_dynamic_parameters_are_rejected_before_evaluation (global_placement.rs:1074-1116)
    calls = []
    rc, st, _, x, _ = step(oracle_lib, [2.0], [2.0], 0.5, None, counting(calls))
    assert abi.STATUS[rc] == "InvalidMomentumParameter" and not calls and st.momentum == 0.5
    rc, st, _, x, _ = step(oracle_lib, [2.0], [2.0], 1.0, float("inf"), counting(calls))
    assert abi.STATUS[rc] == "InvalidLipschitzEstimate" and not calls and x == [2.0]


def test_dynamic_steps_are_bit_deterministic(oracle_lib):
    # dynamic_steps_are_bit_deterministic (global_placement.rs:1118-1139)
    a = step(oracle_lib, [2.0], [2.0], 1.0, None, quadratic(-3.0))
    b = step(oracle_lib, [2.0], [2.0], 1.0, None, quadratic(-3.0))
    assert a[0] == b[0] == 0 and a[3] == b[3] and a[4] == b[4]
    assert (a[1].momentum, a[1].lipschitz) == (b[1].momentum, b[1].lipschitz)
    assert (a[2].objective, a[2].line_search_trials, a[2].evaluations) == \
        (b[2].objective, b[2].line_search_trials, b[2].evaluations)


def test_repeated_dynamic_steps_reach_an_interior_quadratic_minimum(oracle_lib):
    # the dynamic step driven to stationarity on an anisotropic quadratic
    # (the reference pins the same minimum for its monotone variant,
    # global_placement.rs:1141-1154)
    def f(_c, x, n, v, g):
        v[0] = 0.5 * ((x[0] - 2.0) ** 2 + 4.0 * (x[1] + 1.0) ** 2)
        g[0] = x[0] - 2.0
        g[1] = 4.0 * (x[1] + 1.0)
        return 0
    x, xp, mom, lip = [8.0, 7.0], [8.0, 7.0], 1.0, None
    for _ in range(200):
        rc, st, res, x, xp = step(oracle_lib, x, xp, mom, lip, f, lower=(-10.0, -10.0), upper=(10.0, 10.0),
                                  metric=(1.0, 1.0))
        assert rc == 0
        mom, lip = st.momentum, (st.lipschitz if st.has_lipschitz else None)
        if res.accepted == 0:
            break
    assert abs(x[0] - 2.0) <= 1e-6 and abs(x[1] + 1.0) <= 1e-6
