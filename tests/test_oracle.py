"""Pins the FP64 oracle to the reference's own known answers (CPU only).

Each test restates one test of the reference (file:line given) or checks a
closed form / independent library routine, so that a plausible mistake in the
restatement (dropped term, wrong sign or index, transposed operand) fails."""
import ctypes as C

import numpy as np
import pytest
import scipy.fft

import oracle
from texo_amd import abi
from tests import problems

D = C.c_double
P = C.POINTER


def dp(a):
    return abi.ptr(a, C.c_double)


def solve(lib, W, H, density, matrix=False):
    phi = np.zeros(W * H)
    e = C.c_double()
    fn = lib.oracle_poisson_solve_matrix if matrix else lib.oracle_poisson_solve
    rc = fn(W, H, dp(np.ascontiguousarray(density, float)), dp(phi), C.byref(e))
    assert rc == 0
    return phi, e.value


# ---------------------------------------------------------------- Poisson --
def test_constant_density_is_only_the_removed_dc_mode(oracle_lib):
    # electrostatic_placement.rs:1114-1120
    phi, e = solve(oracle_lib, 7, 5, np.full(35, 3.25))
    assert abs(e) <= 1e-24 and np.all(phi == 0.0)


def neumann_laplacian(W, H, v):
    v = v.reshape(H, W)
    r = np.zeros_like(v)
    for y in range(H):
        for x in range(W):
            d = 0.0
            for yy, xx in ((y, x - 1), (y, x + 1), (y - 1, x), (y + 1, x)):
                if 0 <= yy < H and 0 <= xx < W:
                    r[y, x] -= v[yy, xx]
                    d += 1.0
            r[y, x] += d * v[y, x]
    return r.reshape(-1)


@pytest.mark.parametrize("W,H", [(7, 5), (8, 4), (16, 16)])
def test_spectral_solution_matches_the_direct_neumann_operator(oracle_lib, W, H):
    # electrostatic_placement.rs:1122-1135 (7x5), also on FFT-path sizes
    i = np.arange(W * H)
    density = ((i * 17 + 3) % 23) - 11.0
    phi, _ = solve(oracle_lib, W, H, density)
    expected = density - density.mean()
    assert np.abs(neumann_laplacian(W, H, phi) - expected).max() <= 2e-12


@pytest.mark.parametrize("n", [11, 16])
def test_manufactured_cosine_is_one_laplacian_eigenmode(oracle_lib, n):
    # electrostatic_placement.rs:1137-1154
    x = np.arange(n)
    rho = np.cos(np.pi * 3.0 * (x + 0.5) / n)
    lam = 4.0 * np.sin(np.pi * 3.0 / (2.0 * n)) ** 2
    phi, _ = solve(oracle_lib, n, 1, rho)
    assert np.abs(phi - rho / lam).max() <= 2e-13


@pytest.mark.parametrize("W,H", [(127, 96), (128, 64), (64, 64)])
def test_fast_solver_matches_orthonormal_matrix_oracle(oracle_lib, W, H):
    # electrostatic_placement.rs:1239-1264
    i = np.arange(W * H)
    density = ((i * 31 + 7) % 101) - 50.0 + ((i * 13 + 5) % 17) / 17.0
    fast, ef = solve(oracle_lib, W, H, density)
    ref, er = solve(oracle_lib, W, H, density, matrix=True)
    scale = max(np.abs(ref).max(), 1.0)
    assert np.abs(fast - ref).max() <= 2e-11 * scale
    assert abs(ef - er) <= 2e-12 * max(abs(er), 1.0)
    # prime_width_architecture_grid_is_supported_and_repeatable (:1225-1237)
    again, e2 = solve(oracle_lib, W, H, density)
    assert np.array_equal(fast, again) and ef == e2 and ef > 0.0


@pytest.mark.parametrize("W,H", [(127, 96), (256, 256), (5, 9)])
def test_solver_matches_scipy_orthonormal_dct(oracle_lib, W, H):
    # independent textbook check: phi = IDCT2(DCT2(rho_bar) / lambda)
    rng = np.random.default_rng(W * H)
    rho = rng.normal(size=(H, W))
    phi, e = solve(oracle_lib, W, H, rho.reshape(-1))
    ky = np.arange(H)[:, None]
    kx = np.arange(W)[None, :]
    lam = 4 * np.sin(np.pi * kx / (2 * W)) ** 2 + 4 * np.sin(np.pi * ky / (2 * H)) ** 2
    inv = np.where(lam == 0, 0.0, 1.0 / np.where(lam == 0, 1.0, lam))
    centered = rho - rho.mean()
    ref = scipy.fft.idctn(scipy.fft.dctn(centered, type=2, norm="ortho") * inv, type=2, norm="ortho")
    assert np.abs(phi.reshape(H, W) - ref).max() <= 1e-11 * np.abs(ref).max()
    assert abs(e - 0.5 * np.sum(centered * ref)) <= 1e-11 * abs(e)


# ---------------------------------------------------------------- B-spline --
def axis(lib, L, c):
    idx = (C.c_uint32 * 3)()
    w = (C.c_double * 3)()
    d = (C.c_double * 3)()
    lib.oracle_bspline_axis(L, c, idx, w, d)
    return list(idx), list(w), list(d)


def test_quadratic_bspline_preserves_charge_and_has_zero_derivative_sum(oracle_lib):
    # electrostatic_placement.rs:1186-1201
    for c in (0.0, 0.25, 1.0, 1.5, 3.75, 5.0):
        _, w, d = axis(oracle_lib, 6, c)
        assert abs(sum(w) - 1.0) <= 1e-15 and abs(sum(d)) <= 1e-15
    g = np.zeros(30)
    oracle_lib.oracle_deposit_bspline(dp(g), 6, 5, 2.25, 3.5, 7.0)
    oracle_lib.oracle_deposit_bspline(dp(g), 6, 5, 5.0, 0.0, 11.0)
    assert abs(g.sum() - 18.0) <= 1e-14


def test_bspline_indices_follow_the_reference_rule(oracle_lib):
    # quadratic_bspline_axis (electrostatic_placement.rs:950-976): clamp to
    # [0, L-1], centre = min(floor(x + 1/2), L-1), boundary folds.
    rng = np.random.default_rng(0)
    for L in (1, 2, 6, 64):
        for c in np.concatenate([rng.uniform(-3, L + 2, 200), np.arange(-1, L + 1, 0.5)]):
            idx, w, _ = axis(oracle_lib, L, c)
            cc = min(max(c, 0.0), L - 1.0)
            centre = int(min(np.floor(cc + 0.5), L - 1))
            assert idx == [max(centre - 1, 0), centre, min(centre + 1, L - 1)]
            o = cc - centre
            assert w == [0.5 * ((0.5 - o) * (0.5 - o)), 0.75 - o * o, 0.5 * ((0.5 + o) * (0.5 + o))]


def test_bspline_energy_gradient_is_smooth_at_stencil_knots(oracle_lib):
    # electrostatic_placement.rs:1203-1223
    W, H = 13, 9

    def energy(x, y):
        g = np.full(W * H, -2.0 / 117.0)
        oracle_lib.oracle_deposit_bspline(dp(g), W, H, x, y, 2.0)
        return solve(oracle_lib, W, H, g)[1]

    for x, y in ((4.0, 5.0), (4.5, 5.5)):
        g = np.full(W * H, -2.0 / 117.0)
        oracle_lib.oracle_deposit_bspline(dp(g), W, H, x, y, 2.0)
        phi, _ = solve(oracle_lib, W, H, g)
        out = (C.c_double * 3)()
        oracle_lib.oracle_sample_bspline(dp(phi), W, H, x, y, out)
        eps = 1e-6
        fx = (energy(x + eps, y) - energy(x - eps, y)) / (2 * eps)
        fy = (energy(x, y + eps) - energy(x, y - eps)) / (2 * eps)
        assert abs(2 * out[1] - fx) <= 2e-8 and abs(2 * out[2] - fy) <= 2e-8


# ------------------------------------------------------------ density model --
def osession(lib, problem):
    return abi.Session(lib, "oracle", problem)


def test_fixed_occupancy_is_subtracted_and_exhaustion_is_typed(oracle_lib):
    # electrostatic_placement.rs:1316-1347
    W, H, cap = problems.mixed_device_capacity()
    unit = [(1.0, 0.0, [(problems.LUT4, 0.0, 0.0)])]
    ok = problems.density_problem(W, H, cap, unit, fixed_occupancy=[(problems.LUT4, 0, 0)])
    s = osession(oracle_lib, ok)
    # LUT capacity 3 after subtraction (one filler per nonzero bin), demand 1
    assert s.n_fillers == 3 and abs(s.filler_charge[:3].sum() - 2.0) <= 1e-15
    bad = problems.density_problem(W, H, cap, unit, fixed_occupancy=[(problems.LUT4, 0, 0)] * 2)
    with pytest.raises(abi.PlacementError) as err:
        osession(oracle_lib, bad)
    assert err.value.name == "FixedCapacityExhausted"
    assert err.value.kind == problems.LUT4 and err.value.point == (0, 0)


def test_lut_and_register_fields_are_separate(oracle_lib):
    # electrostatic_placement.rs:1349-1388
    W, H, cap = problems.mixed_device_capacity()
    p = problems.density_problem(W, H, cap, [(1.25, 0.25, [(problems.LUT4, 0.0, 0.0), (problems.FF, 0.5, 0.5)])])
    s = osession(oracle_lib, p)
    e1 = s.evaluate(s.initial, 1.0)
    e2 = s.evaluate(s.initial, 1.0)
    assert list(s.field_kind) == [problems.LUT4, problems.FF]
    assert e1["filler_charge"][0] == 3.0 and e1["filler_charge"][1] == 3.0
    assert np.array_equal(e1["energy"], e2["energy"])
    assert np.all(np.isfinite(e1["energy"]))


def test_model_gradient_matches_total_energy_finite_difference(oracle_lib):
    # electrostatic_placement.rs:1390-1428
    W, H, cap = problems.mixed_device_capacity()

    def make(x):
        return problems.density_problem(W, H, cap, [(x, 0.4, [(problems.LUT4, 0.0, 0.0), (problems.FF, 0.3, 0.4)])])

    def total(x):
        s = osession(oracle_lib, make(x))
        return s.evaluate(s.initial, 1.0)["energy"].sum()

    x = 1.2
    s = osession(oracle_lib, make(x))
    e = s.evaluate(s.initial, 1.0)
    grad = e["unit_gradient"][:, 0, 0].sum()
    eps = 1e-6
    assert abs(grad - (total(x + eps) - total(x - eps)) / (2 * eps)) <= 2e-8


def test_capacity_background_and_deposit_preserve_charge(oracle_lib):
    # electrostatic_placement.rs:1430-1452
    W, H, cap = problems.mixed_device_capacity()
    cap = {problems.LUT4: cap[problems.LUT4]}
    p = problems.density_problem(W, H, cap, [(x + 0.25, 0.5, [(problems.LUT4, 0.0, 0.0)]) for x in range(3)])
    s = osession(oracle_lib, p)
    e = s.evaluate(s.initial, 1.0, want_grids=True)
    assert abs(e["density"].sum()) <= 1e-14 and abs(e["net_charge"][0]) <= 1e-14
    assert e["filler_charge"][0] == 1.0 and e["overflow"][0] <= 1.0


def test_overflow_is_independent_of_filler_coordinates(oracle_lib):
    # electrostatic_placement.rs:1454-1484
    W, H, cap = problems.mixed_device_capacity()
    cap = {problems.LUT4: cap[problems.LUT4]}
    s = osession(oracle_lib, problems.density_problem(W, H, cap, [(0.0, 0.0, [(problems.LUT4, 0.0, 0.0)])]))
    base = s.evaluate(s.initial, 1.0)
    x = s.initial.copy()
    U = 1
    for j in range(s.n_fillers):
        x[2 * (U + j)] = 0.0 if j % 2 == 0 else 3.0
        x[2 * (U + j) + 1] = 2.0
    moved = s.evaluate(x, 1.0)
    assert base["overflow"][0] == moved["overflow"][0]
    assert base["energy"][0] != moved["energy"][0]


def test_separating_overlapping_cells_reduces_overflow(oracle_lib):
    # electrostatic_placement.rs:1486-1509
    W, H, cap = problems.mixed_device_capacity()
    cap = {problems.LUT4: cap[problems.LUT4]}
    p = problems.density_problem(W, H, cap, [(0.0, 0.0, [(problems.LUT4, 0.0, 0.0)])] * 2)
    s = osession(oracle_lib, p)
    over = s.evaluate(s.initial, 1.0)["overflow"][0]
    x = s.initial.copy()
    x[2] = 1.0
    sep = s.evaluate(x, 1.0)["overflow"][0]
    assert sep < over


def test_four_lut_macro_is_one_rigid_variable_and_cancels_capacity(oracle_lib):
    # electrostatic_placement.rs:1511-1532 and :1664-1684
    W, H, cap = problems.mixed_device_capacity()
    cap = {problems.LUT4: cap[problems.LUT4]}
    members = [(problems.LUT4, float(o), 0.0) for o in range(4)]
    s = osession(oracle_lib, problems.density_problem(W, H, cap, [(0.25, 0.0, members)]))
    e = s.evaluate(s.initial, 1.0, want_grids=True)
    assert s.n_density_units == 1 and s.n_fillers == 0 and e["filler_charge"][0] == 0.0
    assert abs(e["density"].sum()) <= 1e-14
    s0 = osession(oracle_lib, problems.density_problem(W, H, cap, [(0.0, 0.0, members)]))
    e0 = s0.evaluate(s0.initial, 1.0, want_grids=True)
    assert np.all(e0["density"] == 0.0) and e0["energy"][0] == 0.0 and e0["overflow"][0] == 0.0


def test_initial_fillers_are_capacity_proportional(oracle_lib):
    # electrostatic_placement.rs:1686-1724
    W, H, cap = problems.mixed_device_capacity()
    p = problems.density_problem(W, H, cap, [(1.0, 0.5, [(problems.LUT4, 0.0, 0.0), (problems.FF, 0.0, 0.0)])])
    s = osession(oracle_lib, p)
    assert s.n_fillers == 8
    for f in (0, 1):
        q = s.filler_charge[s.filler_field == f]
        assert q.size == 4 and np.all(np.abs(q - 0.75) <= np.finfo(float).eps) and q.sum() == 3.0
    assert np.all(s.evaluate(s.initial, 1.0)["net_charge"] == 0.0)


def test_filler_gradient_matches_energy_finite_difference(oracle_lib):
    # electrostatic_placement.rs:1751-1786
    W, H, cap = problems.mixed_device_capacity()
    cap = {problems.LUT4: cap[problems.LUT4]}
    s = osession(oracle_lib, problems.density_problem(W, H, cap, [(2.25, 0.5, [(problems.LUT4, 0.0, 0.0)])]))
    U = 1
    j = next(j for j in range(s.n_fillers) if s.initial[2 * (U + j)] == 1.0)
    e = s.evaluate(s.initial, 1.0)

    def energy(x):
        c = s.initial.copy()
        c[2 * (U + j)] = x
        return s.evaluate(c, 1.0)["energy"][0]

    eps = 1e-6
    assert abs(e["filler_gradient"][j, 0] - (energy(1 + eps) - energy(1 - eps)) / (2 * eps)) <= 2e-8


def test_capacity_shortage_is_a_typed_error(oracle_lib):
    # electrostatic_placement.rs:1788-1843
    W, H, cap = problems.mixed_device_capacity()
    p = problems.density_problem(W, H, cap, [(0.0, 0.0, [(abi.KIND_MEMORY, 0.0, 0.0)])])
    with pytest.raises(abi.PlacementError) as err:
        osession(oracle_lib, p)
    assert err.value.name == "InsufficientCapacity" and err.value.kind == abi.KIND_MEMORY
    assert err.value.values == (1.0, 0.0)


# ------------------------------------------------------------ wirelength --
def wide_net(fanout, critical=None):
    """wide_net (analytical_placement.rs:828-847): source + fanout sinks,
    one unit per cell, optional timing weight on the first sink."""
    n = fanout + 1
    net_start = np.array([0, n], np.uint32)
    pins = np.arange(n, dtype=np.uint32)
    arcs = np.array([0, 1], np.uint32) if critical else np.zeros(0, np.uint32)
    extra = np.array([critical - 1], np.uint64) if critical else np.zeros(0, np.uint64)
    return n, net_start, pins, arcs, extra


def wa(lib, positions, fanout, critical=None, gamma=0.25, fixed=None, fixed_xy=None, offsets=None, cells=None):
    n, ns, pins, arcs, extra = wide_net(fanout, critical)
    units = len(positions)
    cells = cells if cells is not None else n
    pos = np.ascontiguousarray(np.array(positions, float).reshape(-1))
    fm = np.zeros(cells, np.uint8) if fixed is None else np.asarray(fixed, np.uint8)
    fx = np.zeros(2 * cells) if fixed_xy is None else np.asarray(fixed_xy, float).reshape(-1)
    of = np.zeros(2 * cells) if offsets is None else np.asarray(offsets, float).reshape(-1)
    v = C.c_double()
    gx, gy = np.zeros(units), np.zeros(units)
    u32 = lambda a: abi.ptr(np.ascontiguousarray(a, np.uint32), C.c_uint32)
    lib.oracle_weighted_average(units, cells, dp(pos), abi.ptr(fm, C.c_uint8), dp(fx), dp(of), 1, u32(ns), u32(pins),
                                u32(pins), len(extra), u32(arcs), u32(arcs), abi.ptr(extra, C.c_uint64), gamma,
                                C.byref(v), dp(gx), dp(gy))
    return v.value, gx, gy


def test_two_pin_weighted_average_matches_closed_form(oracle_lib):
    # analytical_placement.rs:884-900
    v, gx, gy = wa(oracle_lib, [(0.0, 1.0), (4.0, -2.0)], 1)
    g = 0.25
    expected = 4.0 * np.tanh(4.0 / (2 * g)) + 3.0 * np.tanh(3.0 / (2 * g))
    assert abs(v - expected) <= 1e-14 * max(expected, 1.0)
    assert abs(gx.sum()) <= 1e-14 and abs(gy.sum()) <= 1e-14


def test_weighted_average_is_translation_invariant_and_deterministic(oracle_lib):
    # analytical_placement.rs:902-936
    pos = [(-3.0, 7.0), (11.0, -2.0), (1.0, 5.0), (4.0, 15.0), (8.0, 0.0)]
    a = wa(oracle_lib, pos, 4)
    b = wa(oracle_lib, pos, 4)
    c = wa(oracle_lib, [(x + 1e6, y - 2e6) for x, y in pos], 4)
    for r in (b, c):
        assert a[0] == r[0] and np.array_equal(a[1], r[1]) and np.array_equal(a[2], r[2])


def test_weighted_average_gradient_matches_finite_differences(oracle_lib):
    # analytical_placement.rs:938-988 (timed arc of weight 7)
    pos = [(0.00, 0.18), (0.08, 0.02), (0.15, 0.11), (0.22, 0.27)]
    v, gx, gy = wa(oracle_lib, pos, 3, critical=7)
    h = 1e-6
    for u in range(4):
        for axis_, g in ((0, gx), (1, gy)):
            lo = [list(p) for p in pos]
            hi = [list(p) for p in pos]
            lo[u][axis_] -= h
            hi[u][axis_] += h
            fd = (wa(oracle_lib, hi, 3, critical=7)[0] - wa(oracle_lib, lo, 3, critical=7)[0]) / (2 * h)
            assert abs(g[u] - fd) <= 2e-8 * max(abs(fd), 1.0)


def test_timing_overlay_is_a_weighted_two_pin_term(oracle_lib):
    # analytical_placement.rs:1106-1183: overlay = extra weight x 2-pin WA
    pos = [(0.0, 0.0), (3.0, 1.0), (5.0, 2.0), (1.0, 4.0)]
    base = wa(oracle_lib, pos, 3)
    timed = wa(oracle_lib, pos, 3, critical=5)
    two = wa(oracle_lib, pos[:2], 1)
    assert abs((timed[0] - base[0]) - 4 * two[0]) <= 1e-12 * timed[0]


def test_wirelength_preconditioner_counts_nets_and_timing_weight(oracle_lib):
    # analytical_placement.rs:857-882
    n, ns, pins, arcs, extra = wide_net(3, 7)
    d = np.zeros(n)
    b = np.zeros(n)
    u32 = lambda a: abi.ptr(np.ascontiguousarray(a, np.uint32), C.c_uint32)
    oracle_lib.oracle_wirelength_preconditioner(n, 1, u32(ns), u32(pins), 1, u32(arcs), abi.ptr(extra, C.c_uint64),
                                                dp(d), dp(b))
    assert np.all(b == 1.0 / 3.0)
    assert d[0] == 6.0 + 1.0 / 3.0 and d[1] == 6.0 + 1.0 / 3.0 and d[2] == 1.0 / 3.0 and d[3] == 1.0 / 3.0


# -------------------------------------------------------------- Nesterov --
class Quadratic1D:
    def __init__(self, target, fail_at=None, values=None):
        self.target, self.fail_at, self.values, self.calls = target, fail_at, values, 0

    def __call__(self, _ctx, x, n, value, grad):
        self.calls += 1
        if self.fail_at is not None and self.calls == self.fail_at:
            return 7
        if self.values is not None:
            value[0] = self.values(self.calls, x[0])
            grad[0] = 1.0
            return 0
        d = x[0] - self.target
        value[0] = 0.5 * d * d
        grad[0] = d
        return 0


def nesterov(lib, x, xp, momentum, lipschitz, lower, upper, metric, fn):
    n = len(x)
    xa = (C.c_double * n)(*x)
    xpa = (C.c_double * n)(*xp)
    st = oracle.NesterovState(xa, xpa, momentum, lipschitz or 0.0, 1 if lipschitz else 0)
    res = oracle.NesterovResult()
    cb = oracle.EVALUATE_FN(fn)
    arr = lambda v: (C.c_double * n)(*v)
    rc = lib.oracle_nesterov_step(C.byref(st), n, arr(lower), arr(upper), arr(metric), cb, None, C.byref(res))
    return rc, st, res, list(xa), list(xpa)


lib_sig_done = False


@pytest.fixture
def nlib(oracle_lib):
    oracle_lib.oracle_nesterov_step.argtypes = [C.POINTER(oracle.NesterovState), C.c_uint64, P(D), P(D), P(D),
                                                oracle.EVALUATE_FN, C.c_void_p, C.POINTER(oracle.NesterovResult)]
    return oracle_lib


def test_dynamic_step_returns_the_accepted_candidate(nlib):
    # global_placement.rs:746-764
    rc, st, res, x, _ = nesterov(nlib, [2.0], [2.0], 1.0, None, [-10.0], [10.0], [1.0], Quadratic1D(0.0))
    assert rc == 0 and res.accepted == 1 and x == [0.0]
    assert res.objective == 0.0 and res.stationarity == 0.0 and res.coordinate_change == 2.0


def test_diagonal_metric_conditions_an_anisotropic_quadratic(nlib):
    # global_placement.rs:766-796
    def f(_c, x, n, v, g):
        v[0] = 0.5 * x[0] * x[0] + 50.0 * x[1] * x[1]
        g[0] = x[0]
        g[1] = 100.0 * x[1]
        return 0
    rc, st, res, x, _ = nesterov(nlib, [2.0, 2.0], [2.0, 2.0], 1.0, 2.0, [-10.0] * 2, [10.0] * 2, [1.0, 100.0], f)
    assert rc == 0 and res.accepted == 1 and x == [0.0, 0.0] and res.line_search_trials == 1


def test_metric_kkt_residual_scales_and_keeps_bound_signs(oracle_lib):
    # global_placement.rs:798-809
    arr = lambda v: (C.c_double * 4)(*v)
    oracle_lib.oracle_box_metric_kkt_residual.argtypes = [C.c_uint64, P(D), P(D), P(D), P(D), P(D)]
    r = oracle_lib.oracle_box_metric_kkt_residual(4, arr([3.0, -40.0, -50.0, 6.0]), arr([0.0, 0.0, 2.0, 2.0]),
                                                  arr([0.0] * 4), arr([2.0] * 4), arr([9.0, 100.0, 25.0, 4.0]))
    assert r == 4.0


def test_unchanged_candidate_reduces_lipschitz_first(nlib):
    # global_placement.rs:811-828
    rc, st, res, x, _ = nesterov(nlib, [1.0], [1.0], 1.0, 2.0 ** 60, [-10.0], [10.0], [1.0], Quadratic1D(0.0))
    assert rc == 0 and res.accepted == 1 and res.coordinate_change > 0 and res.line_search_trials > 1


def test_overflowing_denominator_is_transactional(nlib):
    # global_placement.rs:830-847
    rc, st, res, x, xp = nesterov(nlib, [1.0], [1.0], 1.0, np.finfo(float).max, [-10.0], [10.0], [4.0],
                                  Quadratic1D(0.0))
    assert abi.STATUS[rc] == "LineSearchOverflow" and x == [1.0] and xp == [1.0]
    assert st.lipschitz == np.finfo(float).max and st.momentum == 1.0


def test_evaluator_failure_leaves_state_bit_identical(nlib):
    # global_placement.rs:875-901
    rc, st, res, x, xp = nesterov(nlib, [2.0], [2.0], 1.0, None, [-10.0], [10.0], [1.0], Quadratic1D(0.0))
    f = Quadratic1D(3.0, fail_at=2)
    rc2, st2, res2, x2, xp2 = nesterov(nlib, x, xp, st.momentum, st.lipschitz, [-10.0], [10.0], [1.0], f)
    assert abi.STATUS[rc2] == "Evaluation" and f.calls == 2
    assert x2 == x and xp2 == xp and st2.momentum == st.momentum and st2.lipschitz == st.lipschitz


def test_stationary_status_carries_the_origin(nlib):
    # global_placement.rs:903-928
    rc, st, res, x, xp = nesterov(nlib, [1.0], [1.0], 1.0, None, [-10.0], [10.0], [1.0], Quadratic1D(1.0))
    assert rc == 0 and res.accepted == 0 and res.payload_evaluation == 1 and res.line_search_trials == 0
    rc, st, res, x, xp = nesterov(nlib, x, xp, st.momentum, st.lipschitz if st.has_lipschitz else None,
                                  [-10.0], [10.0], [1.0], Quadratic1D(2.0))
    assert res.accepted == 1 and x == [2.0]


def test_restart_backtracks_instead_of_repeating_a_rejected_candidate(nlib):
    # global_placement.rs:930-968
    def values(call, x):
        return {1: 0.0, 2: 100.0, 3: 99.98, 4: 1.0, 5: 1.0}.get(call, x)
    f = Quadratic1D(0.0, values=values)
    rc, st, res, x, _ = nesterov(nlib, [0.0], [-10.0], 2.0, 100.0, [-20.0], [20.0], [1.0], f)
    assert rc == 0 and res.accepted == 1 and f.calls == 6 and res.payload_evaluation == 6
    assert res.line_search_trials == 4 and x == [-0.01]


# -------------------------------------------------------------- schedule --
def test_augmented_density_value_and_coefficient(oracle_lib):
    # eplace.rs:1667-1685
    v, c = C.c_double(), C.c_double()
    oracle_lib.oracle_augmented_density(12.5, 50.0, 0.75, C.byref(v), C.byref(c))
    assert abs(v.value - 0.75 * (12.5 + 0.5 * 2000 * 12.5 * 12.5 / 50.0)) <= 1e-12
    assert abs(c.value - 0.75 * (1 + 2000 * 0.25)) <= 1e-12


def test_continuous_rudy_conserves_directional_hpwl(oracle_lib):
    # eplace.rs:1687-1696
    boxes = np.array([1.25, 4.75, 2.50, 5.00, 0.0, 0.0, 0.0, 7.0])
    h, v = np.zeros(100), np.zeros(100)
    assert oracle_lib.oracle_continuous_routing_demand(10, 10, 2, dp(boxes), dp(h), dp(v)) == 0
    assert abs(h.sum() - 3.5) <= 1e-12 and abs(v.sum() - 9.5) <= 1e-12 and h.min() >= 0 and v.min() >= 0


def test_multiplier_growth_endpoints(oracle_lib):
    # eplace.rs:1698-1723
    assert oracle_lib.oracle_multiplier_growth(0.0) == 1.05
    assert oracle_lib.oracle_multiplier_growth(np.finfo(float).max) == 1.06
    assert abs(oracle_lib.oracle_multiplier_growth(1.0) - 1.055) <= 1e-15
    assert np.isnan(oracle_lib.oracle_multiplier_growth(-1.0))


def test_field_gamma_endpoints_and_weighted_mean(oracle_lib):
    # eplace.rs:1725-1749
    for t in (0.10, 0.20):
        assert abs(oracle_lib.oracle_field_wirelength_gamma(t, t, 8.0) - 0.8) <= 1e-14
        assert abs(oracle_lib.oracle_field_wirelength_gamma(1.0, t, 8.0) - 80.0) <= 1e-12
    arr = lambda v: (C.c_double * len(v))(*v)
    assert oracle_lib.oracle_weighted_gamma_mean(2, arr([2.0, 10.0]), arr([1.0, 3.0]), 8.0) == 8.0
    assert oracle_lib.oracle_weighted_gamma_mean(1, arr([80.0]), arr([0.0]), 8.0) == 8.0


def test_overflow_targets_and_rounding(oracle_lib):
    # eplace.rs:1300-1308, lib.rs:4566-4568
    assert oracle_lib.oracle_density_overflow_target(problems.LUT4) == 0.10
    assert oracle_lib.oracle_density_overflow_target(abi.KIND_REGISTER) == 0.10
    assert oracle_lib.oracle_density_overflow_target(abi.KIND_MEMORY) == 0.20
    assert oracle_lib.oracle_density_overflow_target(abi.KIND_DSP) == 0.20
    assert oracle_lib.oracle_density_overflow_target(abi.KIND_IO) == 0.0
    rc = oracle_lib.oracle_rounded_coordinate
    assert [rc(v, 10) for v in (-3.0, 0.49, 0.5, 1.5, 2.5, 8.6, 9.5, 42.0)] == [0, 0, 1, 2, 3, 9, 9, 9]


def test_routability_area_inflation(oracle_lib):
    # eplace.rs:1778-1797, instance_area.rs:64-84
    s2 = np.sqrt(2.0)
    assert abs(oracle_lib.oracle_routability_optimized_area(1.0, s2, 0.0) - 2.0) <= np.finfo(float).eps
    assert oracle_lib.oracle_routability_optimized_area(1.0, 3.0, 0.0) == 2.0
    assert oracle_lib.oracle_routability_optimized_area(1.0, 0.5, 0.25) == 0.25


def test_soft_division_ceiling_bits(oracle_lib):
    # register_clustering.rs:115-121
    sdc = oracle_lib.oracle_soft_division_ceiling
    assert sdc(2.0, 2.0) == 1.0 and sdc(2.5, 2.0) == 1.5 and sdc(3.5, 2.0) == 2.0 and sdc(8.0, 4.0) == 2.0


def clustering(lib, regs, controls, movable):
    n = len(regs)
    area = np.zeros(n)
    u32 = lambda v: abi.ptr(np.array(v, np.uint32), C.c_uint32)
    u64 = lambda v: abi.ptr(np.array(v, np.uint64), C.c_uint64)
    rc = lib.oracle_clustering_areas(n, u32([r[0] for r in regs]), dp(np.array([v for r in regs for v in r[1:]], float)),
                                     len(controls), u32([c[0] for c in controls]),
                                     u64([v for c in controls for v in (c[1], c[2])]), u64([c[3] for c in controls]),
                                     movable, dp(area))
    assert rc == 0
    return area


def test_single_register_clustering_area(oracle_lib):
    # register_clustering.rs:164-175 (erf cancels: a lone register gets 8.0)
    area = clustering(oracle_lib, [(7, 3.25, 4.5)], [(7, 1, 2, 3)], 2)
    assert abs(area[0] - 8.0) < 1e-12


def test_ce_partition_changes_charge_but_other_controls_do_not(oracle_lib):
    # register_clustering.rs:123-162
    regs = [(c, 10.0, 10.0) for c in range(4)]
    a = clustering(oracle_lib, regs, [(0, 1, 2, 3), (1, 1, 2, 3), (2, 1, 2, 4), (3, 9, 2, 3)], 4)
    b = clustering(oracle_lib, regs, [(0, 1, 2, 3), (1, 1, 2, 3), (2, 1, 2, 3), (3, 9, 2, 3)], 4)
    assert a[0] == a[1] and a[0] != b[0] and a[3] == b[3]


def test_fixed_control_records_are_excluded(oracle_lib):
    # register_clustering.rs:177-191
    a = clustering(oracle_lib, [(1, 2.0, 3.0)], [(1, 4, 5, 6), (99, 4, 5, 7)], 2)
    b = clustering(oracle_lib, [(1, 2.0, 3.0)], [(1, 4, 5, 6)], 2)
    assert np.array_equal(a, b)


def test_instance_area_adjustment_conserves_charge(oracle_lib):
    # instance_area.rs:427-662 (charge conservation, no deflation, scaling)
    M = oracle.OracleAreaMember
    members = (M * 3)(M(0, problems.LUT4, 1.0, 2.0, 0.0, 0.0), M(1, problems.LUT4, 1.0, 1.5, 0.0, 0.0),
                      M(2, problems.FF, 1.0, 0.5, 0.0, 0.0))
    fk = np.array([problems.LUT4, problems.LUT4, problems.FF], np.uint32)
    fa = np.array([0.6, 0.4, 2.0])
    ma, fo = np.zeros(3), np.zeros(3)
    reset = C.c_int()
    rc = oracle_lib.oracle_adjust_instance_areas(3, members, 3, abi.ptr(fk, C.c_uint32), dp(fa), dp(ma), dp(fo),
                                                 C.byref(reset))
    assert rc == 0 and reset.value == 1
    # LUT desire 1.5 > filler 1.0: scaled, fillers exhausted; FF: no growth
    assert abs(ma[0] + ma[1] + fo[0] + fo[1] - 3.0) <= 1e-14
    assert ma[2] == 1.0 and fo[2] == 2.0 and fo[0] == 0.0 and fo[1] == 0.0
    assert ma[0] >= 1.0 and ma[1] >= 1.0
