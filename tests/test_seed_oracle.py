"""The coarse quadratic seed's oracle (CPU): linearize_axis, the axis
equations' component gauges and the Jacobi-PCG solve, pinned to the
reference's own tests (analytical_placement.rs:1185-1390,
lib.rs:14240-14441)."""
import ctypes as C

import numpy as np
import pytest

from texo_amd import abi

D, U32, U64, U8 = C.c_double, C.c_uint32, C.c_uint64, C.c_uint8


def arr(ctype, values):
    return (ctype * max(len(values), 1))(*values)


def declare(lib):
    P = C.POINTER
    lib.oracle_linearize_axis.argtypes = [U32, U32, P(D), P(U8), P(D), P(D), U32, P(U32), P(U32), P(U32), U32,
                                          P(U32), P(U32), P(U64), U64, P(U32), P(U32), P(D), P(U8)]
    lib.oracle_linearize_axis.restype = U64
    lib.oracle_solve_axis.argtypes = [U32, P(D), P(U32), P(U32), P(D), P(D), P(U8), P(D), U32, P(D), P(U64)]
    lib.oracle_solve_axis.restype = C.c_int


@pytest.fixture(scope="module")
def lib(oracle_lib):
    declare(oracle_lib)
    return oracle_lib


def wide_net(fanout, critical=None):
    """wide_net (analytical_placement.rs:828-847): one unit per cell."""
    n = fanout + 1
    arcs = [(0, 0, 1, 1, critical - 1)] if critical else []
    return n, [0, n], list(range(n)), arcs


def linearize(lib, positions, fanout, critical=None, fixed=None, offsets=None, cell_unit=None):
    n, ns, pins, arcs = wide_net(fanout, critical)
    cells = n
    cu = cell_unit or list(range(cells))
    units = max(cu) + 1
    fm = fixed or [0] * cells
    cap = 4 * n + 8
    l, r, w, t = (U32 * cap)(), (U32 * cap)(), (D * cap)(), (U8 * cap)()
    k = lib.oracle_linearize_axis(units, cells, arr(D, positions), arr(U8, fm), arr(D, [0.0] * cells),
                                  arr(D, offsets or [0.0] * cells), 1, arr(U32, ns), arr(U32, [cu[p] for p in pins]),
                                  arr(U32, pins), len(arcs), arr(U32, [cu[v] for a in arcs for v in (a[1], a[3])]),
                                  arr(U32, [v for a in arcs for v in (a[1], a[3])]), arr(U64, [a[4] for a in arcs]),
                                  cap, l, r, w, t)
    return [(l[i], r[i], w[i], bool(t[i])) for i in range(k)]


def test_b2b_influence_is_continuous_across_fanout(lib):
    # b2b_influence_is_continuous_across_legacy_fanout_boundary (analytical_placement.rs:1185-1206)
    def total(f):
        return sum(e[2] for e in linearize(lib, [0.0] * (f + 1), f) if not e[3])
    a, b, c = total(256), total(257), total(402)
    assert a > 0 and b > 0 and c > 0
    assert abs(b / a - 1) < 0.01 and abs(c / a - 1) < 0.01


def test_timed_overlay_is_sink_local(lib):
    # timed_overlay_is_sink_local_and_not_diluted_by_fanout (analytical_placement.rs:1208-1232)
    def overlay(f):
        pos = [0.0] * (f + 1)
        pos[1] = 8.0
        return [e for e in linearize(lib, pos, f, critical=64) if e[3]]
    narrow, wide = overlay(2), overlay(257)
    assert len(narrow) == 1 and len(wide) == 1
    assert narrow[0][:2] == (0, 1) and wide[0][:2] == (0, 1)
    assert abs(narrow[0][2] - wide[0][2]) <= np.finfo(float).eps
    assert narrow[0][2] == 64.0 * 63 / 8.0


def test_full_u64_timing_weight(lib):
    # full_u64_timing_weight_linearizes_without_narrowing_or_panicking (analytical_placement.rs:1234-1246)
    e = [e for e in linearize(lib, [0.0, 1.0], 1, critical=2 ** 64) if e[3]]
    # extra weight u64::MAX (critical - 1)
    assert np.isfinite(e[0][2]) and e[0][2] > 64.0 * (2 ** 32 - 1)


def test_collapsed_bounds_are_deterministic_and_distinct(lib):
    # collapsed_bounds_are_deterministic_and_span_distinct_units (analytical_placement.rs:1248-1259)
    a = linearize(lib, [0.0] * 5, 4)
    b = linearize(lib, [0.0] * 5, 4)
    assert a == b and all(e[0] != e[1] for e in a)


def test_legal_spread_changes_bounds(lib):
    # legal_spread_changes_bounds_selected_from_a_collapsed_tie (analytical_placement.rs:1261-1280)
    def has(edges, l, r):
        return any((e[0], e[1]) in ((l, r), (r, l)) for e in edges)
    collapsed = linearize(lib, [0.0] * 4, 3)
    legal = linearize(lib, [5.0, 0.0, 10.0, 6.0], 3)
    assert has(collapsed, 0, 3) and not has(collapsed, 1, 2)
    assert has(legal, 1, 2) and not has(legal, 0, 3)


def test_same_unit_net_adds_no_edge(lib):
    # same_unit_geometry_is_scored_but_remains_outside_continuous_system (analytical_placement.rs:1355-1390)
    assert linearize(lib, [1.0], 1, critical=5, offsets=[0.0, 6.0], cell_unit=[0, 0]) == []


def solve(lib, diag, adj, rhs, anchored, initial, center):
    n = len(diag)
    start = np.cumsum([0] + [len(a) for a in adj]).astype(np.uint32)
    other = [o for a in adj for o, _ in a]
    weight = [w for a in adj for _, w in a]
    out = (D * n)()
    it = U64()
    rc = lib.oracle_solve_axis(n, arr(D, diag), arr(U32, start), arr(U32, other), arr(D, weight), arr(D, rhs),
                               arr(U8, anchored), arr(D, initial), center, out, C.byref(it))
    return rc, np.array(out[:n]), it.value


def test_component_gauge_centers_each_component(lib):
    # component_gauge_preserves_relative_optimum_and_centers_each_component (lib.rs:14261-14292)
    rc, x, _ = solve(lib, [1.0, 1.0, 0.0], [[(1, 1.0)], [(0, 1.0)], []], [4.0, -4.0, 0.0], [0, 0, 0], [10.0] * 3, 10)
    assert rc == 0
    assert abs(x[0] - 12) < 1e-10 and abs(x[1] - 8) < 1e-10 and abs(x[2] - 10) < 1e-10
    assert (x[0] - x[1] - 4.0) ** 2 < 1e-20


def test_fixed_component_needs_no_gauge(lib):
    # fixed_connected_component_uses_absolute_constraint_without_a_gauge (lib.rs:14294-14311)
    rc, x, _ = solve(lib, [1.0, 2.0], [[], []], [3.0, 10.0], [1, 1], [100.0] * 2, 100)
    assert rc == 0 and abs(x[0] - 3) < 1e-10 and abs(x[1] - 5) < 1e-10


def test_explicit_anchor_prevents_a_gauge(lib):
    # explicit_anchor_prevents_an_artificial_component_gauge (lib.rs:14313-14323):
    # floating pair plus add_anchor(1, 2.0, 7.0)
    rc, x, _ = solve(lib, [1.0, 3.0], [[(1, 1.0)], [(0, 1.0)]], [4.0, -4.0 + 14.0], [0, 1], [100.0] * 2, 100)
    assert rc == 0 and abs(x[0] - 11) < 1e-10 and abs(x[1] - 7) < 1e-10


def path_equations(n, reverse):
    diag = [0.0] * n
    adj = [[] for _ in range(n)]
    for l in range(n - 1):
        diag[l] += 1.0
        diag[l + 1] += 1.0
        adj[l].append((l + 1, 1.0))
        adj[l + 1].append((l, 1.0))
    if reverse:
        adj = [a[::-1] for a in adj]
    expected = [float(i) for i in range(n)]
    rhs = []
    for u in range(n):
        s = diag[u] * expected[u]
        for o, w in adj[u]:
            s = s - w * expected[o]
        rhs.append(s)
    return diag, adj, rhs


def test_long_floating_laplacian_is_order_deterministic(lib):
    # long_floating_laplacian_converges_and_is_insertion_order_deterministic (lib.rs:14358-14379)
    f = solve(lib, *path_equations(129, False), [0] * 129, [64.0] * 129, 64)
    r = solve(lib, *path_equations(129, True), [0] * 129, [64.0] * 129, 64)
    assert f[0] == 0 and r[0] == 0
    idx = np.arange(129.0)
    assert np.abs(f[1] - idx).max() < 1e-7 and np.abs(r[1] - idx).max() < 1e-7
    assert np.abs(f[1] - r[1]).max() < 1e-10


def test_jacobi_pcg_is_not_truncated(lib):
    # jacobi_pcg_is_not_truncated_at_the_legacy_hundred_iteration_budget (lib.rs:14381-14441)
    n = 128
    d = 10.0 ** (-12.0 + 24.0 * np.arange(n) / (n - 1))
    # the former raw CG with a 100-step budget leaves a material residual
    x, r = np.zeros(n), np.ones(n)
    p, rr0 = r.copy(), 128.0
    rr = rr0
    for _ in range(100):
        q = p * d
        pq = p @ q
        if pq <= np.finfo(float).eps:
            break
        a = rr / pq
        x += a * p
        r -= a * q
        nx = r @ r
        p = r + nx / rr * p
        rr = nx
    assert np.sqrt(rr / rr0) > 1e-4
    rc, sol, it = solve(lib, list(d), [[] for _ in range(n)], [1.0] * n, [1] * n, [0.0] * n, 0)
    assert rc == 0
    assert np.sqrt(np.sum((1.0 - d * sol) ** 2) / rr0) < 1e-12
