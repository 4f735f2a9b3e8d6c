"""The coarse quadratic seed on the GPU (texo_gp_quadratic_seed,
texo_gp_solve_axis) against the reference's known answers
(lib.rs:14261-14441) and the oracle's FP64 restatement.

The matrix-vector products are the reference's bit for bit (each row folded
in insertion order); the dot products are summed over blocks in a fixed
order, which differs from the oracle's sequential sums, so the solutions
agree within the solver's own tolerance (squared relative residual 1e-8,
lib.rs:4594), not to the bit: the test bounds are stated with each check."""
import ctypes as C
import time

import numpy as np
import pytest

from texo_amd import abi, seed, synth
from tests.test_seed_oracle import declare, path_equations, solve as oracle_solve

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def olib(oracle_lib):
    declare(oracle_lib)
    oracle_lib.oracle_quadratic_seed.argtypes = [C.POINTER(abi.texo_gp_problem), C.POINTER(C.c_double),
                                                 C.POINTER(abi.texo_gp_report)]
    oracle_lib.oracle_quadratic_seed.restype = C.c_int
    return oracle_lib


def test_component_gauge_centers_each_component():
    # component_gauge_preserves_relative_optimum_and_centers_each_component (lib.rs:14261-14292)
    x, _ = seed.solve_axis([1.0, 1.0, 0.0], [[(1, 1.0)], [(0, 1.0)], []], [4.0, -4.0, 0.0], [0, 0, 0], [10.0] * 3, 10)
    assert abs(x[0] - 12) < 1e-10 and abs(x[1] - 8) < 1e-10 and abs(x[2] - 10) < 1e-10
    assert (x[0] - x[1] - 4.0) ** 2 < 1e-20


def test_fixed_component_needs_no_gauge():
    # fixed_connected_component_uses_absolute_constraint_without_a_gauge (lib.rs:14294-14311)
    x, _ = seed.solve_axis([1.0, 2.0], [[], []], [3.0, 10.0], [1, 1], [100.0] * 2, 100)
    assert abs(x[0] - 3) < 1e-10 and abs(x[1] - 5) < 1e-10


def test_explicit_anchor_prevents_a_gauge():
    # explicit_anchor_prevents_an_artificial_component_gauge (lib.rs:14313-14323)
    x, _ = seed.solve_axis([1.0, 3.0], [[(1, 1.0)], [(0, 1.0)]], [4.0, 10.0], [0, 1], [100.0] * 2, 100)
    assert abs(x[0] - 11) < 1e-10 and abs(x[1] - 7) < 1e-10


def test_long_floating_laplacian_is_order_deterministic(olib):
    # long_floating_laplacian_converges_and_is_insertion_order_deterministic (lib.rs:14358-14379)
    idx = np.arange(129.0)
    got = []
    for reverse in (False, True):
        d, adj, rhs = path_equations(129, reverse)
        x, it = seed.solve_axis(d, adj, rhs, [0] * 129, [64.0] * 129, 64)
        rc, xo, ito = oracle_solve(olib, d, adj, rhs, [0] * 129, [64.0] * 129, 64)
        print(f"reverse={reverse}: gpu {it} iterations, oracle {ito}", flush=True)
        assert rc == 0 and np.abs(x - idx).max() < 1e-7
        # the same Krylov path: the oracle's iteration count, its solution
        # within the tolerance the reference asserts between orders
        assert it == ito and np.abs(x - xo).max() < 1e-10
        got.append(x)
    assert np.abs(got[0] - got[1]).max() < 1e-10


def test_jacobi_pcg_is_not_truncated():
    # jacobi_pcg_is_not_truncated_at_the_legacy_hundred_iteration_budget (lib.rs:14381-14441)
    n = 128
    d = 10.0 ** (-12.0 + 24.0 * np.arange(n) / (n - 1))
    x, _ = seed.solve_axis(d, [[] for _ in range(n)], [1.0] * n, [1] * n, [0.0] * n, 0)
    assert np.sqrt(np.sum((1.0 - d * x) ** 2) / n) < 1e-12


def test_breakdown_reports_did_not_converge_like_the_oracle(olib):
    # a negative diagonal makes p.Ap <= 0: solve_quadratic stops and
    # finish_quadratic_solve reports the residual (lib.rs:4618-4630, 4702-4752)
    args = ([-1.0, 2.0], [[], []], [3.0, 4.0], [1, 1], [0.0, 0.0], 0)
    with pytest.raises(abi.PlacementError) as err:
        seed.solve_axis(*args)
    rc, _, ito = oracle_solve(olib, *args)
    assert err.value.name == "DidNotConverge" and abi.STATUS[rc] == "DidNotConverge"
    assert err.value.index == ito


def random_equations(rng, n, degree, anchored_fraction):
    """A random connected-ish Laplacian plus anchors, rows in insertion order."""
    diag = np.zeros(n)
    adj = [[] for _ in range(n)]
    rhs = rng.normal(0, 50, n)
    for _ in range(n * degree // 2):
        a, b = rng.integers(0, n, 2)
        if a == b:
            continue
        w = float(rng.uniform(0.1, 64.0))
        diag[a] += w
        diag[b] += w
        adj[a].append((int(b), w))
        adj[b].append((int(a), w))
    anchored = (rng.random(n) < anchored_fraction).astype(np.uint8)
    diag += anchored * rng.uniform(0.5, 8.0, n)
    return list(diag), adj, list(rhs), list(anchored)


@pytest.mark.parametrize("n,degree,anchored", [(1, 0, 1.0), (300, 4, 0.05), (5000, 6, 0.01), (20000, 3, 0.0)])
def test_solve_axis_matches_the_oracle(olib, n, degree, anchored):
    rng = np.random.default_rng(n)
    d, adj, rhs, anc = random_equations(rng, n, degree, anchored)
    init = list(rng.uniform(0, 100, n))
    x, it = seed.solve_axis(d, adj, rhs, anc, init, 50)
    rc, xo, ito = oracle_solve(olib, d, adj, rhs, anc, init, 50)
    err = np.abs(x - xo).max() if n else 0.0
    print(f"n={n}: gpu {it} iterations, oracle {ito}, max |dx| {err:.3e}", flush=True)
    assert rc == 0
    # one PCG step more or less at the stopping threshold at most
    assert abs(it - ito) <= 1
    # on the same Krylov path the dot products' summation order is the only
    # difference (measured <= 3e-11); a step more or less at the stopping
    # threshold leaves both within the solver's residual of ~100 coordinates
    assert err <= (1e-8 if it == ito else 1e-3)


@pytest.mark.parametrize("name", ["plain_1k", "ecp5_25k", "ecp5_85k"])
def test_quadratic_seed_matches_the_oracle(olib, name):
    p = getattr(synth, name)()
    t0 = time.perf_counter()
    xy, it = seed.quadratic_seed(p)
    t_gpu = time.perf_counter() - t0
    xo = np.zeros(2 * p.n_units)
    rep = abi.texo_gp_report()
    t0 = time.perf_counter()
    rc = olib.oracle_quadratic_seed(C.byref(p.struct()), xo.ctypes.data_as(C.POINTER(C.c_double)), C.byref(rep))
    t_cpu = time.perf_counter() - t0
    xo = xo.reshape(-1, 2)
    err = np.abs(xy - xo).max()
    print(f"{name}: gpu {it} iterations {t_gpu * 1e3:.1f} ms, oracle {rep.iterations} iterations "
          f"{t_cpu * 1e3:.1f} ms, max |dxy| {err:.3e}", flush=True)
    assert rc == 0
    assert abs(it - rep.iterations) <= 2
    # the same iterations: the dot products' order is the only difference
    # (measured <= 1e-11); otherwise 1e-3 of a site on grids of ~200 a side
    assert err <= (1e-8 if it == rep.iterations else 1e-3)
