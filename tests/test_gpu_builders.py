"""GP input builders on the GPU against the oracle: the routing capacity map
bit for bit, the routing demand map within its fixed-point quantum
(nets covering a tile * 2^-40) and run-to-run identical."""
import numpy as np
import pytest

import oracle
from texo_amd import builders, synth
from tests.test_builders import capacity_case

pytestmark = pytest.mark.gpu
FX = 2.0 ** -40


def test_capacity_map_known_answer_and_random_match_oracle():
    lib = oracle.load()
    W, H, orient, xy, cap, pf, pt, pb = capacity_case()
    h, v = builders.routing_capacity_map(W, H, orient, xy, cap, pf, pt, pb)
    assert h[0, 0] == 2 and h[0, 1] == 0 and v[1, 2] == 4 and h.sum() == 2 and v.sum() == 4
    rng = np.random.default_rng(3)
    for W, H, nw, npip in ((126, 95, 200_000, 500_000), (40, 30, 5000, 3000), (1, 1, 70_000, 70_000)):
        print(f"capacity map {W}x{H}, {nw} wires, {npip} pips", flush=True)
        orient = rng.integers(0, 3, nw)
        xy = np.stack([rng.integers(0, W, nw), rng.integers(0, H, nw)], 1)
        capv = rng.integers(0, 65536, nw)
        pf, pt = rng.integers(0, nw, npip), rng.integers(0, nw, npip)
        pb = rng.random(npip) < 0.1
        got = builders.routing_capacity_map(W, H, orient, xy, capv, pf, pt, pb)
        ref = builders.routing_capacity_map(W, H, orient, xy, capv, pf, pt, pb, lib=lib, prefix="oracle")
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])


def placed_netlist(problem, rng):
    """The problem's nets with every cell at an integer tile (a legal-looking
    placement: cells jittered around their start coordinates), a few cells
    unplaced and a few clock-driven nets."""
    W, H = problem.width, problem.height
    unit_xy = problem.unit_initial_xy.reshape(-1, 2)
    cell_xy = unit_xy[problem.cell_unit] + problem.cell_offset_xy.reshape(-1, 2)
    cell_xy = np.where(problem.cell_fixed[:, None] != 0, problem.cell_fixed_xy.reshape(-1, 2), cell_xy)
    pts = np.clip(np.rint(cell_xy + rng.normal(0, W / 6, cell_xy.shape)), 0, [W - 1, H - 1]).astype(np.uint32)
    unplaced = rng.random(len(pts)) < 0.01
    pts[unplaced] = 0xFFFFFFFF
    pin_xy = pts[problem.pin_cell]
    clock = (rng.random(problem.n_nets) < 0.002).astype(np.uint8)
    return problem.net_pin_start, pin_xy, clock


@pytest.mark.parametrize("name", ["ecp5_25k", "ecp5_85k"])
def test_demand_map_matches_oracle(name):
    lib = oracle.load()
    h, v, inc, exc = builders.routing_demand_map(4, 3, [0, 2], [(0, 0), (3, 2)], [0])
    assert inc == 1 and abs(h[1, 1] - 0.25) < 1e-9 and abs(v[1, 1] - 1 / 6) < 1e-9
    p = synth.CONFIGS[name]()
    rng = np.random.default_rng(8)
    ns, pxy, clock = placed_netlist(p, rng)
    print(f"demand map {name}: {p.n_nets} nets on {p.width}x{p.height}", flush=True)
    g = builders.routing_demand_map(p.width, p.height, ns, pxy, clock)
    g2 = builders.routing_demand_map(p.width, p.height, ns, pxy, clock)
    r = builders.routing_demand_map(p.width, p.height, ns, pxy, clock, lib=lib, prefix="oracle")
    assert g[2:] == r[2:] and r[3] > 0
    assert np.array_equal(g[0], g2[0]) and np.array_equal(g[1], g2[1])
    tol = p.n_nets * FX
    print(f"  max error h {np.abs(g[0] - r[0]).max():.3e} v {np.abs(g[1] - r[1]).max():.3e} (bound {tol:.3e})",
          flush=True)
    assert np.abs(g[0] - r[0]).max() <= tol and np.abs(g[1] - r[1]).max() <= tol
    assert np.array_equal(g[0] > 0, r[0] > 0) and np.array_equal(g[1] > 0, r[1] > 0)
