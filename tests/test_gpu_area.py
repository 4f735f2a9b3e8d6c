"""Area-round pieces on the GPU against the oracle: register-clustering
areas (register_clustering.rs:17-81) on a 20k-register control-set group,
and the whole area proposal of an area round (propose_area_adjustment,
eplace.rs:627-729: RUDY utilisation, routability areas, clustering areas and
adjust_instance_areas) at the same coordinates.

Tolerances: the clustering sums add the pairs in the reference's order, so
areas differ only through the device erf (relative 1e-12 measured bound);
the area proposal inherits the fixed-point RUDY's nets * 2^-40 (relative
1e-9)."""
import ctypes as C

import numpy as np
import pytest

import oracle
import texo_amd
from texo_amd import abi, synth
from tests.test_gpu_parity import area_round_problem

pytestmark = pytest.mark.gpu


def clustering(lib, name, regs_cell, regs_xy, ctrl, movable, device=True):
    n = len(regs_cell)
    areas = np.zeros(max(n, 1))
    u32 = lambda a: abi.ptr(np.ascontiguousarray(a, np.uint32), C.c_uint32)
    u64 = lambda a: abi.ptr(np.ascontiguousarray(a, np.uint64), C.c_uint64)
    xy = np.ascontiguousarray(regs_xy, float).reshape(-1)
    args = [n, u32(regs_cell), abi.ptr(xy, C.c_double), len(ctrl[0]), u32(ctrl[0]),
            u64(np.stack([ctrl[1], ctrl[2]], 1).reshape(-1)), u64(ctrl[3]), movable, abi.ptr(areas, C.c_double)]
    fn = getattr(lib, name)
    rc = fn(0, *args) if device else fn(*args)
    return rc, areas[:n]


def declare(lib):
    P, u32, u64, d = C.POINTER, C.c_uint32, C.c_uint64, C.c_double
    lib.texo_gp_clustering_areas.argtypes = [C.c_int32, u64, P(u32), P(d), u64, P(u32), P(u64), P(u64), u64, P(d)]
    lib.texo_gp_clustering_areas.restype = C.c_int


@pytest.mark.parametrize("n_regs,groups,ces", [(20000, 1, 4), (6000, 5, 3), (700, 2, 20)])
def test_clustering_areas_match_oracle(n_regs, groups, ces):
    # one control-set group of 20k registers (the 84k-cell configuration's
    # flip-flop count is 23.5k); several groups; a group with more
    # clock-enable classes than the kernel holds (host path)
    print(f"clustering {n_regs} registers, {groups} groups, {ces} CE classes", flush=True)
    g, o = texo_amd.load(), oracle.load()
    declare(g)
    rng = np.random.default_rng(n_regs)
    cells = rng.permutation(n_regs * 2)[:n_regs]
    xy = rng.uniform(0, 256, (n_regs, 2))
    # clusters of registers so that windows overlap (sigma = sqrt(1e-5 * 84k) ~ 0.92)
    xy[: n_regs // 2] = rng.normal(128, 3.0, (n_regs // 2, 2))
    ctrl = (cells, rng.integers(0, groups, n_regs), np.full(n_regs, 7), rng.integers(0, ces, n_regs))
    rc_g, a_g = clustering(g, "texo_gp_clustering_areas", cells, xy, ctrl, 84000)
    o.oracle_set_threads(16)
    n = n_regs
    areas = np.zeros(n)
    u32 = lambda a: abi.ptr(np.ascontiguousarray(a, np.uint32), C.c_uint32)
    u64 = lambda a: abi.ptr(np.ascontiguousarray(a, np.uint64), C.c_uint64)
    xyf = np.ascontiguousarray(xy, float).reshape(-1)
    rc_o = o.oracle_clustering_areas(n, u32(cells), abi.ptr(xyf, C.c_double), n, u32(ctrl[0]),
                                     u64(np.stack([ctrl[1], ctrl[2]], 1).reshape(-1)), u64(ctrl[3]), 84000,
                                     abi.ptr(areas, C.c_double))
    o.oracle_set_threads(1)
    assert rc_g == 0 and rc_o == 0
    err = np.abs(a_g - areas) / np.abs(areas)
    print(f"  max relative difference {err.max():.3e}; areas {areas.min():.3f}..{areas.max():.3f}", flush=True)
    assert err.max() <= 1e-12


def test_area_proposal_matches_oracle():
    p = area_round_problem()
    g = texo_amd.Session(p)
    o = abi.Session(oracle.load(), "oracle", p)
    o.iterate(40)
    x = o.coordinates()
    pg, po = g.area_proposal(x), o.area_proposal(x)
    rm = np.abs(pg["member_areas"] - po["member_areas"]) / po["member_areas"]
    rf = np.abs(pg["filler_areas"] - po["filler_areas"]) / np.maximum(po["filler_areas"], 1e-300)
    print(f"area proposal: members {rm.max():.3e} fillers {rf.max():.3e} reset {pg['requires_reset']} "
          f"increase {pg['total_relative_increase']:.6f} vs {po['total_relative_increase']:.6f}", flush=True)
    assert pg["requires_reset"] == po["requires_reset"] and po["requires_reset"]
    assert rm.max() <= 1e-9 and rf.max() <= 1e-9
    assert abs(pg["total_relative_increase"] - po["total_relative_increase"]) <= 1e-9 * po["total_relative_increase"]
    # at least some members grew from routability or clustering
    assert (po["member_areas"] > 1.0).sum() > 0
    g.close()
    o.close()
