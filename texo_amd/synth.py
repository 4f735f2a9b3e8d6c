"""Seeded synthetic placement problems for the benchmark and parity tests.

Each builder returns an `abi.Problem` in the flattened form texo-pnr's
`eplace::place` receives (crates/texo-pnr/src/eplace.rs:242-253):

* units and cells: unit-area LUT4 and FF cells, two-LUT carry macros with
  member offsets (0,0)/(0,1), fixed DP16KD blocks (Memory) and fixed IO cells;
* capacity per exchangeable kind so that sum(capacity) * 0.9 equals the demand
  (target density 0.9), spread uniformly over the usable bins, zero in the two
  memory columns; memory BELs exist only where the fixed blocks sit, so the
  fixed occupancy consumes all Memory capacity and no Memory field arises
  (eplace.rs:263-271);
* nets: driver first, pins drawn uniformly among all cells;
* sink weights 1 + round(10 * c^k) on the reference's 1/1024 fixed point
  (crates/texo-flow/src/lib.rs:3355-3433), c ~ Beta(0.5, 4), giving timed arcs
  with extra weight w - 1 (analytical_placement.rs:145-157);
* initial coordinates clumped around the die centre with a Gaussian of
  sigma = side/12, mimicking the quadratic seed that precedes eplace.
"""
from __future__ import annotations

import math
from typing import Optional

import numpy as np

from . import abi

TIMING_SCALE_BITS = 10
TIMING_WEIGHT = 10


def scaled_binary_fraction(numerator: int, denominator: int, bits: int) -> int:
    """floor(numerator / denominator * 2^bits), saturating at 2^bits
    (crates/texo-flow/src/lib.rs:3413-3433)."""
    if numerator >= denominator:
        return 1 << bits
    remainder, scaled = numerator, 0
    for _ in range(bits):
        scaled <<= 1
        complement = denominator - remainder
        if remainder >= complement:
            remainder -= complement
            scaled |= 1
        else:
            remainder *= 2
    return scaled


def criticality_weight(path_delay_ps: int, worst_path_delay_ps: int, exponent: int) -> int:
    """ecp5_placement_criticality_weight (crates/texo-flow/src/lib.rs:3355-3390)."""
    scale = 1 << TIMING_SCALE_BITS
    exponent = max(exponent, 1)
    if worst_path_delay_ps == 0:
        scaled = 0
    else:
        scaled = scaled_binary_fraction(min(path_delay_ps, worst_path_delay_ps), worst_path_delay_ps,
                                        TIMING_SCALE_BITS)
    numerator = scaled ** exponent
    denominator = scale ** exponent
    if numerator.bit_length() <= 128 and denominator.bit_length() <= 128 \
            and (TIMING_WEIGHT * numerator).bit_length() <= 128:
        extra = (TIMING_WEIGHT * numerator + denominator // 2) // denominator
    else:
        # scaled_fraction_power fallback (texo-flow/src/lib.rs:3395-3411)
        result, factor, e = scale, scaled, exponent
        while e:
            if e & 1:
                result = (result * factor + scale // 2) // scale
            e >>= 1
            if e:
                factor = (factor * factor + scale // 2) // scale
        extra = (TIMING_WEIGHT * result + scale // 2) // scale
    return 1 + min(extra, TIMING_WEIGHT)


def _weight_table(exponent: int) -> np.ndarray:
    """Weight for every 10-bit criticality code (the only input bits that the
    fixed-point formula reads)."""
    scale = 1 << TIMING_SCALE_BITS
    return np.array([criticality_weight(code, scale, exponent) for code in range(scale + 1)], np.uint64)


def _spread_capacity(rng, demand: int, usable: np.ndarray, target_density: float = 0.9) -> np.ndarray:
    """Integer capacity per bin with sum = round(demand / target_density),
    as uniform as possible over the usable bins."""
    total = int(round(demand / target_density))
    bins = np.flatnonzero(usable)
    cap = np.zeros(usable.size, np.uint32)
    base, extra = divmod(total, bins.size)
    cap[bins] = base
    if extra:
        cap[rng.choice(bins, size=extra, replace=False)] += 1
    return cap


def _degrees_geometric(rng, n: int, p: float, cap: Optional[int]) -> np.ndarray:
    d = 2 + (rng.geometric(p, size=n) - 1)
    if cap is not None:
        d = np.minimum(d, cap)
    return d.astype(np.int64)


def _degrees_rent(rng, n: int, p: float, dmax: int) -> np.ndarray:
    d = np.arange(2, dmax + 1)
    w = d.astype(np.float64) ** (-(1.0 + p))
    return rng.choice(d, size=n, p=w / w.sum()).astype(np.int64)


def make_problem(*, seed: int, side: int, n_lut: int, n_ff: int, n_carry_units: int, n_memory: int,
                 n_io: int, n_nets: int, degree: str, degree_p: float, degree_cap: Optional[int],
                 high_fanout: int = 0, high_fanout_range=(256, 512), weight_exponent: Optional[int] = None,
                 routing_capacity: Optional[int] = None, register_controls: bool = False,
                 region_lut_sites: Optional[float] = None, name: str = "synthetic") -> abi.Problem:
    rng = np.random.default_rng(seed)
    W = H = side
    G = W * H

    # --- cells and units ---------------------------------------------------
    # unit order: LUT, FF, carry, memory (fixed), IO (fixed)
    n_units = n_lut + n_ff + n_carry_units + n_memory + n_io
    n_cells = n_lut + n_ff + 2 * n_carry_units + n_memory + n_io
    lut_kind = abi.KIND_LUT(4)
    unit_fixed = np.zeros(n_units, np.uint8)
    cell_unit = np.zeros(n_cells, np.uint32)
    cell_fixed = np.zeros(n_cells, np.uint8)
    cell_fixed_xy = np.zeros((n_cells, 2))
    cell_offset = np.zeros((n_cells, 2))
    member_counts = np.zeros(n_units, np.int64)

    u = 0
    c = 0
    # single-cell LUT and FF units
    single = n_lut + n_ff
    cell_unit[:single] = np.arange(single)
    member_counts[:single] = 1
    member_kind = [np.full(n_lut, lut_kind, np.uint16), np.full(n_ff, abi.KIND_REGISTER, np.uint16)]
    member_off = [np.zeros((single, 2))]
    member_cell = [np.arange(single, dtype=np.uint32)]
    u, c = single, single
    # carry units: two LUT cells, offsets (0,0), (0,1)
    cu = np.arange(u, u + n_carry_units)
    cell_unit[c:c + 2 * n_carry_units] = np.repeat(cu, 2)
    cell_offset[c + 1:c + 2 * n_carry_units:2, 1] = 1.0
    member_counts[u:u + n_carry_units] = 2
    member_kind.append(np.full(2 * n_carry_units, lut_kind, np.uint16))
    off = np.zeros((2 * n_carry_units, 2))
    off[1::2, 1] = 1.0
    member_off.append(off)
    member_cell.append(np.arange(c, c + 2 * n_carry_units, dtype=np.uint32))
    u += n_carry_units
    c += 2 * n_carry_units

    # memory columns
    mem_cols = [W // 3, (2 * W) // 3] if n_memory else []
    usable = np.ones((H, W), bool)
    if region_lut_sites is not None:
        # logic capacity confined to a central square holding about
        # `region_lut_sites` LUT sites per bin at the 0.9 target density
        s = min(side, int(math.ceil(math.sqrt((n_lut + 2 * n_carry_units) / (0.9 * region_lut_sites)))) + 1)
        lo = (side - s) // 2
        usable[:] = False
        usable[lo:lo + s, lo:lo + s] = True
    for x in mem_cols:
        usable[:, x] = False
    mem_xy = np.zeros((n_memory, 2))
    if n_memory:
        per_col = (n_memory + 1) // 2
        for i in range(n_memory):
            col = mem_cols[i % 2]
            row = (i // 2) * H // per_col
            mem_xy[i] = (col, row)
    mu = np.arange(u, u + n_memory)
    unit_fixed[mu] = 1
    cell_unit[c:c + n_memory] = mu
    cell_fixed[c:c + n_memory] = 1
    cell_fixed_xy[c:c + n_memory] = mem_xy
    u += n_memory
    c += n_memory

    # IO cells on the die boundary
    io_xy = np.zeros((n_io, 2))
    if n_io:
        perimeter = 2 * (W - 1) + 2 * (H - 1)
        for i in range(n_io):
            t = (i * perimeter) // n_io
            if t < W - 1:
                io_xy[i] = (t, 0)
            elif t < W - 1 + H - 1:
                io_xy[i] = (W - 1, t - (W - 1))
            elif t < 2 * (W - 1) + H - 1:
                io_xy[i] = (W - 1 - (t - (W - 1 + H - 1)), H - 1)
            else:
                io_xy[i] = (0, H - 1 - (t - (2 * (W - 1) + H - 1)))
    iu = np.arange(u, u + n_io)
    unit_fixed[iu] = 1
    cell_unit[c:c + n_io] = iu
    cell_fixed[c:c + n_io] = 1
    cell_fixed_xy[c:c + n_io] = io_xy
    u += n_io
    c += n_io
    assert u == n_units and c == n_cells

    unit_member_start = np.zeros(n_units + 1, np.uint32)
    unit_member_start[1:] = np.cumsum(member_counts)
    member_kind = np.concatenate(member_kind)
    member_off = np.concatenate(member_off)
    member_cell = np.concatenate(member_cell)

    # --- capacity ---------------------------------------------------------
    usable_flat = usable.reshape(-1)
    lut_cap = _spread_capacity(rng, n_lut + 2 * n_carry_units, usable_flat)
    cap_kinds = [lut_kind]
    caps = [lut_cap]
    if n_ff:
        cap_kinds.append(abi.KIND_REGISTER)
        caps.append(_spread_capacity(rng, n_ff, usable_flat))
    fixed_kind = np.zeros(0, np.uint16)
    fixed_xy = np.zeros(0, np.uint32)
    if n_memory:
        mem_cap = np.zeros(G, np.uint32)
        for x, y in mem_xy.astype(np.int64):
            mem_cap[y * W + x] += 1
        cap_kinds.append(abi.KIND_MEMORY)
        caps.append(mem_cap)
        fixed_kind = np.full(n_memory, abi.KIND_MEMORY, np.uint16)
        fixed_xy = mem_xy.astype(np.uint32).reshape(-1)

    # --- bounds and initial coordinates -----------------------------------
    bounds = np.zeros((n_units, 4))
    bounds[:, 2] = W - 1
    bounds[:, 3] = H - 1

    def kind_box(cap):
        ys, xs = np.nonzero(cap.reshape(H, W))
        return xs.min(), ys.min(), xs.max(), ys.max()

    lb = kind_box(lut_cap)
    bounds[:n_lut] = lb
    if n_ff:
        bounds[n_lut:n_lut + n_ff] = kind_box(caps[1])
    if n_carry_units:
        bounds[n_lut + n_ff:n_lut + n_ff + n_carry_units] = (lb[0], lb[1], lb[2], lb[3] - 1)
    initial = np.zeros((n_units, 2))
    sigma = side / 12.0
    movable = n_lut + n_ff + n_carry_units
    initial[:movable] = (W - 1) / 2.0 + rng.normal(0.0, sigma, size=(movable, 2))
    initial[:movable] = np.clip(initial[:movable], bounds[:movable, :2], bounds[:movable, 2:])
    fixed_units = np.arange(movable, n_units)
    fixed_cells = np.arange(n_cells - n_memory - n_io, n_cells)
    initial[fixed_units] = cell_fixed_xy[fixed_cells]
    bounds[fixed_units, :2] = cell_fixed_xy[fixed_cells]
    bounds[fixed_units, 2:] = cell_fixed_xy[fixed_cells]

    # --- nets --------------------------------------------------------------
    if degree == "geometric":
        deg = _degrees_geometric(rng, n_nets, degree_p, degree_cap)
    elif degree == "rent":
        deg = _degrees_rent(rng, n_nets, degree_p, degree_cap or 64)
    else:
        raise ValueError(degree)
    if high_fanout:
        lo, hi = high_fanout_range
        deg = np.concatenate([deg, rng.integers(lo, hi + 1, size=high_fanout)])
    deg = np.minimum(deg, n_cells)
    net_pin_start = np.zeros(deg.size + 1, np.uint32)
    net_pin_start[1:] = np.cumsum(deg)
    pin_cell = rng.integers(0, n_cells, size=int(deg.sum()), dtype=np.int64).astype(np.uint32)

    # --- timing weights -> timed arcs -------------------------------------
    arc_cells = np.zeros(0, np.uint32)
    arc_extra = np.zeros(0, np.uint64)
    if weight_exponent is not None:
        table = _weight_table(weight_exponent)
        crit = rng.beta(0.5, 4.0, size=pin_cell.size)
        codes = np.minimum((crit * (1 << TIMING_SCALE_BITS)).astype(np.int64), 1 << TIMING_SCALE_BITS)
        w = table[codes]
        is_sink = np.ones(pin_cell.size, bool)
        is_sink[net_pin_start[:-1]] = False
        net_of_pin = np.repeat(np.arange(deg.size), deg)
        sel = np.flatnonzero(is_sink & (w > 1))
        drivers = pin_cell[net_pin_start[net_of_pin[sel]]]
        arc_cells = np.stack([drivers, pin_cell[sel]], axis=1).astype(np.uint32).reshape(-1)
        arc_extra = (w[sel] - 1).astype(np.uint64)

    routing_h = routing_v = None
    if routing_capacity is not None:
        routing_h = np.full(G, routing_capacity, np.uint32)
        routing_v = np.full(G, routing_capacity, np.uint32)

    reg_cell = np.zeros(0, np.uint32)
    reg_clk = np.zeros(0, np.uint64)
    reg_ce = np.zeros(0, np.uint64)
    if register_controls and n_ff:
        reg_cell = np.arange(n_lut, n_lut + n_ff, dtype=np.uint32)
        clock = rng.integers(0, 4, size=n_ff)
        lsr = rng.integers(0, 2, size=n_ff)
        reg_clk = np.stack([clock, lsr], axis=1).astype(np.uint64).reshape(-1)
        reg_ce = rng.integers(0, 16, size=n_ff).astype(np.uint64)

    return abi.Problem(
        width=W, height=H, cell_unit=cell_unit, cell_fixed=cell_fixed, cell_fixed_xy=cell_fixed_xy,
        cell_offset_xy=cell_offset, unit_fixed=unit_fixed, unit_initial_xy=initial, unit_bounds=bounds,
        unit_member_start=unit_member_start, member_kind=member_kind, member_offset_xy=member_off,
        member_cell=member_cell, capacity_kind=np.array(cap_kinds, np.uint16),
        capacity=np.concatenate(caps), net_pin_start=net_pin_start, pin_cell=pin_cell,
        fixed_occupancy_kind=fixed_kind, fixed_occupancy_xy=fixed_xy, arc_cells=arc_cells,
        arc_extra_weight=arc_extra, routing_horizontal=routing_h, routing_vertical=routing_v,
        register_cell=reg_cell, register_clock_lsr=reg_clk, register_ce=reg_ce,
        meta={"name": name, "seed": seed, "side": side, "cells": n_cells, "units": n_units,
              "nets": int(deg.size), "pins": int(pin_cell.size), "arcs": int(arc_extra.size),
              "target_density": 0.9},
    )


def plain_1k(seed: int = 0) -> abi.Problem:
    """1,024 movable unit-area LUT cells + 64 fixed IO cells, 1,200 nets of
    degree 2 + Geometric(0.5) capped at 8, 64x64 bins."""
    return make_problem(seed=seed, side=64, n_lut=1024, n_ff=0, n_carry_units=0, n_memory=0, n_io=64,
                        n_nets=1200, degree="geometric", degree_p=0.5, degree_cap=8, name="plain_1k")


def ecp5_25k(seed: int = 1, **kw) -> abi.Problem:
    """24k cells (16,800 LUT4, 6,720 FF, 240 two-LUT carry units), 56 fixed
    DP16KD, 30k Rent-like nets (p=0.6, max 64), 128x128 bins."""
    return make_problem(seed=seed, side=128, n_lut=16800, n_ff=6720, n_carry_units=240, n_memory=56, n_io=0,
                        n_nets=30000, degree="rent", degree_p=0.6, degree_cap=64, name="ecp5_25k", **kw)


def ecp5_85k(seed: int = 2, weight_exponent: int = 4, routing_capacity: int = 8, **kw) -> abi.Problem:
    """84k cells (58,800 LUT4, 23,520 FF, 840 two-LUT carry units), 208 fixed
    DP16KD, 100k nets of degree 2 + Geometric(0.4) plus 16 nets of 256-512
    pins, timing weights 1 + 10 c^4 with c ~ Beta(0.5, 4), 256x256 bins,
    uniform routing capacity for RUDY."""
    return make_problem(seed=seed, side=256, n_lut=58800, n_ff=23520, n_carry_units=840, n_memory=208, n_io=0,
                        n_nets=100000, degree="geometric", degree_p=0.4, degree_cap=64, high_fanout=16,
                        weight_exponent=weight_exponent, routing_capacity=routing_capacity, name="ecp5_85k",
                        **kw)


def serving_job(seed: int, cells: int, weight_exponent: int) -> abi.Problem:
    """One job of the mixed-size serving stream: grid side =
    clamp(nextpow2(sqrt(cells/4)), 64, 256), the ECP5 cell mix, nets
    1.2 per cell of degree 2 + Geometric(0.4); logic capacity in a central
    square at about 2 LUT sites per bin (a design occupies part of the die)."""
    side = 1 << max(0, math.ceil(math.log2(max(1.0, math.sqrt(cells / 4.0)))))
    side = min(max(side, 64), 256)
    n_carry_units = max(0, int(cells * 0.01))
    n_ff = int(cells * 0.28)
    n_lut = cells - n_ff - 2 * n_carry_units
    n_memory = min(208, max(2, cells // 400))
    return make_problem(seed=seed, side=side, n_lut=n_lut, n_ff=n_ff, n_carry_units=n_carry_units,
                        n_memory=n_memory, n_io=0, n_nets=int(cells * 1.2), degree="geometric",
                        degree_p=0.4, degree_cap=64, weight_exponent=weight_exponent,
                        region_lut_sites=2.0, name=f"job{seed}")


CONFIGS = {"plain_1k": plain_1k, "ecp5_25k": ecp5_25k, "ecp5_85k": ecp5_85k}


def multistart(problem: abi.Problem, start: int) -> abi.Problem:
    """The same design with the movable units' initial coordinates redrawn
    (Gaussian around the die centre, sigma = side/12, clipped to each unit's
    bounds) from seed `start`: one start of a multi-start placement."""
    import dataclasses
    rng = np.random.default_rng(1_000_003 * (start + 1))
    W, H = problem.width, problem.height
    movable = problem.unit_fixed == 0
    init = problem.unit_initial_xy.reshape(-1, 2).copy()
    b = problem.unit_bounds.reshape(-1, 4)
    sigma = max(W, H) / 12.0
    draw = (np.array([W - 1, H - 1]) / 2.0) + rng.normal(0.0, sigma, size=(int(movable.sum()), 2))
    init[movable] = np.clip(draw, b[movable, :2], b[movable, 2:])
    meta = dict(problem.meta, start=start)
    return dataclasses.replace(problem, unit_initial_xy=init.reshape(-1), meta=meta)


def serving_stream(n_jobs: int, seed: int = 2026, rate: float = 20.0, min_cells: int = 1000, max_cells: int = 84000):
    """Mixed-size serving stream: cell counts log-uniform in [min, max],
    weight exponents from {2, 4, 6}, per-job seeds, Poisson arrivals at `rate`
    jobs/s.  Returns [(arrival_seconds, problem)]."""
    rng = np.random.default_rng(seed)
    cells = np.exp(rng.uniform(math.log(min_cells), math.log(max_cells), n_jobs)).astype(int)
    exps = rng.choice([2, 4, 6], size=n_jobs)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n_jobs))
    arrivals -= arrivals[0]
    return [(float(arrivals[i]), serving_job(seed * 1000 + i, int(cells[i]), int(exps[i]))) for i in range(n_jobs)]


# ---------------------------------------------------------------------------
# Incremental (ECO) re-placement on a resident design
# ---------------------------------------------------------------------------
def _assign_sites(unit_xy: np.ndarray, site_bins: np.ndarray, W: int) -> np.ndarray:
    """Legal sites for units: the units are cut into row bands by y (band k
    holds as many units as row band k of the sites has sites) and matched to
    the sites by x inside each band, so the placement keeps the units'
    coarse 2-D order.  `site_bins` are row-major bin indices (one per site,
    repeated by capacity), at least as many as units."""
    n = unit_xy.shape[0]
    out = np.zeros((n, 2))
    if n == 0:
        return out
    pick = site_bins[np.linspace(0, site_bins.size - 1, n).round().astype(np.int64)]
    sx, sy = pick % W, pick // W
    order_y = np.argsort(unit_xy[:, 1], kind="stable")
    rows, counts = np.unique(sy, return_counts=True)
    start = 0
    for row, cnt in zip(rows, counts):
        units = order_y[start:start + cnt]
        units = units[np.argsort(unit_xy[units, 0], kind="stable")]
        xs = np.sort(sx[sy == row])
        out[units, 0] = xs
        out[units, 1] = row
        start += cnt
    return out


def eco_design(seed: int = 5, cells: int = 250_000, side: int = 512, nets: int = 300_000,
               high_fanout: int = 16) -> abi.Problem:
    """The resident design of the incremental stream: the ECP5 cell mix
    (70% LUT4, 28% FF, 2% of cells in two-LUT carry units) with fixed DP16KD
    blocks, nets of degree 2 + Geometric(0.4) plus `high_fanout` nets of
    256-512 pins, on side x side bins, with a legal placement as its
    coordinates: every movable unit's members sit on sites of their kind, no
    bin over its capacity (a carry unit takes a LUT site at (x, y) and one at
    (x, y + 1))."""
    n_carry_units = int(cells * 0.01)
    n_ff = int(cells * 0.28)
    n_lut = cells - n_ff - 2 * n_carry_units
    n_memory = int(round(208 * cells / 84_000))
    p = make_problem(seed=seed, side=side, n_lut=n_lut, n_ff=n_ff, n_carry_units=n_carry_units, n_memory=n_memory,
                     n_io=0, n_nets=nets, degree="geometric", degree_p=0.4, degree_cap=64, high_fanout=high_fanout,
                     name="eco_design")
    W, H = p.width, p.height
    G = W * H
    kinds = list(p.capacity_kind)
    lut_cap = p.capacity[kinds.index(abi.KIND_LUT(4)) * G:(kinds.index(abi.KIND_LUT(4)) + 1) * G].astype(np.int64)
    ff_cap = p.capacity[kinds.index(abi.KIND_REGISTER) * G:(kinds.index(abi.KIND_REGISTER) + 1) * G].astype(np.int64)
    init = p.unit_initial_xy.reshape(-1, 2).copy()
    sites = init.copy()
    carry = np.arange(n_lut + n_ff, n_lut + n_ff + n_carry_units)
    # carry units on vertically adjacent LUT sites, pairs starting on even rows
    b = np.arange(G - W)
    pairs = b[(lut_cap[:G - W] > 0) & (lut_cap[W:] > 0) & ((b // W) % 2 == 0)]
    if pairs.size < n_carry_units:
        raise ValueError("not enough vertically adjacent LUT sites for the carry units")
    sites[carry] = _assign_sites(init[carry], pairs, W)
    used = (sites[carry, 1] * W + sites[carry, 0]).astype(np.int64)
    np.subtract.at(lut_cap, used, 1)
    np.subtract.at(lut_cap, used + W, 1)
    lut_sites = np.repeat(np.arange(G), lut_cap)
    ff_sites = np.repeat(np.arange(G), ff_cap)
    lut_units, ff_units = np.arange(n_lut), np.arange(n_lut, n_lut + n_ff)
    sites[lut_units] = _assign_sites(init[lut_units], lut_sites, W)
    sites[ff_units] = _assign_sites(init[ff_units], ff_sites, W)
    import dataclasses
    meta = dict(p.meta, name="eco_design", resident=True)
    return dataclasses.replace(p, unit_initial_xy=sites.reshape(-1), meta=meta)


class EcoStream:
    """Incremental requests on a resident design: each request re-places a
    connected subset of the movable units (grown breadth-first over the nets
    from a random unit) with every other unit fixed at its resident site, as
    texo-pnr does for a unit with a single placement choice: its cells become
    fixed coordinates and its members fixed occupancy that the density model
    subtracts from the capacity (eplace.rs:263-271)."""

    def __init__(self, design: abi.Problem):
        self.design = design
        p = design
        U = p.unit_fixed.size
        self.movable = np.flatnonzero(p.unit_fixed == 0)
        deg = np.diff(p.net_pin_start.astype(np.int64))
        net_of_pin = np.repeat(np.arange(deg.size), deg)
        pin_unit = p.cell_unit[p.pin_cell].astype(np.int64)
        # unit -> nets and net -> units (CSR)
        order = np.argsort(pin_unit, kind="stable")
        self.unit_start = np.zeros(U + 1, np.int64)
        np.add.at(self.unit_start, pin_unit + 1, 1)
        self.unit_start = np.cumsum(self.unit_start)
        self.unit_nets = net_of_pin[order]
        self.net_start = p.net_pin_start.astype(np.int64)
        self.net_units = pin_unit

    def subset(self, rng, size: int) -> np.ndarray:
        """A connected set of `size` movable units, breadth-first from a
        random movable unit (each level is complete before the next starts,
        so any prefix of the order is connected); a start whose component
        is smaller than `size` (a unit without nets) is drawn again."""
        fixed = self.design.unit_fixed != 0
        best = np.zeros(0, np.int64)
        for _ in range(64):
            seen = np.zeros(fixed.size, bool)
            start = int(rng.choice(self.movable))
            order = [np.array([start])]
            seen[start] = True
            frontier = order[0]
            total = 1
            while total < size and frontier.size:
                nets = [self.unit_nets[self.unit_start[u]:self.unit_start[u + 1]] for u in frontier]
                nets = np.unique(np.concatenate(nets))
                if nets.size == 0:
                    break
                nxt = np.unique(np.concatenate([self.net_units[self.net_start[n]:self.net_start[n + 1]] for n in nets]))
                nxt = nxt[~seen[nxt] & ~fixed[nxt]]
                seen[nxt] = True
                order.append(nxt)
                total += nxt.size
                frontier = nxt
            found = np.concatenate(order)[:size]
            if found.size >= size:
                return found
            if found.size > best.size:
                best = found
        return best

    def request(self, seed: int, fraction: Optional[float] = None) -> abi.Problem:
        """One incremental request: a connected subset of 1-5% of the movable
        units (or `fraction`) is ripped up and re-placed: it starts clumped
        around its resident centroid (a Gaussian of sigma = side/12 for the
        side of a square holding the subset at density 0.9, clipped to each
        unit's bounds, as the initial coordinates of a full placement are);
        every other unit is fixed at its resident site."""
        import dataclasses
        rng = np.random.default_rng(seed)
        p = self.design
        frac = float(rng.uniform(0.01, 0.05)) if fraction is None else fraction
        keep = self.subset(rng, max(1, int(round(frac * self.movable.size))))
        U = p.unit_fixed.size
        fix = np.zeros(U, bool)
        fix[self.movable] = True
        fix[keep] = False
        xy = p.unit_initial_xy.reshape(-1, 2)
        init = xy.copy()
        sigma = math.sqrt(keep.size / 0.9) / 12.0
        b = p.unit_bounds.reshape(-1, 4)
        init[keep] = np.clip(xy[keep].mean(axis=0) + rng.normal(0.0, sigma, size=(keep.size, 2)), b[keep, :2],
                             b[keep, 2:])
        unit_fixed = p.unit_fixed.copy()
        unit_fixed[fix] = 1
        bounds = p.unit_bounds.reshape(-1, 4).copy()
        bounds[fix, :2] = xy[fix]
        bounds[fix, 2:] = xy[fix]
        cell_fixed = p.cell_fixed.copy()
        cell_xy = p.cell_fixed_xy.reshape(-1, 2).copy()
        cells = np.flatnonzero(fix[p.cell_unit])
        cell_fixed[cells] = 1
        cell_xy[cells] = xy[p.cell_unit[cells]] + p.cell_offset_xy.reshape(-1, 2)[cells]
        # members of the newly fixed units occupy their sites
        ms = p.unit_member_start.astype(np.int64)
        counts = np.diff(ms)
        owner_all = np.repeat(np.arange(U), counts)
        members = np.flatnonzero(fix[owner_all])
        owner = owner_all[members]
        occ_xy = (xy[owner] + p.member_offset_xy.reshape(-1, 2)[members]).astype(np.uint32)
        occ_kind = np.concatenate([p.fixed_occupancy_kind, p.member_kind[members]])
        occ_xy = np.concatenate([p.fixed_occupancy_xy.reshape(-1), occ_xy.reshape(-1)])
        meta = dict(p.meta, name=f"eco{seed}", eco_seed=seed, movable_units=int(keep.size))
        return dataclasses.replace(p, unit_fixed=unit_fixed, unit_initial_xy=init.reshape(-1),
                                   unit_bounds=bounds.reshape(-1), cell_fixed=cell_fixed,
                                   cell_fixed_xy=cell_xy.reshape(-1), fixed_occupancy_kind=occ_kind,
                                   fixed_occupancy_xy=occ_xy, meta=meta)
