"""Small hand-built problems restating the reference's unit-test fixtures
(DensityUnit/Device builders of crates/texo-pnr/src/electrostatic_placement.rs
tests, wide_net of analytical_placement.rs tests) in the flattened ABI form."""
from __future__ import annotations

import numpy as np

from texo_amd import abi

LUT4 = abi.KIND_LUT(4)
FF = abi.KIND_REGISTER


def density_problem(W, H, capacity, units, nets=None, fixed_units=None, charges=None, routing=None,
                    fixed_occupancy=None):
    """units: list of (origin_x, origin_y, [(kind, ox, oy), ...]); each member
    is its own cell.  capacity: {kind: uint32[H*W]}.  nets: list of lists of
    cell ids (driver first); by default one net joins consecutive cells.
    fixed_units: list of (x, y) fixed single-cell Logic units appended after."""
    cell_unit, offsets, member_kind, member_off, member_cell = [], [], [], [], []
    unit_member_start = [0]
    initial, bounds = [], []
    for u, (x, y, members) in enumerate(units):
        for kind, ox, oy in members:
            cid = len(cell_unit)
            cell_unit.append(u)
            offsets.append((ox, oy))
            member_kind.append(kind)
            member_off.append((ox, oy))
            member_cell.append(cid)
        unit_member_start.append(len(member_kind))
        initial.append((x, y))
        bounds.append((0, 0, W - 1, H - 1))
    n_mov = len(units)
    fixed_units = fixed_units or []
    cell_fixed = [0] * len(cell_unit)
    fixed_xy = [(0.0, 0.0)] * len(cell_unit)
    for i, (x, y) in enumerate(fixed_units):
        cell_unit.append(n_mov + i)
        offsets.append((0.0, 0.0))
        cell_fixed.append(1)
        fixed_xy.append((x, y))
        unit_member_start.append(len(member_kind))
        initial.append((x, y))
        bounds.append((x, y, x, y))
    n_units = n_mov + len(fixed_units)
    n_cells = len(cell_unit)
    if nets is None:
        nets = [[c, c + 1] for c in range(n_cells - 1)] or [[0, 0]]
    net_pin_start = np.cumsum([0] + [len(n) for n in nets])
    pin_cell = [c for n in nets for c in n]
    kinds = sorted(capacity)
    cap = np.concatenate([np.asarray(capacity[k], np.uint32).reshape(-1) for k in kinds])
    fo = fixed_occupancy or []
    return abi.Problem(
        width=W, height=H, cell_unit=cell_unit, cell_fixed=cell_fixed, cell_fixed_xy=np.array(fixed_xy, float),
        cell_offset_xy=np.array(offsets, float), unit_fixed=[0] * n_mov + [1] * len(fixed_units),
        unit_initial_xy=np.array(initial, float), unit_bounds=np.array(bounds, float),
        unit_member_start=unit_member_start, member_kind=member_kind,
        member_offset_xy=np.array(member_off, float).reshape(-1, 2), member_cell=member_cell,
        capacity_kind=kinds, capacity=cap, net_pin_start=net_pin_start, pin_cell=pin_cell,
        fixed_occupancy_kind=[k for k, _, _ in fo], fixed_occupancy_xy=[v for _, x, y in fo for v in (x, y)],
        routing_horizontal=None if routing is None else routing[0],
        routing_vertical=None if routing is None else routing[1])


def mixed_device_capacity():
    """4x3 device: LUT4 at y=0, FF at y=1 (electrostatic_placement.rs:1303-1314)."""
    W, H = 4, 3
    lut = np.zeros((H, W), np.uint32)
    ff = np.zeros((H, W), np.uint32)
    lut[0, :] = 1
    ff[1, :] = 1
    return W, H, {LUT4: lut, FF: ff}


def grid_problem(W, H, kinds=(LUT4,)):
    """A problem whose only purpose is to own a W x H Poisson grid with one
    field per kind: one capacity unit per bin, one unit-charge cell per kind."""
    cap = {k: np.ones((H, W), np.uint32) for k in kinds}
    units = [(W / 2.0, H / 2.0, [(k, 0.0, 0.0)]) for k in kinds]
    return density_problem(W, H, cap, units)
