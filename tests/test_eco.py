"""Incremental (ECO) requests on a resident design (texo_amd.synth.EcoStream):
the request generator keeps the resident placement legal and the re-placed
subset connected, and a request placed by the oracle keeps every other unit
at its site (a single-choice unit is fixed, eplace.rs:263-271)."""
import ctypes as C

import numpy as np

import oracle
from texo_amd import abi, synth


def small_design():
    return synth.eco_design(seed=7, cells=6000, side=64, nets=7200, high_fanout=2)


def occupancy_within_capacity(p):
    G = p.width * p.height
    kind = np.asarray(p.fixed_occupancy_kind)
    xy = p.fixed_occupancy_xy.reshape(-1, 2).astype(np.int64)
    for i, k in enumerate(p.capacity_kind):
        cap = p.capacity[i * G:(i + 1) * G].astype(np.int64)
        sel = kind == k
        if np.any(np.bincount(xy[sel, 1] * p.width + xy[sel, 0], minlength=G) > cap):
            return False
    return True


def connected(p, units):
    """The units form one component of the netlist restricted to them."""
    inside = np.zeros(p.unit_fixed.size, bool)
    inside[units] = True
    pin_unit = p.cell_unit[p.pin_cell].astype(np.int64)
    parent = {int(u): int(u) for u in units}

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a
    starts = p.net_pin_start.astype(np.int64)
    for n in range(starts.size - 1):
        us = [int(u) for u in pin_unit[starts[n]:starts[n + 1]] if inside[u]]
        for u in us[1:]:
            parent[find(u)] = find(us[0])
    return len({find(int(u)) for u in units}) == 1


def test_resident_placement_is_legal():
    d = small_design()
    G = d.width * d.height
    xy = d.unit_initial_xy.reshape(-1, 2)
    assert np.array_equal(xy, np.round(xy))
    counts = np.diff(d.unit_member_start.astype(np.int64))
    owner = np.repeat(np.arange(d.unit_fixed.size), counts)
    movable_members = d.unit_fixed[owner] == 0
    mxy = (xy[owner] + d.member_offset_xy.reshape(-1, 2))[movable_members].astype(np.int64)
    kinds = d.member_kind[movable_members]
    for i, k in enumerate(d.capacity_kind):
        cap = d.capacity[i * G:(i + 1) * G].astype(np.int64)
        sel = kinds == k
        assert np.all(np.bincount(mxy[sel, 1] * d.width + mxy[sel, 0], minlength=G) <= cap)


def test_requests_fix_the_rest_legally_and_keep_the_subset_connected():
    d = small_design()
    s = synth.EcoStream(d)
    movable = int((d.unit_fixed == 0).sum())
    for seed in range(4):
        r = s.request(seed)
        keep = np.flatnonzero((r.unit_fixed == 0) & (d.unit_fixed == 0))
        assert 0.01 * movable - 1 <= keep.size <= 0.05 * movable + 1
        assert connected(d, keep)
        assert occupancy_within_capacity(r)
        fixed = (r.unit_fixed != 0) & (d.unit_fixed == 0)
        b = r.unit_bounds.reshape(-1, 4)
        xy = d.unit_initial_xy.reshape(-1, 2)
        assert np.array_equal(b[fixed, :2], xy[fixed]) and np.array_equal(b[fixed, 2:], xy[fixed])
        cells = np.flatnonzero(fixed[r.cell_unit])
        assert np.all(r.cell_fixed[cells] == 1)
        assert np.array_equal(r.cell_fixed_xy.reshape(-1, 2)[cells],
                              xy[r.cell_unit[cells]] + r.cell_offset_xy.reshape(-1, 2)[cells])


def test_oracle_keeps_fixed_units_on_their_sites():
    d = small_design()
    r = synth.EcoStream(d).request(1)
    t = np.zeros(2 * r.n_units, np.uint32)
    rep = abi.texo_gp_report()
    rc = oracle.load().oracle_place(C.byref(r.struct()), C.byref(abi.options(fixed_iterations=30)),
                                    abi.ptr(t, C.c_uint32), C.byref(rep))
    assert rc == 0 and rep.iterations == 30
    t = t.reshape(-1, 2)
    fixed = r.unit_fixed != 0
    assert np.array_equal(t[fixed], d.unit_initial_xy.reshape(-1, 2)[fixed].astype(np.uint32))
    keep = ~fixed
    start = r.unit_initial_xy.reshape(-1, 2)[keep]
    assert t[keep].astype(float).std(axis=0).min() > start.std(axis=0).max()  # the clump spreads
