"""The allocation-free routing demand map (texo_gp_routing_demand_map_noalloc)
against the allocating entry point, through the C ABI: two calls in a row,
the second on a smaller input, must give the same bits as the allocating
call; inputs beyond the device scratch are refused with a clear error."""
import ctypes as C

import numpy as np
import pytest

import texo_amd
from texo_amd import abi, builders, synth
from tests.test_gpu_builders import placed_netlist

pytestmark = pytest.mark.gpu


def noalloc(width, height, net_pin_start, pin_xy, net_clock):
    lib = texo_amd.load()
    f = lib.texo_gp_routing_demand_map_noalloc
    f.argtypes = lib.texo_gp_routing_demand_map.argtypes
    f.restype = C.c_int
    return builders.routing_demand_map(width, height, net_pin_start, pin_xy, net_clock, lib=_Alias(lib, f))


class _Alias:
    """A library view whose routing_demand_map entry is the allocation-free one."""

    def __init__(self, lib, fn):
        self.lib, self.fn = lib, fn

    def __getattr__(self, name):
        return self.fn if name == "texo_gp_routing_demand_map" else getattr(self.lib, name)


def test_noalloc_matches_the_allocating_entry_point_on_two_calls():
    rng = np.random.default_rng(21)
    big, small = synth.ecp5_85k(), synth.plain_1k()
    for p in (big, small):  # the second call is smaller: stale sums must not leak in
        ns, pxy, clock = placed_netlist(p, rng)
        print(f"noalloc demand map: {p.n_nets} nets on {p.width}x{p.height}", flush=True)
        a = builders.routing_demand_map(p.width, p.height, ns, pxy, clock)
        b = noalloc(p.width, p.height, ns, pxy, clock)
        assert a[2:] == b[2:]
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_noalloc_refuses_inputs_beyond_its_scratch():
    # 512x512 tiles need 8 MiB; 8M two-pin nets stage 160 MiB of pins
    W = H = 512
    n = 8_000_000
    ns = np.arange(0, 2 * n + 1, 2, dtype=np.uint32)
    pxy = np.zeros((2 * n, 2), np.uint32)
    clock = np.zeros(n, np.uint8)
    with pytest.raises(abi.PlacementError) as err:
        noalloc(W, H, ns, pxy, clock)
    assert err.value.name == "OutOfMemory"
    assert "device scratch" in texo_amd.last_error()
