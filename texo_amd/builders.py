"""Python mirror of the GP input builders and the provenance record
(include/texo_gp.h, "GP input builders"): numpy in, numpy out, over the C ABI
of libtexo_gp.so.  Follows crates/texo-pnr/src/routing_demand.rs:185-313,
crates/texo-flow/src/lib.rs:3185-3201 and :3355-3433, and
crates/texo-cli/src/checkpoint.rs:87-100."""
from __future__ import annotations

import ctypes as C
import json

import numpy as np

from . import abi, load

ALGORITHM = {"imported": 0, "routability_electrostatic": 1, "timing_routability_electrostatic": 2}


def declare(lib, prefix: str = "texo_gp"):
    P, u8, u16, u32, u64, d = C.POINTER, C.c_uint8, C.c_uint16, C.c_uint32, C.c_uint64, C.c_double
    dev = [C.c_int32] if prefix == "texo_gp" else []
    lib[f"{prefix}_channel_orientation"].argtypes = [C.c_char_p]
    lib[f"{prefix}_channel_orientation"].restype = C.c_int
    f = lib[f"{prefix}_routing_capacity_map"]
    f.argtypes = dev + [u32, u32, u32, P(u8), P(u32), P(u16), u32, P(u32), P(u32), P(u8), P(u32), P(u32)]
    f.restype = C.c_int
    f = lib[f"{prefix}_routing_demand_map"]
    f.argtypes = dev + [u32, u32, u32, P(u32), P(u32), P(u8), P(d), P(d), P(u64), P(u64)]
    f.restype = C.c_int
    f = lib[f"{prefix}_timing_placement_weights"]
    f.argtypes = [u64, P(u32), P(u32), P(u64), P(u64), u32, P(u32), P(u32), P(u64)]
    f.restype = u64


class _Named:
    """Item access to a CDLL's functions by name."""

    def __init__(self, lib):
        self.lib = lib

    def __getitem__(self, name):
        return getattr(self.lib, name)


def _lib():
    lib = load()
    if not getattr(lib, "_builders_declared", False):
        declare(_Named(lib))
        lib.texo_gp_placement_model.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(abi.texo_gp_options), C.c_char_p,
                                                C.c_size_t]
        lib.texo_gp_placement_model.restype = C.c_int
        lib._builders_declared = True
    return lib


def _check(rc: int):
    if rc != 0:
        raise abi.PlacementError.from_status(rc)


def channel_orientation(name: str) -> int:
    return _lib().texo_gp_channel_orientation(name.encode())


def routing_capacity_map(width, height, wire_orientation, wire_xy, wire_capacity, pip_from, pip_to,
                         pip_blocked=None, device: int = 0, lib=None, prefix: str = "texo_gp"):
    """(horizontal, vertical) uint32[height, width]."""
    lib = lib or _lib()
    wo = np.ascontiguousarray(wire_orientation, np.uint8)
    wxy = np.ascontiguousarray(wire_xy, np.uint32).reshape(-1)
    wc = np.ascontiguousarray(wire_capacity, np.uint16)
    pf, pt = np.ascontiguousarray(pip_from, np.uint32), np.ascontiguousarray(pip_to, np.uint32)
    pb = None if pip_blocked is None else np.ascontiguousarray(pip_blocked, np.uint8)
    h = np.zeros(width * height, np.uint32)
    v = np.zeros(width * height, np.uint32)
    args = [width, height, wo.size, abi.ptr(wo, C.c_uint8), abi.ptr(wxy, C.c_uint32), abi.ptr(wc, C.c_uint16),
            pf.size, abi.ptr(pf, C.c_uint32), abi.ptr(pt, C.c_uint32), abi.ptr(pb, C.c_uint8),
            abi.ptr(h, C.c_uint32), abi.ptr(v, C.c_uint32)]
    fn = getattr(lib, f"{prefix}_routing_capacity_map")
    _check(fn(device, *args) if prefix == "texo_gp" else fn(*args))
    return h.reshape(height, width), v.reshape(height, width)


def routing_demand_map(width, height, net_pin_start, pin_xy, net_clock, device: int = 0, lib=None,
                       prefix: str = "texo_gp"):
    """(horizontal, vertical float64[height, width], included_nets, excluded_clock_nets)."""
    lib = lib or _lib()
    ns = np.ascontiguousarray(net_pin_start, np.uint32)
    pxy = np.ascontiguousarray(pin_xy, np.uint32).reshape(-1)
    nc = np.ascontiguousarray(net_clock, np.uint8)
    h = np.zeros(width * height)
    v = np.zeros(width * height)
    inc, exc = C.c_uint64(), C.c_uint64()
    args = [width, height, ns.size - 1, abi.ptr(ns, C.c_uint32), abi.ptr(pxy, C.c_uint32), abi.ptr(nc, C.c_uint8),
            abi.ptr(h, C.c_double), abi.ptr(v, C.c_double), C.byref(inc), C.byref(exc)]
    fn = getattr(lib, f"{prefix}_routing_demand_map")
    _check(fn(device, *args) if prefix == "texo_gp" else fn(*args))
    return h.reshape(height, width), v.reshape(height, width), inc.value, exc.value


def _u128_words(values):
    out = np.zeros(2 * len(values), np.uint64)
    for i, x in enumerate(values):
        out[2 * i] = x & 0xFFFFFFFFFFFFFFFF
        out[2 * i + 1] = x >> 64
    return out


def timing_placement_weights(edges, exponent: int, lib=None, prefix: str = "texo_gp"):
    """edges: iterable of (net, sink, path_delay_ps, worst_path_delay_ps) with
    Python ints (up to 128 bits).  Returns {(net, sink): weight}."""
    lib = lib or _lib()
    edges = list(edges)
    n = len(edges)
    net = np.array([e[0] for e in edges], np.uint32)
    sink = np.array([e[1] for e in edges], np.uint32)
    path, worst = _u128_words([e[2] for e in edges]), _u128_words([e[3] for e in edges])
    on, os_, ow = np.zeros(max(n, 1), np.uint32), np.zeros(max(n, 1), np.uint32), np.zeros(max(n, 1), np.uint64)
    k = getattr(lib, f"{prefix}_timing_placement_weights")(
        n, abi.ptr(net, C.c_uint32), abi.ptr(sink, C.c_uint32), abi.ptr(path, C.c_uint64), abi.ptr(worst, C.c_uint64),
        exponent, abi.ptr(on, C.c_uint32), abi.ptr(os_, C.c_uint32), abi.ptr(ow, C.c_uint64))
    return {(int(on[i]), int(os_[i])): int(ow[i]) for i in range(k)}


def placement_model(algorithm: str, weight_exponent: int, options=None) -> dict:
    """The checkpoint placement-model record with the engine's provenance."""
    lib = _lib()
    opts = options if options is not None else abi.options()
    need = -lib.texo_gp_placement_model(ALGORITHM[algorithm], weight_exponent, C.byref(opts), None, 0)
    if need <= 0:
        raise abi.PlacementError.from_status(-need)
    buf = C.create_string_buffer(need)
    _check(lib.texo_gp_placement_model(ALGORITHM[algorithm], weight_exponent, C.byref(opts), buf, need))
    return json.loads(buf.value.decode())
