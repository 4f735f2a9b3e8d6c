"""Python mirror of the legalization auction (include/texo_gp.h,
"legalization auction"): the relaxed unit-to-object assignment of
legalization::project (assign_implicitly,
crates/texo-pnr/src/legalization/auction.rs:123-253) over explicit candidate
domains, run on the GPU through libtexo_gp.so."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi, load


def declare(lib, prefix: str = "texo_gp"):
    """Declares the auction entry point; `prefix` "oracle" declares the CPU
    restatement's (the reference's sequential queue), which takes no device."""
    P, u32, u64, i64 = C.POINTER, C.c_uint32, C.c_uint64, C.c_int64
    dev = [C.c_int32] if prefix == "texo_gp" else []
    f = getattr(lib, f"{prefix}_auction")
    f.argtypes = dev + [u64, u64, u64, P(u32), P(u32), P(u64), P(i64), P(u64)]
    f.restype = C.c_int


def _lib():
    lib = load()
    if not getattr(lib, "_auction_declared", False):
        declare(lib)
        lib._auction_declared = True
    return lib


def domains_csr(domains) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(start, object, base cost) of per-bidder lists of (object, base cost)."""
    start = np.zeros(len(domains) + 1, np.uint32)
    start[1:] = np.cumsum([len(d) for d in domains])
    obj = np.array([o for d in domains for o, _ in d], np.uint32)
    cost = np.array([c for d in domains for _, c in d], np.uint64)
    return start, obj, cost


def auction(start, obj, cost, n_objects: int, max_base_cost: int, device: int = 0):
    """(assignment: int64[n_bidders], object or -1; stats: bids that took an
    object, candidates examined, unmatched, rounds).  Raises PlacementError
    with the reference's AuctionError name on malformed domains or range
    exhaustion."""
    lib = _lib()
    start = np.ascontiguousarray(start, np.uint32)
    n = start.size - 1
    obj = np.ascontiguousarray(obj if len(obj) else np.zeros(1), np.uint32)
    cost = np.ascontiguousarray(cost if len(cost) else np.zeros(1), np.uint64)
    out = np.zeros(max(n, 1), np.int64)
    stats = np.zeros(4, np.uint64)
    rc = lib.texo_gp_auction(device, n, n_objects, max_base_cost, abi.ptr(start, C.c_uint32),
                             abi.ptr(obj, C.c_uint32), abi.ptr(cost, C.c_uint64), abi.ptr(out, C.c_int64),
                             abi.ptr(stats, C.c_uint64))
    if rc != 0:
        raise abi.PlacementError.from_status(rc)
    return out[:n], stats
