"""Incremental (ECO) requests through the HIP path: a request re-places a
connected subset of a resident design with every other unit fixed at its
site.  It tracks the FP64 oracle, keeps the fixed units on their sites bit for
bit, and gives the same result alone and in a serving mix with full
multi-start jobs of the design and other jobs."""
import ctypes as C
import functools
import time

import numpy as np
import pytest

import oracle
import texo_amd
from texo_amd import abi, synth

pytestmark = pytest.mark.gpu

ITERS = 100


@functools.lru_cache(maxsize=None)
def stream():
    return synth.EcoStream(synth.eco_design(seed=5, cells=20000, side=128, nets=24000, high_fanout=2))


@functools.lru_cache(maxsize=None)
def oracle_result(seed):
    r = stream().request(seed)
    t = np.zeros(2 * r.n_units, np.uint32)
    rep = abi.texo_gp_report()
    rc = oracle.load().oracle_place(C.byref(r.struct()), C.byref(abi.options(fixed_iterations=ITERS)),
                                    abi.ptr(t, C.c_uint32), C.byref(rep))
    assert rc == 0
    return t.reshape(-1, 2), rep.as_dict()


def fixed_on_sites(request, design, targets):
    fixed = request.unit_fixed != 0
    return np.array_equal(targets[fixed], design.unit_initial_xy.reshape(-1, 2)[fixed].astype(np.uint32))


@pytest.mark.parametrize("precision", [abi.DCT_FP32, abi.DCT_BF16X3])
def test_request_tracks_the_oracle(precision):
    s = stream()
    r = s.request(3)
    print(f"eco request 3 ({r.meta['movable_units']} movable units) precision={precision}", flush=True)
    tg, rg = texo_amd.place(r, dct_precision=precision, fixed_iterations=ITERS)
    to, ro = oracle_result(3)
    print("gpu", rg, "oracle", ro, flush=True)
    assert rg["iterations"] == ITERS
    assert fixed_on_sites(r, s.design, tg)
    assert abs(rg["final_hpwl"] - ro["final_hpwl"]) <= 0.005 * ro["final_hpwl"]
    assert max(abs(a - b) for a, b in zip(rg["field_overflow"], ro["field_overflow"])) <= 0.02


def test_requests_in_a_mix_are_bit_identical_to_alone():
    s = stream()
    jobs = [(s.request(4), abi.DCT_BF16X3), (s.request(5), abi.DCT_BF16X3),
            (synth.multistart(s.design, 1), abi.DCT_FP16_SCALED), (s.request(6), abi.DCT_FP32),
            (synth.serving_job(21, 3000, 4), abi.DCT_BF16X3)]
    iters = 40
    print("eco mix alone", [p.meta["name"] for p, _ in jobs], flush=True)
    ref = [texo_amd.place(p, dct_precision=prec, fixed_iterations=iters) for p, prec in jobs]
    with texo_amd.Server() as srv:
        ids = [srv.submit(p, abi.options(dct_precision=prec, fixed_iterations=iters)) for p, prec in jobs]
        got = [srv.wait(i) for i in ids]
    for (ta, ra), (tb, rb) in zip(ref, got):
        assert np.array_equal(ta, tb)
        for k in ("iterations", "evaluations", "line_search_trials", "final_hpwl", "final_wirelength",
                  "field_overflow", "field_energy"):
            assert ra[k] == rb[k], k
    for (p, _), (t, _) in zip(jobs, got):
        if p.meta["name"].startswith("eco"):
            assert fixed_on_sites(p, s.design, t)


@functools.lru_cache(maxsize=None)
def resident():
    # the incremental stream's design: 250k cells on 512x512 bins, 300k nets
    return synth.EcoStream(synth.eco_design())


def test_request_on_the_resident_250k_design():
    # one request of 1-5% of the units for 100 iterations, BF16x3 transform
    d = resident().design
    r = resident().request(11)
    print(f"eco 250k request: {r.meta['movable_units']} movable units, {r.fixed_occupancy_kind.size} fixed sites",
          flush=True)
    t0 = time.perf_counter()
    t, rep = texo_amd.place(r, dct_precision=abi.DCT_BF16X3, fixed_iterations=ITERS)
    print(f"placed in {time.perf_counter() - t0:.2f} s", rep, flush=True)
    assert rep["status"] == "OK" and rep["iterations"] == ITERS
    assert fixed_on_sites(r, d, t)
    keep = r.unit_fixed == 0
    b = r.unit_bounds.reshape(-1, 4)[keep]
    assert np.all(t[keep] >= np.floor(b[:, :2])) and np.all(t[keep] <= np.ceil(b[:, 2:]))
    start = r.unit_initial_xy.reshape(-1, 2)[keep]
    assert t[keep].astype(float).std(axis=0).min() > start.std(axis=0).max()


def test_requests_on_512_bins_are_bit_identical_batched_and_alone():
    # on 512x512 bins a request's transform has 64 wide tiles per field, so
    # a server round over several requests runs the wide split-MFMA kernels
    # (K = 512) where a request alone runs the 32x32 ones in most rounds; both
    # keep one summation order, so the results agree bit for bit
    s = resident()
    jobs = [(s.request(31), abi.DCT_BF16X3), (s.request(32), abi.DCT_BF16X3), (s.request(33), abi.DCT_FP16_SCALED),
            (s.request(34), abi.DCT_BF16X3)]
    iters = 30
    print("eco 512 mix alone", [(p.meta["movable_units"], prec) for p, prec in jobs], flush=True)
    ref = [texo_amd.place(p, dct_precision=prec, fixed_iterations=iters) for p, prec in jobs]
    with texo_amd.Server() as srv:
        ids = [srv.submit(p, abi.options(dct_precision=prec, fixed_iterations=iters)) for p, prec in jobs]
        got = [srv.wait(i) for i in ids]
    for (ta, ra), (tb, rb) in zip(ref, got):
        print(ra["final_hpwl"], rb["final_hpwl"], ra["field_energy"][:2], rb["field_energy"][:2], flush=True)
        assert np.array_equal(ta, tb)
        for k in ("iterations", "evaluations", "line_search_trials", "final_hpwl", "final_wirelength",
                  "field_overflow", "field_energy"):
            assert ra[k] == rb[k], k


@pytest.mark.parametrize("precision", [abi.DCT_FP32, abi.DCT_BF16X3, abi.DCT_FP16_SCALED])
def test_request_field_error_on_512_bins(precision):
    # the incremental stream is specified at <= 1e-3 relative field error:
    # the field at every charge (units and fillers) of a 250k-cell / 512x512
    # request, against the FP64 oracle, at its start and at spread coordinates
    r = resident().request(12)
    print(f"eco 512 field error precision={precision}, {r.meta['movable_units']} movable units", flush=True)
    g = texo_amd.Session(r, abi.options(dct_precision=precision))
    o = abi.Session(oracle.load(), "oracle", r)
    rng = np.random.default_rng(4)
    for x in (g.initial, np.clip(g.initial + rng.normal(0, 3.0, g.initial.size), g.lower, g.upper)):
        eg, eo = g.evaluate(x, 4.0), o.evaluate(x, 4.0)
        a = np.concatenate([eg["unit_gradient"].ravel(), eg["filler_gradient"].ravel()])
        b = np.concatenate([eo["unit_gradient"].ravel(), eo["filler_gradient"].ravel()])
        field = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        energy = float(np.max(np.abs(eg["energy"] - eo["energy"]) / eo["energy"]))
        print(f"  field rel {field:.3e} energy rel {energy:.3e}", flush=True)
        assert field <= 1e-3 and energy <= 1e-3
    g.close()
    o.close()
