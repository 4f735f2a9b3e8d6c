"""The origin evaluation of an iteration with momentum evaluates the value
alone (global_placement.rs:143-182 reads the origin's gradient only when y
is the origin or after a restart).  The three engine modes, full origin
evaluations (TEXO_GP_ORIGIN_VALUE_ONLY=0), value only (default) and value
only with every gradient check taken by a separate full evaluation (=2),
must give the same run bit for bit: targets, counts, final wirelength and
HPWL.  The 1,024-cell configuration at 1,000 fixed iterations restarts the
momentum (the oracle restarts 9 times over that length), so the restart
path's full origin evaluation is exercised too."""

import os

import numpy as np
import pytest

import texo_amd
from texo_amd import abi, synth

pytestmark = pytest.mark.gpu

MODES = ("0", "1", "2")


def _run_modes(fn):
    out = {}
    old = os.environ.get("TEXO_GP_ORIGIN_VALUE_ONLY")
    try:
        for m in MODES:
            os.environ["TEXO_GP_ORIGIN_VALUE_ONLY"] = m
            out[m] = fn()
    finally:
        if old is None:
            os.environ.pop("TEXO_GP_ORIGIN_VALUE_ONLY", None)
        else:
            os.environ["TEXO_GP_ORIGIN_VALUE_ONLY"] = old
    return out


KEYS = ("iterations", "evaluations", "line_search_trials", "area_rounds", "final_wirelength", "final_hpwl")


@pytest.mark.parametrize("precision", [abi.DCT_FP32, abi.DCT_FP16_SCALED])
@pytest.mark.parametrize("fixed", [0, 1000])
def test_value_only_origin_runs_are_bit_identical(precision, fixed):
    p = synth.plain_1k()
    runs = _run_modes(lambda: texo_amd.place(p, dct_precision=precision, fixed_iterations=fixed))
    t0, r0 = runs["0"]
    print({m: {k: r[k] for k in KEYS[:3]} for m, (_, r) in runs.items()}, flush=True)
    for m in MODES[1:]:
        t, r = runs[m]
        assert np.array_equal(t, t0), m
        for k in KEYS:
            assert r[k] == r0[k], (m, k, r[k], r0[k])


def test_value_only_origin_batch_is_bit_identical():
    base = synth.ecp5_25k()
    starts = [synth.multistart(base, s) for s in range(3)]

    def run():
        with texo_amd.Batch() as b:
            ids = [b.add(q, abi.options(dct_precision=abi.DCT_FP16_SCALED, fixed_iterations=60)) for q in starts]
            b.run()
            return [b.result(i) for i in ids]

    runs = _run_modes(run)
    for m in MODES[1:]:
        for (t, r), (t0, r0) in zip(runs[m], runs["0"]):
            assert np.array_equal(t, t0), m
            for k in KEYS:
                assert r[k] == r0[k], (m, k)


def _run_skip(fn):
    out = {}
    old = os.environ.get("TEXO_GP_SKIP_REJECTED")
    try:
        for m in ("0", "1"):
            os.environ["TEXO_GP_SKIP_REJECTED"] = m
            out[m] = fn()
    finally:
        if old is None:
            os.environ.pop("TEXO_GP_SKIP_REJECTED", None)
        else:
            os.environ["TEXO_GP_SKIP_REJECTED"] = old
    return out


@pytest.mark.parametrize("precision", [abi.DCT_FP32, abi.DCT_FP16_SCALED, abi.DCT_BF16X3])
@pytest.mark.parametrize("fixed", [0, 1000])
def test_skipping_rejected_trial_gradients_is_bit_identical(precision, fixed):
    """A rejected line-search trial needs only its value and the reference's
    finite-gradient check; its gradient blocks are skipped when every entry
    is bounded.  Runs with and without the skip must agree bit for bit."""
    p = synth.plain_1k()
    runs = _run_skip(lambda: texo_amd.place(p, dct_precision=precision, fixed_iterations=fixed))
    (t0, r0), (t1, r1) = runs["0"], runs["1"]
    assert np.array_equal(t1, t0)
    for k in KEYS:
        assert r1[k] == r0[k], (k, r1[k], r0[k])


def test_skipping_rejected_trial_gradients_batch_is_bit_identical():
    base = synth.ecp5_25k()
    starts = [synth.multistart(base, s) for s in range(3)]

    def run():
        with texo_amd.Batch() as b:
            ids = [b.add(q, abi.options(dct_precision=abi.DCT_FP16_SCALED, fixed_iterations=60)) for q in starts]
            b.run()
            return [b.result(i) for i in ids]

    runs = _run_skip(run)
    for (t, r), (t0, r0) in zip(runs["1"], runs["0"]):
        assert np.array_equal(t, t0)
        for k in KEYS:
            assert r[k] == r0[k], k
