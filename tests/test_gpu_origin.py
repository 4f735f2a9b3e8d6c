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


def _timed_tasks(kernel, env):
    """Evaluations per timed launch of `kernel` over 30 iterations of a
    3-start batch of the 24k-cell design, with the engine switch `env`."""
    base = synth.ecp5_25k()
    starts = [synth.multistart(base, s) for s in range(3)]
    key, value = env
    old = os.environ.get(key)
    os.environ[key] = value
    try:
        with texo_amd.Batch() as b:
            for q in starts:
                b.add(q, abi.options(dct_precision=abi.DCT_FP16_SCALED))
            b.iterate(5)
            b.set_kernel_timing(kernel)
            b.iterate(30)
            _, launches, tasks = b.kernel_timing()
            b.set_kernel_timing(None)
    finally:
        if old is None:
            os.environ.pop(key, None)
        else:
            os.environ[key] = old
    assert launches > 0
    return tasks


def test_rejected_trials_skip_their_gradient_blocks():
    """The gradient launches cover fewer evaluations with the skip on: the
    first trial of most iterations is rejected, and its blocks do not run."""
    on = _timed_tasks("gather", ("TEXO_GP_SKIP_REJECTED", "1"))
    off = _timed_tasks("gather", ("TEXO_GP_SKIP_REJECTED", "0"))
    print("gradient evaluations per launch: skip", on, "no skip", off, flush=True)
    assert on < 0.9 * off


def test_value_only_origins_count_as_partial_wa_evaluations():
    """A value-only origin evaluation stores no derivatives: the WA timer
    counts it as U / (U + movable pins) of a full evaluation."""
    vo = _timed_tasks("wa", ("TEXO_GP_ORIGIN_VALUE_ONLY", "1"))
    full = _timed_tasks("wa", ("TEXO_GP_ORIGIN_VALUE_ONLY", "0"))
    print("WA evaluations per launch: value-only", vo, "full", full, flush=True)
    assert vo < 0.95 * full
