"""HIP path vs the FP64 oracle at each seam of the optimizer, through the C
ABI: the combined objective and its gradient (eplace.rs:1022-1044), the
diagonal preconditioner (eplace.rs:1046-1100), the rounded checkpoint targets
(eplace.rs:954-971, lib.rs:4566-4568), the Nesterov step counts
(global_placement.rs:121-335) and the multiplier / gamma / momentum schedule
(eplace.rs:1126-1439) over the first iterations.

Tolerances: integer and index work bit-exact; FP64 host and elementwise
arithmetic in the reference's order bit-exact; anything downstream of the
Poisson transform within its field error (FP32 transform: relative 1e-5).
"""
import ctypes as C

import numpy as np
import pytest

import oracle
import texo_amd
from texo_amd import abi, synth

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def pair(problem, precision=abi.DCT_FP32, **kw):
    g = texo_amd.Session(problem, abi.options(dct_precision=precision, **kw))
    o = abi.Session(oracle.load(), "oracle", problem, abi.options(**kw))
    return g, o


@pytest.mark.parametrize("name", ["plain_1k", "ecp5_25k"])
@pytest.mark.parametrize("precision,tol", [(abi.DCT_FP32, 1e-5), (abi.DCT_BF16X3, 1e-3),
                                           (abi.DCT_FP16_SCALED, 1e-5)])
def test_combined_objective_matches_oracle(name, precision, tol):
    print(f"combined objective {name} precision={precision}", flush=True)
    g, o = pair(synth.CONFIGS[name](), precision)
    rng = np.random.default_rng(11)
    x = np.clip(g.initial + rng.normal(0, 2.0, g.initial.size), g.lower, g.upper)
    e = o.evaluate(x, 6.0)
    scales = np.maximum(e["energy"], 1.0)
    for mult in (np.full(g.n_fields, 0.0), np.full(g.n_fields, 2.5e-3), np.linspace(1e-3, 4e-3, g.n_fields)):
        vg, gg = g.combined(x, 6.0, mult, scales)
        vo, go = o.combined(x, 6.0, mult, scales)
        print(f"  multipliers {mult}: value rel {abs(vg - vo) / abs(vo):.3e} gradient rel {rel(gg, go):.3e}",
              flush=True)
        assert abs(vg - vo) <= tol * abs(vo)
        assert rel(gg, go) <= tol
        if not mult.any():
            # with zero multipliers the objective is the FP64 wirelength alone
            assert abs(vg - vo) <= 1e-12 * abs(vo) and rel(gg, go) <= 1e-12
    g.close()
    o.close()


@pytest.mark.parametrize("name", ["plain_1k", "ecp5_25k", "ecp5_85k"])
def test_preconditioner_matches_oracle_bit_for_bit(name):
    g, o = pair(synth.CONFIGS[name]())
    for mult in (np.zeros(g.n_fields), np.full(g.n_fields, 3.7e-3), np.linspace(0.5, 2.0, g.n_fields)):
        dg = g.preconditioner(mult)
        do = o.preconditioner(mult)
        assert np.array_equal(dg, do), (name, mult, np.abs(dg - do).max())
        assert dg.min() >= 1.0  # the floor of eplace.rs:1097
    with pytest.raises(abi.PlacementError) as err:
        g.preconditioner(np.full(g.n_fields, -1.0))
    assert err.value.name == "InvalidNormalization"
    g.close()
    o.close()


def test_round_targets_match_rounded_coordinate_bit_for_bit():
    p = synth.plain_1k()
    g = texo_amd.Session(p)
    lib = oracle.load()
    rng = np.random.default_rng(5)
    U, W, H = p.n_units, p.width, p.height
    x = rng.uniform(-3.0, max(W, H) + 3.0, 2 * U)
    # ties at .5 (half away from zero), the clamp edges, and just beside them
    special = np.array([-0.5, -0.49, 0.0, 0.49999999999999994, 0.5, 1.5, 2.5, W - 1.5, W - 0.5, W - 0.51, 1e9, -1e9])
    x[:2 * special.size] = np.repeat(special, 2)
    xs = np.concatenate([x, np.zeros(g.n_coordinates - 2 * U)])
    t = g.round_targets(xs)
    ref = np.array([lib.oracle_rounded_coordinate(float(v), H if i & 1 else W) for i, v in enumerate(x)],
                   np.uint32).reshape(-1, 2)
    assert np.array_equal(t, ref)
    g.close()


def test_gamma_weights_match_oracle():
    for name in ("plain_1k", "ecp5_85k"):
        g, o = pair(synth.CONFIGS[name]())
        assert g.schedule()["gamma_weight"] == o.schedule()["gamma_weight"]
        g.close()
        o.close()


@pytest.mark.parametrize("name,iterations", [("plain_1k", 10), ("ecp5_25k", 10)])
def test_nesterov_counts_and_schedule_match_oracle(name, iterations):
    # the FP32 transform's field error (~1e-6) must not flip a line-search
    # decision over the first iterations: iteration, evaluation and trial
    # counts equal the oracle's; momentum follows from the accept/restart
    # sequence alone (bit-equal); multipliers, E0, gamma and the Lipschitz
    # estimate carry the field error
    cfg = {"plain_1k": {"gamma_base": 4.0}}.get(name, {})
    g, o = pair(synth.CONFIGS[name](), **cfg)
    for step in (1, iterations - 1):
        rg, ro = g.iterate(step), o.iterate(step)
        print("gpu", rg.as_dict(), "oracle", ro.as_dict(), flush=True)
        assert (rg.iterations, rg.evaluations, rg.line_search_trials) == \
            (ro.iterations, ro.evaluations, ro.line_search_trials)
        sg, so = g.schedule(), o.schedule()
        print("gpu", sg, "oracle", so, flush=True)
        assert sg["momentum"] == so["momentum"]
        assert sg["gamma_weight"] == so["gamma_weight"]
        for k in ("multiplier", "initial_energy"):
            assert np.allclose(sg[k], so[k], rtol=1e-5, atol=0), k
        for k in ("multiplier_step", "gamma", "lipschitz"):
            assert abs(sg[k] - so[k]) <= 1e-5 * abs(so[k]), k
    assert abs(rg.final_hpwl - ro.final_hpwl) <= 1e-5 * ro.final_hpwl
    g.close()
    o.close()


@pytest.mark.parametrize("name", ["plain_1k", "ecp5_25k", "ecp5_85k"])
def test_reported_hpwl_matches_oracle_at_the_same_coordinates(name):
    # The report's final HPWL is an FP64 reduction on the device over the
    # job's nets at its current coordinates; the oracle's continuous_hpwl
    # (pin coordinate map of analytical_placement.rs:589-597) sums the same
    # nets sequentially: equal within the summation-order rounding.
    print(f"reported hpwl {name}", flush=True)
    problem = synth.CONFIGS[name]()
    with texo_amd.Batch() as b:
        b.add(problem, abi.options(dct_precision=abi.DCT_FP32))
        b.iterate(3)
        _, rep = b.result(0, targets=False)
        x = b.coordinates(0)
    olib = oracle.load()
    o = abi.Session(olib, "oracle", problem)
    ho = olib.oracle_session_hpwl(o.handle, x.ctypes.data_as(C.POINTER(C.c_double)))
    o.close()
    print(f"  device {rep['final_hpwl']!r} oracle {ho!r}", flush=True)
    assert ho > 0.0
    assert abs(rep["final_hpwl"] - ho) <= 1e-12 * ho
