"""Rounds replayed from captured HIP graphs (TEXO_GP_GRAPHS=1) give the same
bits as rounds issued launch by launch (the default): single jobs, a lockstep
batch and the server."""
import os

import numpy as np
import pytest

import texo_amd
from texo_amd import abi, synth

pytestmark = pytest.mark.gpu


def with_graphs(enabled, fn):
    old = os.environ.get("TEXO_GP_GRAPHS")
    os.environ["TEXO_GP_GRAPHS"] = "1" if enabled else "0"
    try:
        return fn()
    finally:
        if old is None:
            del os.environ["TEXO_GP_GRAPHS"]
        else:
            os.environ["TEXO_GP_GRAPHS"] = old


@pytest.mark.parametrize("name,precision", [("plain_1k", abi.DCT_FP32), ("ecp5_25k", abi.DCT_FP16_SCALED)])
def test_graph_rounds_are_bit_identical(name, precision):
    p = synth.CONFIGS[name]()

    def run():
        s = texo_amd.Session(p, abi.options(dct_precision=precision))
        r = s.iterate(40).as_dict()
        x = s.coordinates()
        s.close()
        return r, x

    ra, xa = with_graphs(True, run)
    rb, xb = with_graphs(False, run)
    assert np.array_equal(xa, xb)
    for k in ("iterations", "evaluations", "line_search_trials", "final_hpwl", "final_wirelength", "field_overflow"):
        assert ra[k] == rb[k], k


def test_graph_rounds_in_a_batch_and_the_server():
    base = synth.ecp5_25k()
    starts = [synth.multistart(base, s) for s in range(4)]

    def batch():
        b = texo_amd.Batch()
        for q in starts:
            b.add(q, abi.options(dct_precision=abi.DCT_BF16X3))
        b.iterate(25)
        out = [b.coordinates(i) for i in range(len(starts))]
        b.close()
        return out

    ga, gb = with_graphs(True, batch), with_graphs(False, batch)
    assert all(np.array_equal(a, b) for a, b in zip(ga, gb))

    def serve():
        with texo_amd.Server(max_batch=1) as srv:
            t, rep = srv.wait(srv.submit(synth.plain_1k(), abi.options(fixed_iterations=60)))
            return t, rep, srv.stats()["graph_rounds"]

    ta, rpa, na = with_graphs(True, serve)
    tb, rpb, nb = with_graphs(False, serve)
    print(f"server rounds replayed from graphs: {na} (direct: {nb})", flush=True)
    assert np.array_equal(ta, tb) and rpa["final_hpwl"] == rpb["final_hpwl"]
    assert na > 0 and nb == 0
