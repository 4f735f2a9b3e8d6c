"""Continuous-batching server (texo_gp_server_*): a job's result is
bit-identical to the same job placed alone, whatever else shares its rounds
(other grid sizes, other transform precisions, jobs that join mid-run)."""
import time

import numpy as np
import pytest

import texo_amd
from texo_amd import abi, synth
from tests import problems

pytestmark = pytest.mark.gpu

ITERS = 40


def mix():
    return [
        (synth.plain_1k(3), abi.DCT_FP32),
        (synth.serving_job(11, 2000, 4), abi.DCT_FP32),
        (synth.serving_job(12, 6000, 2), abi.DCT_BF16X3),
        (synth.serving_job(13, 2000, 1), abi.DCT_BF16X3),
    ]


def alone(problem, precision, fixed=ITERS):
    return texo_amd.place(problem, dct_precision=precision, fixed_iterations=fixed)


def same(a, b):
    (ta, ra), (tb, rb) = a, b
    assert np.array_equal(ta, tb)
    for k in ("iterations", "evaluations", "line_search_trials", "area_rounds", "final_hpwl",
              "final_wirelength", "field_overflow", "field_energy"):
        assert ra[k] == rb[k], k


def test_batch_is_bit_identical_to_alone():
    jobs = mix()
    print("serving mix alone", [p.meta["name"] for p, _ in jobs], flush=True)
    ref = [alone(p, prec) for p, prec in jobs]
    with texo_amd.Server() as srv:
        ids = [srv.submit(p, abi.options(dct_precision=prec, fixed_iterations=ITERS)) for p, prec in jobs]
        got = [srv.wait(i) for i in ids]
        st = srv.stats()
    print("stats", st, flush=True)
    for a, b in zip(ref, got):
        same(a, b)
    assert st["completed"] == len(jobs) and st["failed"] == 0 and st["active"] == 0
    # lockstep: the batch needs about as many rounds as its longest job, not the sum
    assert st["rounds"] < sum(r["evaluations"] for _, r in ref)


def test_job_joining_mid_run_is_bit_identical():
    (p0, prec0), (p1, prec1) = mix()[1], mix()[0]
    ref0, ref1 = alone(p0, prec0, 60), alone(p1, prec1)
    with texo_amd.Server() as srv:
        a = srv.submit(p0, abi.options(dct_precision=prec0, fixed_iterations=60))
        deadline = time.time() + 60
        while srv.stats()["rounds"] < 20 and time.time() < deadline:
            time.sleep(0.001)
        b = srv.submit(p1, abi.options(dct_precision=prec1, fixed_iterations=ITERS))
        same(ref1, srv.wait(b))
        same(ref0, srv.wait(a))


def test_full_placement_with_stop_rules_in_a_batch():
    p = synth.plain_1k(5)
    ref = texo_amd.place(p)
    with texo_amd.Server(max_batch=2) as srv:
        ids = [srv.submit(p), srv.submit(synth.plain_1k(6)), srv.submit(p)]
        got = [srv.wait(i) for i in ids]
    same(ref, got[0])
    same(ref, got[2])
    assert got[0][1]["field_overflow"][0] <= 0.10


def test_rejected_problem_fails_at_submit_and_poll_releases_results():
    W, H, cap = problems.mixed_device_capacity()
    bad = problems.density_problem(W, H, cap, [(0.0, 0.0, [(abi.KIND_MEMORY, 0.0, 0.0)])])
    with texo_amd.Server() as srv:
        with pytest.raises(abi.PlacementError) as e:
            srv.submit(bad)
        assert e.value.name == "InsufficientCapacity"
        job = srv.submit(synth.plain_1k(3), abi.options(fixed_iterations=5))
        res = None
        deadline = time.time() + 60
        while res is None and time.time() < deadline:
            res = srv.poll(job)
            time.sleep(0.001)
        assert res is not None and res[1]["iterations"] == 5 and res[1]["elapsed_seconds"] > 0.0
        rc = srv.lib.texo_gp_server_poll(srv.handle, job, None, None)
        assert abi.STATUS[rc] == "InvalidArgument"  # released by the first poll


def test_lockstep_batch_iterates_bit_identically_to_sessions():
    jobs = mix()
    with texo_amd.Batch() as batch:
        for p, prec in jobs:
            batch.add(p, abi.options(dct_precision=prec))
        batch.iterate(25)
        batch.iterate(15)
        got = [batch.coordinates(i) for i in range(len(jobs))]
        reps = [batch.result(i, targets=False)[1] for i in range(len(jobs))]
    for (p, prec), x, rep in zip(jobs, got, reps):
        s = texo_amd.Session(p, abi.options(dct_precision=prec))
        r = s.iterate(40)
        assert np.array_equal(s.coordinates(), x)
        assert r.evaluations == rep["evaluations"] and r.final_hpwl == rep["final_hpwl"]
        s.close()


def test_lockstep_batch_run_matches_place():
    jobs = [synth.plain_1k(7), synth.plain_1k(8)]
    ref = [texo_amd.place(p) for p in jobs]
    with texo_amd.Batch() as batch:
        for p in jobs:
            batch.add(p)
        batch.run()
        got = [batch.result(i) for i in range(len(jobs))]
    for a, b in zip(ref, got):
        same(a, b)


def test_multistart_jobs_share_their_design_and_match_alone():
    # starts of one design read one shared copy of its design block (from
    # the second start on); each start, and a different design beside them,
    # iterates exactly as it does alone
    base = synth.serving_job(31, 5000, 4)
    jobs = [(synth.multistart(base, s), abi.DCT_FP16_SCALED) for s in range(3)]
    jobs.insert(1, (synth.serving_job(32, 3000, 2), abi.DCT_BF16X3))
    print("multi-start batch", [p.meta.get("start") for p, _ in jobs], flush=True)
    with texo_amd.Batch() as batch:
        for p, prec in jobs:
            batch.add(p, abi.options(dct_precision=prec))
        batch.iterate(30)
        got = [batch.coordinates(i) for i in range(len(jobs))]
        reps = [batch.result(i, targets=False)[1] for i in range(len(jobs))]
    for (p, prec), x, rep in zip(jobs, got, reps):
        s = texo_amd.Session(p, abi.options(dct_precision=prec))
        r = s.iterate(30)
        assert np.array_equal(s.coordinates(), x)
        assert r.evaluations == rep["evaluations"] and r.final_hpwl == rep["final_hpwl"]
        s.close()


def test_second_start_joining_a_running_design_matches_alone():
    # the second start arrives while the first runs and promotes the first
    # one's design block to a shared copy mid-run; the third arrives after
    # the first two finished (a new owner)
    base = synth.serving_job(33, 4000, 4)
    a, b, c = (synth.multistart(base, s) for s in range(3))
    ref = [alone(p, abi.DCT_FP16_SCALED, 50) for p in (a, b, c)]
    opts = abi.options(dct_precision=abi.DCT_FP16_SCALED, fixed_iterations=50)
    with texo_amd.Server() as srv:
        ia = srv.submit(a, opts)
        deadline = time.time() + 60
        while srv.stats()["rounds"] < 15 and time.time() < deadline:
            time.sleep(0.001)
        ib = srv.submit(b, opts)
        same(ref[0], srv.wait(ia))
        same(ref[1], srv.wait(ib))
        ic = srv.submit(c, opts)
        same(ref[2], srv.wait(ic))
        st = srv.stats()
    assert st["completed"] == 3 and st["failed"] == 0


def test_wide_transform_batch_is_bit_identical_to_single_jobs():
    # 16 starts of the 84k-cell / 256^2 design: each launch of a job group
    # (8 starts) has at least 256 wide transform tiles, so the batch runs the
    # wide FP16-scaled kernel where a single job runs the 32x32 one; the wide
    # kernel keeps the 32x32 kernel's summation order, so every start matches
    # its single-job run bit for bit
    base = synth.ecp5_85k()
    starts = [synth.multistart(base, s) for s in range(16)]
    opts = abi.options(dct_precision=abi.DCT_FP16_SCALED)
    print("wide transform batch of 16 starts", flush=True)
    with texo_amd.Batch() as batch:
        for p in starts:
            batch.add(p, opts)
        batch.iterate(4)
        reps = [batch.result(i, targets=False)[1] for i in range(len(starts))]
    for k in (0, 7, 15):
        s = texo_amd.Session(starts[k], opts)
        r = s.iterate(4).as_dict()
        s.close()
        print(k, "batch", reps[k]["final_hpwl"], reps[k]["field_energy"][:2], "alone", r["final_hpwl"],
              r["field_energy"][:2], flush=True)
        for key in ("iterations", "evaluations", "line_search_trials", "final_hpwl", "final_wirelength",
                    "field_overflow", "field_energy"):
            assert reps[k][key] == r[key], key
