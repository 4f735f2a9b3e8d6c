"""Failure paths of the HIP placement path, forced from the host so that
every kernel still completes normally: an injected evaluation failure, the
iteration cap with and without best effort (eplace.rs:989-999), a job over
the server's HBM budget, and cancellation."""
import ctypes as C
import time

import numpy as np
import pytest

import oracle
import texo_amd
from texo_amd import abi, synth

pytestmark = pytest.mark.gpu


def test_injected_evaluation_failure_leaves_the_iterate_unchanged():
    # an evaluator failure inside a step leaves the optimizer state bit for
    # bit as it was (global_placement.rs:875-901): the coordinates after the
    # failure equal those of a clean run stopped before the failing iteration
    p = synth.plain_1k()
    for inject in (9, 16, 23):
        bad = texo_amd.Session(p, abi.options(inject_failure_at_evaluation=inject))
        with pytest.raises(abi.PlacementError) as err:
            bad.iterate(50)
        assert err.value.name == "Injected"
        x_bad = bad.coordinates()
        clean = texo_amd.Session(p)
        # find the iteration the failing evaluation belonged to: the clean
        # run's evaluation count passes `inject` during it
        done, x_done = 0, clean.coordinates()
        while True:
            r = clean.iterate(1)
            if r.evaluations >= inject:
                failed_iteration = r.iterations
                break
            done = r.iterations
            x_done = clean.coordinates()
        print(f"inject {inject}: failed in iteration {failed_iteration}", flush=True)
        assert failed_iteration == done + 1
        assert np.array_equal(x_bad, x_done)
        bad.close()
        clean.close()


def test_iteration_cap_fails_strict_and_keeps_best_effort_checkpoint():
    # stop_without_convergence (eplace.rs:989-999): a strict solve that hits
    # the iteration cap reports DidNotConverge with its overflows; a
    # best-effort solve returns the lowest-overflow checkpoint.  Both agree
    # with the oracle under the same cap.
    lib = oracle.load()
    for best_effort in (False, True):
        p = synth.plain_1k()
        p.best_effort = best_effort
        opts = abi.options(max_iterations=12)
        tg = np.zeros(2 * p.n_units, np.uint32)
        rg = abi.texo_gp_report()
        rc = texo_amd.load().texo_gp_place(C.byref(p.struct()), C.byref(opts), abi.ptr(tg, C.c_uint32),
                                           C.byref(rg))
        to = np.zeros(2 * p.n_units, np.uint32)
        ro = abi.texo_gp_report()
        rco = lib.oracle_place(C.byref(p.struct()), C.byref(opts), abi.ptr(to, C.c_uint32), C.byref(ro))
        print(best_effort, rc, rg.as_dict(), rco, ro.as_dict(), flush=True)
        assert rc == rco
        if best_effort:
            assert rc == 0 and rg.iterations == ro.iterations == 12
            # the checkpoint's rounded targets carry the field error only at
            # rounding ties
            assert np.mean(tg == to) >= 0.99 and np.abs(tg.astype(int) - to.astype(int)).max() <= 1
        else:
            assert abi.STATUS[rc] == "DidNotConverge"
            assert rg.err_index == ro.err_index == 12
            assert abs(rg.err_value[0] - ro.err_value[0]) <= 1e-4


def test_job_over_the_hbm_budget_is_rejected():
    p = synth.plain_1k()
    with texo_amd.Server(hbm_budget_bytes=1 << 20) as srv:
        with pytest.raises(abi.PlacementError) as err:
            srv.submit(p)
        assert err.value.name == "OutOfMemory"
    with texo_amd.Server(hbm_budget_bytes=4 << 30) as srv:
        t, rep = srv.wait(srv.submit(p, abi.options(fixed_iterations=5)))
        assert rep["status"] == "OK"


def test_cancel_pending_and_running_jobs():
    long_job = synth.ecp5_25k()
    small = synth.plain_1k()
    with texo_amd.Server(max_batch=1) as srv:
        running = srv.submit(long_job, abi.options(fixed_iterations=100000))
        pending = srv.submit(long_job, abi.options(fixed_iterations=100000))
        assert srv.cancel(pending) == 0
        _, rep = srv.wait(pending, raise_on_error=False)
        assert rep["status"] == "Cancelled"
        # let the running job take a few rounds, then stop it
        for _ in range(200):
            if srv.stats()["rounds"] >= 5:
                break
            time.sleep(0.01)
        assert srv.cancel(running) == 0
        _, rep = srv.wait(running, raise_on_error=False)
        assert rep["status"] == "Cancelled" and rep["iterations"] < 100000
        st = srv.stats()
        assert st["active"] == 0 and st["hbm_in_use"] == 0
        # the server keeps serving; a finished job cannot be cancelled
        job = srv.submit(small, abi.options(fixed_iterations=5))
        t, rep = srv.wait(job)
        assert rep["status"] == "OK" and rep["iterations"] == 5
        assert srv.cancel(job) == abi.STATUS_CODE["InvalidArgument"]
