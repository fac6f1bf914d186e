"""Serving (continuous batching): jobs that arrive while others run are
admitted as memory frees up, and each result is the same, bit for bit, as the
job run alone."""

import numpy as np
import pytest

from reyn_amd import scenes
from reyn_amd.serve import Server, make_jobs, run_fifo, same_results

pytestmark = pytest.mark.gpu


def tiny_jobs(count=6, seed=21):
    return scenes.config5_jobs(count, seed=seed, n_lo=1500, n_hi=5000, steps_lo=4, steps_hi=14)


@pytest.mark.parametrize("solver,kernel", [("iisph", "B3"), ("pcisph", "W6")])
def test_served_results_equal_results_alone(ctx, solver, kernel):
    spec = tiny_jobs()
    opts = dict(solver=solver, kernel=kernel, rho_0=1000.0)
    # arrivals close together so that jobs join while others run
    arrivals = np.linspace(0.0, 0.02, len(spec))
    served = make_jobs(spec, arrivals)
    srv = Server(ctx, **opts)
    # a budget of the two largest jobs forces queueing and later admission
    need = sorted(srv.job_bytes(sc) for sc, _ in spec)
    srv.budget = need[-1] + need[-2]
    print(f"{solver}: jobs {[(sc.n, st) for sc, st in spec]}, budget {srv.budget}", flush=True)
    rep = srv.run(served)
    srv.close()
    assert rep["peak_bytes"] <= srv.budget
    assert rep["peak_resident"] >= 2  # jobs overlapped
    assert all(j.info["steps"] == j.steps for j in served)
    assert all(j.admitted >= j.arrival and j.finished >= j.admitted for j in served)
    alone = make_jobs(spec, arrivals)
    run_fifo(ctx, alone, **opts)
    same, diff = same_results(served, alone)
    assert not diff, f"jobs {diff} differ between the mix and alone"


def test_size_class_server_results_equal_results_alone(ctx):
    """Size classes on their own streams and host threads: every result is
    still the job alone, bit for bit."""
    from reyn_amd.serve import SizeClassServer

    spec = tiny_jobs(count=8, seed=23)
    opts = dict(solver="iisph", kernel="B3", rho_0=1000.0)
    arrivals = np.linspace(0.0, 0.01, len(spec))
    served = make_jobs(spec, arrivals)
    sizes = sorted(sc.n for sc, _ in spec)
    srv = SizeClassServer(ctx, thresholds=(sizes[3] + 1,), device_setup=True, **opts)
    rep = srv.run(served)
    srv.close()
    assert [c["jobs"] for c in rep["classes"]] == [4, 4]
    assert all(j.info["steps"] == j.steps for j in served)
    alone = make_jobs(spec, arrivals)
    run_fifo(ctx, alone, device_setup=True, **opts)
    same, diff = same_results(served, alone)
    assert not diff, f"jobs {diff} differ between the size classes and alone"
