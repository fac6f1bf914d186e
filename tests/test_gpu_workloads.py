"""The batched and served workloads at their stated sizes, through the C-ABI:

- configuration 4: the full batch of 64 PCISPH scenes of 10k-300k particles
  (seed 4, about 5.5M particles), every step after the first replayed as
  captured HIP graphs; a sample of scenes run alone gives the same bits as in
  the mix; every scene's iterations lie within [min_iter, max_iter] and its
  state stays finite and in its box; one step of a scene from its batch state
  against the FP64 oracle.
- configuration 5: eight IISPH jobs drawn with the configuration's defaults
  (5k-500k particles, 200-2000 steps, box sides U[1,3], a sphere with p 0.5)
  on a compressed arrival clock, served by continuous batching; every result
  equals the job alone bit for bit; one step of the largest job (the same
  kernels as serving) against the FP64 oracle.

The FP64 bars are the single-scene ones: rho, p and x within 1e-5 and v
within 1e-4 of each field's largest value.
"""

import numpy as np
import pytest

from oracle import parity
from reyn_amd import scenes
from reyn_amd.serve import Server, make_jobs, run_fifo, same_results
from reyn_amd.sim import Batch

pytestmark = pytest.mark.gpu

TOL = 1e-5
TOL_DIFF = 1e-4
FIELDS = ("xx", "xy", "xz", "vx", "vy", "vz", "rho")


def check_step(errs, what):
    print(f"{what}: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    assert errs["rho"] <= TOL and errs["x"] <= TOL, (what, errs)
    assert errs.get("p", 0.0) <= TOL and errs["v"] <= TOL_DIFF, (what, errs)


def test_config4_full_batch(ctx):
    scs = scenes.config4_batch()
    n = sum(sc.n for sc in scs)
    print(f"configuration 4: {len(scs)} scenes, {n} particles ({min(sc.n for sc in scs)}..{max(sc.n for sc in scs)})",
          flush=True)
    opts = dict(solver="pcisph", kernel="W6", eta_rho_max=0.001, min_iter=3, max_iter=100, cell_ratio=1.0,
                check_every=2)
    mix = Batch(ctx, scs, graph=True, **opts)
    mix.step(3)
    order = np.argsort([sc.n for sc in scs])
    mid = int(order[len(order) // 2])
    check_step(parity.batch_scene_step_parity(mix, mid), f"scene {mid} ({scs[mid].n} particles), step 4 vs FP64")
    for s, sc in enumerate(scs):
        inf = mix.info(s)
        assert inf.steps == 4 and 3 <= inf.iterations <= 100, (s, inf.iterations)
        d = mix.download(s)
        x = np.stack([d["xx"], d["xy"], d["xz"]], 1)
        assert np.all(np.isfinite(x)) and np.all(np.isfinite(d["vx"])), s
        assert np.all(x >= sc.bound_min) and np.all(x <= sc.bound_max), s
    for s in (int(order[0]), mid, int(order[-1])):
        alone = Batch(ctx, [scs[s]], graph=True, **opts)
        alone.step(4)
        a, b = alone.download(0), mix.download(s)
        for k in FIELDS:
            assert np.array_equal(a[k], b[k]), (s, k)
        assert (alone.info(0).iterations, alone.info(0).dt) == (mix.info(s).iterations, mix.info(s).dt)
        alone.close()
    mix.close()


def test_config5_served_jobs(ctx):
    spec = scenes.config5_jobs(8)
    print("configuration 5 jobs: " + ", ".join(f"{sc.n} x {st}" for sc, st in spec), flush=True)
    opts = dict(solver="iisph", kernel="B3", rho_0=1000.0, eta_rho_max=0.001, omega=0.5, min_iter=3, max_iter=100,
                cell_ratio=1.0, check_every=2)
    # arrivals closer together than the jobs take, so jobs join while others run
    arrivals = np.linspace(0.0, 0.4, len(spec))
    served = make_jobs(spec, arrivals)
    srv = Server(ctx, device_setup=True, **opts)
    rep = srv.run(served)
    srv.close()
    print(f"served: p50 {rep['p50_s']:.2f} s, p99 {rep['p99_s']:.2f} s, peak resident {rep['peak_resident']}",
          flush=True)
    assert rep["peak_resident"] >= 2
    assert all(j.info["steps"] == j.steps for j in served)
    alone = make_jobs(spec, arrivals)
    run_fifo(ctx, alone, device_setup=True, **opts)
    same, diff = same_results(served, alone)
    assert not diff, f"jobs {diff} differ between the mix and alone"
    big = max(range(len(spec)), key=lambda k: spec[k][0].n)
    b = Batch(ctx, [spec[big][0]], **opts)
    b.step(2)
    check_step(parity.batch_scene_step_parity(b, 0), f"job {big} ({spec[big][0].n} particles), step 3 vs FP64")
    b.close()


def test_brick_order_batch_scene_against_fp64(ctx):
    """A batch in the brick cell order: one step of a scene from its batch
    state against the FP64 oracle's step over the oracle's brick-ordered grid."""
    scs = scenes.config4_batch(count=4, seed=9, n_lo=5000, n_hi=20000)
    opts = dict(solver="iisph", kernel="B3", eta_rho_max=0.001, min_iter=3, max_iter=100, cell_ratio=1.0,
                check_every=2, cell_order="brick")
    b = Batch(ctx, scs, **opts)
    b.step(3)
    check_step(parity.batch_scene_step_parity(b, 2), f"brick-order batch scene 2 ({scs[2].n} particles) vs FP64")
    b.close()
