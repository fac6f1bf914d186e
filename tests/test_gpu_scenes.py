"""Device-side scene generation: the fluid lattice (with the hash jitter and
the sphere cut), the wall shell and the sphere shells generated on the GPU
equal the host generator's points bit for bit, and a scene added to a batch
from its description runs exactly as the same scene added from host
arrays."""

import numpy as np
import pytest

from reyn_amd import scenes
from reyn_amd.sim import Batch, generate_dam_break

pytestmark = pytest.mark.gpu


def _cases():
    yield "config1", scenes.config1_sesph()
    yield "config3-side48", scenes.config3_iisph(side=48)
    for k, (sc, _) in enumerate(scenes.config5_jobs(6, seed=12, n_lo=2000, n_hi=60000)):
        yield f"job{k}", sc


@pytest.mark.parametrize("k", range(8))
def test_device_scene_equals_host_scene(ctx, k):
    name, sc = list(_cases())[k]
    print(f"{name}: {sc.n} fluid, {sc.n_bdy} boundary, sphere {sc.meta.get('sphere')}", flush=True)
    x, xb = generate_dam_break(ctx, scenes.dam_break_desc(sc))
    assert x.shape == sc.x.shape and xb.shape == sc.xb.shape
    assert np.array_equal(x, sc.x), f"fluid differs at {np.argwhere(x != sc.x)[:5]}"
    assert np.array_equal(xb, sc.xb), f"boundary differs at {np.argwhere(xb != sc.xb)[:5]}"


@pytest.mark.parametrize("solver,kernel", [("iisph", "B3"), ("pcisph", "W6")])
def test_batch_scenes_from_descriptions(ctx, solver, kernel):
    scs = [sc for sc, _ in scenes.config5_jobs(3, seed=8, n_lo=2000, n_hi=8000)]
    host = Batch(ctx, scs, solver=solver, kernel=kernel)
    dev = Batch(ctx, [], solver=solver, kernel=kernel, rho_0=scs[0].rho_0)
    ids = dev.add_dam_breaks(scs)
    host.step(6)
    dev.step(6)
    for k, i in enumerate(ids):
        a, b = host.download(k), dev.download(i)
        for f in a:
            assert np.array_equal(a[f], b[f]), (k, f)
        ia, ib = host.info(k), dev.info(i)
        assert (ia.steps, ia.iterations, ia.dt, ia.time) == (ib.steps, ib.iterations, ib.dt, ib.time)
    host.close()
    dev.close()
