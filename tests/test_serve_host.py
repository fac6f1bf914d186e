"""The serving loop's admission and continuous batching, driven against a
stand-in batch on the CPU: the HBM budget holds at every admission, jobs are
admitted in arrival order, jobs join while others run, and each job leaves
after exactly its steps."""

import numpy as np
import pytest

from reyn_amd import scenes, serve


class FakeBatch:
    log = []

    def __init__(self, ctx, scenes_, **options):
        self.prm = None
        self.resident = {}
        self.next = 0
        self.steps = {}

    def add(self, scs):
        ids = []
        for sc in scs:
            self.resident[self.next] = sc
            self.steps[self.next] = 0
            ids.append(self.next)
            self.next += 1
        FakeBatch.log.append(("add", [sc.meta["job"] for sc in scs]))
        return ids

    def remove(self, ids):
        for i in ids:
            del self.resident[i]
        FakeBatch.log.append(("remove", list(ids)))

    def step(self, k):
        assert self.resident
        for i in self.resident:
            self.steps[i] += k

    def download(self, i):
        return {"xx": np.array([float(i)])}

    def info(self, i):
        class I:
            pass
        o = I()
        o.steps, o.iterations, o.dt, o.time = self.steps[i], 3, 1e-3, 1e-3 * self.steps[i]
        return o

    def close(self):
        pass


@pytest.fixture
def fake(monkeypatch):
    FakeBatch.log = []
    monkeypatch.setattr(serve, "Batch", FakeBatch)
    monkeypatch.setattr(serve, "device_memory", lambda ctx: (100 << 30, 100 << 30))
    monkeypatch.setattr(serve, "scene_bytes", lambda ctx, prm, sc: sc.n * 1000)
    return FakeBatch


def test_admission_respects_budget_and_order(fake):
    spec = scenes.config5_jobs(8, seed=3, n_lo=1000, n_hi=4000, steps_lo=3, steps_hi=9)
    jobs = serve.make_jobs(spec, np.zeros(len(spec)))
    srv = serve.Server(None, reserve_bytes=0)
    srv.budget = 2 * max(sc.n for sc, _ in spec) * 1000
    rep = srv.run(jobs)
    assert rep["peak_bytes"] <= srv.budget
    admitted = [k for op, ks in fake.log if op == "add" for k in ks]
    assert admitted == list(range(len(spec)))  # FIFO
    assert all(j.info["steps"] == j.steps for j in jobs)
    assert rep["jobs"] == len(spec) and rep["particle_steps"] == sum(sc.n * st for sc, st in spec)


def test_jobs_join_a_running_batch(fake):
    spec = scenes.config5_jobs(4, seed=4, n_lo=1000, n_hi=2000, steps_lo=50, steps_hi=60)
    # the later jobs arrive while the first runs
    jobs = serve.make_jobs(spec, [0.0, 0.001, 0.002, 0.003])
    t = [0.0]

    def clock():  # every call advances 0.5 ms
        t[0] += 0.0005
        return t[0]

    srv = serve.Server(None, reserve_bytes=0)
    rep = srv.run(jobs, clock=clock)
    adds = [ks for op, ks in fake.log if op == "add"]
    assert len(adds) >= 2 and rep["peak_resident"] >= 2
    assert all(j.taken == j.steps for j in jobs)


def test_job_over_budget_is_refused(fake):
    spec = scenes.config5_jobs(2, seed=5, n_lo=1000, n_hi=2000, steps_lo=3, steps_hi=4)
    srv = serve.Server(None, reserve_bytes=0)
    srv.budget = 10
    with pytest.raises(RuntimeError):
        srv.run(serve.make_jobs(spec, [0.0, 0.0]))


def test_config5_jobs_are_seeded():
    a = scenes.config5_jobs(5, seed=9)
    b = scenes.config5_jobs(5, seed=9)
    for (sa, ka), (sb, kb) in zip(a, b):
        assert ka == kb and np.array_equal(sa.x, sb.x) and np.array_equal(sa.xb, sb.xb)
    assert all(200 <= k <= 2000 for _, k in a)
    assert all(4000 <= sc.n <= 520000 for sc, _ in a)


def test_poisson_arrivals():
    t = serve.poisson_arrivals(20000, 50.0, seed=1)
    assert np.all(np.diff(t) > 0)
    assert abs(np.diff(t).mean() - 1 / 50.0) < 0.05 / 50.0


def test_size_classes_split_and_budget(fake, monkeypatch):
    """Jobs go to their size class; every class's lane admits under the one
    budget and each job leaves after exactly its steps."""
    monkeypatch.setattr(serve, "Context", lambda device: None)
    spec = scenes.config5_jobs(10, seed=6, n_lo=1000, n_hi=20000, steps_lo=3, steps_hi=8)
    jobs = serve.make_jobs(spec, np.zeros(len(spec)))
    srv = serve.SizeClassServer(types_ns_ctx(), thresholds=(5000,), reserve_bytes=0)
    srv.budget = 3 * max(sc.n for sc, _ in spec) * 1000
    rep = srv.run(jobs)
    small = sum(1 for sc, _ in spec if sc.n < 5000)
    assert [c["jobs"] for c in rep["classes"]] == [small, len(spec) - small]
    assert rep["peak_bytes"] <= srv.budget
    assert all(j.taken == j.steps and j.finished is not None for j in jobs)
    assert [srv.class_of(j) for j in jobs] == [0 if j.scene.n < 5000 else 1 for j in jobs]


def types_ns_ctx():
    class Ctx:
        device = 0

        def close(self):
            pass
    return Ctx()
