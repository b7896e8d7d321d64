"""CPU: host logic of pooled analyses and the synthetic workloads.

A scripted stand-in for the device (no GPU) checks that driving analyses
through a pool — interleaved, resumed in whatever order the pool returns
them — makes exactly the calls, in the same per-model order, that solving
each alone makes, including rejected increments and auto-increment.
"""
import ctypes
import math

import numpy as np
import pytest


class FakeDev:
    """Answers NR iterations from a script keyed on (increment, it)."""

    def __init__(self, model, device=0):
        self.model = model
        self.calls = []
        self._keep = ctypes.c_int(0)
        self.handle = ctypes.c_void_p(ctypes.addressof(self._keep))
        self.inc = 0
        self.seed = int(model.ne)

    @staticmethod
    def opts(**kw):
        return dict(kw)

    def set_scheme(self, scheme):
        assert scheme == "FE"

    def set_dofs(self):
        self.calls.append(("set_dofs",))

    def stage_begin(self, U, F):
        self.calls.append(("stage_begin", len(U)))

    def increment_begin(self, dT, acr, add):
        self.inc += 1
        self.calls.append(("inc", round(dT, 12), add))

    def _result(self, it):
        # increment k converges at iteration 1 + (k + seed) % 3; every fifth increment fails
        k = self.inc
        fail = (k + self.seed) % 5 == 4
        need = 1 + (k + self.seed) % 3
        res = 0.0 if it >= need and not fail else 10.0 ** (1 - it) * (2.0 if fail else 1.0)
        if fail and it >= 2:
            res = 1e3   # growing residual: the NR loop gives up
        return {"res": res, "err": 1.0, "relres": 0.0, "pcg_iterations": 10 * it, "replacements": 1,
                "fallback": False, "status": 0, "message": None}

    def nr_iteration(self, it, opts):
        self.calls.append(("nr", it))
        return self._result(it)

    def nr_prepare(self, it):
        self.calls.append(("nr", it))
        self._pending = it

    def nr_finish(self, si):
        return self._result(self._pending)

    def nr_finish_begin(self, si):
        self.calls.append(("finish_begin",))

    def nr_finish_end(self):
        return self._result(self._pending)

    @staticmethod
    def nr_prepare_many(devs, its):
        for d, it in zip(devs, its):
            d.nr_prepare(it)

    @staticmethod
    def nr_finish_begin_many(devs, sis):
        for d, si in zip(devs, sis):
            d.nr_finish_begin(si)

    def increment_commit(self):
        self.calls.append(("commit",))

    def increment_reject(self):
        self.calls.append(("reject",))

    def device_bytes(self):
        return 1000 + self.model.nnodes

    def set_coarse(self, *a):
        pass

    def close(self):
        self.closed = True

    def stage_get(self, last=False):
        return np.zeros(self.model.ndofs), np.zeros(self.model.ndofs)

    def state_get(self, committed=False):
        return np.zeros_like(self.model.state)


class FakePool:
    """Returns members after a member-dependent number of steps."""

    cap = 64

    def __init__(self, *a, **k):
        self.members = {}
        self.t = 0
        self.attached = set()

    def attach(self, dev):
        self.attached.add(dev.handle.value)

    def detach(self, dev):
        self.attached.discard(dev.handle.value)

    def release_left(self):
        pass

    def forget(self, dev):
        self.attached.discard(dev.handle.value)

    def close(self):
        pass

    def add(self, dev, opts):
        assert dev.handle.value in self.attached
        assert dev.handle.value not in self.members
        self.members[dev.handle.value] = (dev, self.t + 1 + dev.seed % 4)

    def size(self):
        return len(self.members)

    def step(self):
        self.step_begin()
        return self.step_end()

    def step_done(self):
        return True

    def step_begin(self):
        # members added before step_end start at the next step, as on the device
        self.t += 1
        self.flight = list(self.members.items())

    def step_end(self):
        out = [(h, object()) for h, (d, t) in self.flight if t <= self.t]
        for h, _ in out:
            del self.members[h]
        self.flight = []
        return out


def analyses(am, n=5):
    out = []
    for i in range(n):
        mesh = am.Mesh(am.Block([[0, 0, 0], [1, 0.2, 0.2]], nx=3 + i, ny=1, nz=1, cellshape="HEX8"))
        model = am.FEModel(mesh, [("all", am.MechSolid, am.LinearElastic(E=1e9, nu=0.3))])
        ana = am.MechAnalysis(model)
        am.addstage(ana, [("x==0", am.SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", am.SurfaceBC(tz=-1.0))], nincs=3)
        am.addstage(ana, [("x==0", am.SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", am.SurfaceBC(ty=1.0))], nincs=2)
        out.append(ana)
    return out


@pytest.mark.parametrize("autoinc", [False, True])
def test_pooled_driver_makes_the_alone_calls(amaru, monkeypatch, autoinc):
    from amaru_amd import analysis, pool as P
    monkeypatch.setattr(analysis, "DeviceModel", FakeDev)
    alone = analyses(amaru)
    res_alone = [amaru.solve(a, autoinc=autoinc) for a in alone]
    pooled = analyses(amaru)
    fp = FakePool()
    args = P.solver_args(autoinc=autoinc)
    jobs = [P.Job(a, args, fp) for a in pooled]
    live = {}
    res_pool = [None] * len(jobs)
    for i, j in enumerate(jobs):
        if j.start():
            res_pool[i] = j.result
        else:
            live[j.ana.dev.handle.value] = i
    assert fp.size() == len(jobs)
    while live:
        for h, si in fp.step():
            i = live.pop(h)
            if jobs[i].resume(si):
                res_pool[i] = jobs[i].result
            else:
                live[jobs[i].ana.dev.handle.value] = i
    assert res_pool == res_alone
    for a, b in zip(alone, pooled):
        assert a.dev.calls == b.dev.calls
        assert a.stats == b.stats
        assert [s.status for s in a.stages] == [s.status for s in b.stages]
    # the script really exercises rejections
    assert any(("reject",) in a.dev.calls for a in alone)


def test_run_analyses_interleaves_like_alone(amaru, monkeypatch):
    """run_analyses resumes the members a step returns with queued-then-
    completed finishes; per model the calls are those of solving alone."""
    from amaru_amd import analysis, pool as P
    monkeypatch.setattr(analysis, "DeviceModel", FakeDev)
    alone = analyses(amaru)
    res_alone = [amaru.solve(a, autoinc=True) for a in alone]
    pooled = analyses(amaru)
    res_pool = P.run_analyses(pooled, pool=FakePool(), autoinc=True)
    assert res_pool == res_alone
    for a, b in zip(alone, pooled):
        assert a.dev.calls == [c for c in b.dev.calls if c != ("finish_begin",)]
        assert a.stats == b.stats


@pytest.mark.parametrize("on_workers", [False, True], ids=["loop", "workers"])
def test_server_isolates_failures_and_returns_budget(amaru, monkeypatch, on_workers):
    """A request whose construction raises fails alone; the others finish
    with the results of solving alone; the admitted bytes return to zero."""
    from amaru_amd import analysis, pool as P
    monkeypatch.setattr(analysis, "DeviceModel", FakeDev)
    monkeypatch.setattr(P, "DeviceModel", FakeDev)
    monkeypatch.setattr(P, "Pool", FakePool)
    alone = analyses(amaru, 4)
    ref = [amaru.solve(a) for a in alone]

    def builder(i):
        def build():
            if i == 2:
                raise ValueError("bad request")
            return analyses(amaru, 4)[i], {}
        return build
    reqs = [P.Request(i, 0.001 * i, builder(i)) for i in range(4)]
    est = max(P.estimate_device_bytes(a.model) for a in alone)
    budget = int(2.5 * est)   # room for two models at once
    # on_workers: the analyses' starts and NR resumes run on worker threads
    srv = P.Server(hbm_budget=budget, workers=2, start_on_workers=on_workers, resume_on_workers=on_workers)
    done = srv.serve(reqs)
    bad = [r for r in done if r.rid == 2][0]
    assert not bad.ok and "bad request" in bad.msg
    for r in done:
        if r.rid != 2:
            assert (r.ok, r.msg) == ref[r.rid]
            # the server creates the device model before the analysis starts, so
            # stage 1 re-checks the DOF map once (a no-op on the device)
            calls = [c for c in r.ana.dev.calls if c != ("finish_begin",)]
            assert calls[0] == ("set_dofs",) and calls[1:] == alone[r.rid].dev.calls
            assert r.ana.dev.closed
            assert r.t_arrival <= r.t_admit <= r.t_start <= r.t_done
    assert srv.in_use == 0
    # admission never exceeds the budget (how many overlap depends on timing)
    assert est <= srv.peak_in_use <= budget


def test_server_largest_first_order(amaru, monkeypatch):
    """order="largest_first": requests waiting together are built (and
    admitted) in decreasing work hint; results are those of solving alone."""
    from amaru_amd import analysis, pool as P
    monkeypatch.setattr(analysis, "DeviceModel", FakeDev)
    monkeypatch.setattr(P, "DeviceModel", FakeDev)
    monkeypatch.setattr(P, "Pool", FakePool)
    built = []

    def builder(i):
        def build():
            built.append(i)
            return analyses(amaru, 4)[i], {}
        return build
    works = [1.0, 5.0, 3.0, 4.0]
    reqs = [P.Request(i, 0.0, builder(i), work=works[i]) for i in range(4)]
    srv = P.Server(hbm_budget=1 << 40, workers=1, order="largest_first")
    done = srv.serve(reqs)
    assert built == [1, 3, 2, 0]
    ref = [amaru.solve(a) for a in analyses(amaru, 4)]
    for r in done:
        assert (r.ok, r.msg) == ref[r.rid]
    with pytest.raises(ValueError):
        P.Server(hbm_budget=1, order="random")


def test_cantilever_batch_sizes(amaru):
    from amaru_amd import workloads as W
    ps = W.cantilever_params(0, 256)
    N = np.array([p["nelem"] for p in ps])
    assert N.min() >= 500 and N.max() <= 20000
    assert np.all([10e9 <= p["E"] <= 200e9 and 0.2 <= p["nu"] <= 0.35 and abs(p["P"]) <= 1e3 for p in ps])
    assert W.cantilever_params(0, 256) == ps   # seeded
    for nel, ne in [(500, 480), (20000, 20580)]:
        ana = W.cantilever(dict(nelem=nel, E=1e10, nu=0.3, P=1.0))
        m = ana.model
        m_ = max(1, int(round((nel / 60.0) ** (1.0 / 3.0))))
        assert m.ne == 6 * 10 * m_ ** 3 == ne
        assert m.nnodes == (20 * m_ + 1) * (2 * m_ + 1) ** 2


def test_served_requests(amaru):
    from amaru_amd import workloads as W
    ps = W.request_params(0, 400)
    d = np.array([p["dofs"] for p in ps])
    assert d.min() >= 1e3 and d.max() <= 3e5
    vm = np.mean([p["vm"] for p in ps])
    assert 0.24 < vm < 0.36
    hexes = np.mean([p["shape"] == "HEX8" for p in ps])
    assert 0.42 < hexes < 0.58
    t = W.arrival_times(0, 2000, rate=5.0)
    assert np.all(np.diff(t) > 0)
    assert abs(2000 / t[-1] - 5.0) < 0.5
    for p in (dict(dofs=1e3, shape="HEX8", vm=True, E=2e11, nu=0.3), dict(dofs=3e5, shape="TET10", vm=False,
                                                                          E=2e11, nu=0.3)):
        ana, kw = W.box_request(p)
        m = ana.model
        ndofs, nu = m.configure_dofs(ana.stages[0].bcs)
        assert 0.3 * p["dofs"] < ndofs < 2.0 * p["dofs"]
        assert ana.stages[0].nincs == (5 if p["vm"] else 1)


def test_device_bytes_estimate_is_monotone(amaru):
    from amaru_amd import pool as P, workloads as W
    prev = 0
    for dofs in (1e3, 1e4, 1e5):
        ana, _ = W.box_request(dict(dofs=dofs, shape="TET10", vm=False, E=1e11, nu=0.3))
        b = P.estimate_device_bytes(ana.model)
        assert b > prev
        prev = b
    # Tet10 at 1e5 dofs: Ke (900 doubles per element) dominates, a few hundred MB
    assert 1e8 < prev < 2e9
    assert math.isfinite(prev)


def test_server_counts_unfreed_models_in_admission(amaru, monkeypatch):
    """Finished models are freed at idle points; until then their bytes
    count against the budget, so admitted plus unfreed bytes never exceed
    it (the server frees them early when a request would not fit)."""
    from amaru_amd import analysis, pool as P
    monkeypatch.setattr(analysis, "DeviceModel", FakeDev)
    monkeypatch.setattr(P, "DeviceModel", FakeDev)
    monkeypatch.setattr(P, "Pool", FakePool)
    alone = analyses(amaru, 6)
    est = max(P.estimate_device_bytes(a.model) for a in alone)
    budget = int(2.5 * est)
    reqs = [P.Request(i, 0.0, (lambda i=i: (analyses(amaru, 6)[i], {}))) for i in range(6)]
    srv = P.Server(hbm_budget=budget, workers=2)
    done = srv.serve(reqs)
    ref = [amaru.solve(a) for a in alone]
    assert [(r.ok, r.msg) for r in sorted(done, key=lambda r: r.rid)] == ref
    assert est <= srv.peak_held <= budget
    assert srv.in_use == 0
