"""GPU: pooled (batched) linear solves and the continuous-batching server.

Bar: every model's displacements, reactions and IP states are bit-identical
whether its analysis runs alone (``solve``) or with others in one pool
(``run_analyses``, ``Server``), in every precision mode; pooled solutions
match the FP64 direct solve to 1e-8 relative.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

pytestmark = pytest.mark.gpu


def small_set(am):
    """Analyses of different element types, sizes, materials and stage counts."""
    out = []

    def box(shape, n, mat, nips=0, L=1.0):
        mesh = am.Mesh(am.Block([[0, 0, 0], [L, 0.2, 0.2]], nx=n[0], ny=n[1], nz=n[2], cellshape=shape))
        return am.FEModel(mesh, [("all", am.MechSolid, mat)], nips=nips)

    le = am.LinearElastic(E=200e9, nu=0.3)
    vm = am.VonMises(E=200e9, nu=0.3, fy=250e6, H=2e9)
    clamp = ("x==0", am.SurfaceBC(ux=0, uy=0, uz=0))
    # 0: Hex8 elastic, traction
    a = am.MechAnalysis(box("HEX8", (12, 3, 3), le))
    am.addstage(a, [clamp, ("x==1", am.SurfaceBC(tz=-1e5))])
    out.append((a, {}))
    # 1: Tet10 elastic, two stages
    a = am.MechAnalysis(box("TET10", (8, 2, 2), am.LinearElastic(E=70e9, nu=0.25)))
    am.addstage(a, [clamp, ("x==1", am.SurfaceBC(tz=-2e5))])
    am.addstage(a, [clamp, ("x==1", am.SurfaceBC(ty=1e5))])
    out.append((a, {"precond": "two_level"}))
    # 2: Hex20 27-IP elastic
    a = am.MechAnalysis(box("HEX20", (6, 2, 2), le, nips=27))
    am.addstage(a, [clamp, ("x==1", am.SurfaceBC(tz=-1e5))])
    out.append((a, {}))
    # 3: Hex8 von Mises, prescribed tip deflection in 4 increments
    a = am.MechAnalysis(box("HEX8", (10, 2, 2), vm))
    am.addstage(a, [clamp, ("x==1", am.SurfaceBC(uz=-0.02))], nincs=4)
    out.append((a, {"precond": "two_level"}))
    # 4: Tet10 von Mises with auto-increment
    a = am.MechAnalysis(box("TET10", (6, 2, 2), vm))
    am.addstage(a, [clamp, ("x==1", am.SurfaceBC(uz=-0.015))], nincs=3)
    out.append((a, {"autoinc": True}))
    return out


def results(ana):
    m = ana.model
    return m.uvals.copy(), m.fvals.copy(), m.state.copy(), dict(ana.stats)


@pytest.mark.parametrize("precision,graphs", [("fp64", True), ("mixed64acc", True), ("mixed", True),
                                              ("fp64", False), ("mixed64acc", False)])
def test_pool_bit_identical_to_alone(amaru, precision, graphs):
    """Also: a step's iterations launched as one captured graph (replayed
    while members join and leave) or kernel by kernel give the same bits."""
    from amaru_amd import pool as P
    alone = []
    for ana, kw in small_set(amaru):
        ok, msg = amaru.solve(ana, precision=precision, pcg_rtol=1e-10, **kw)
        assert ok, msg
        alone.append(results(ana))
        ana.dev.close()
    # pooled: one runner per distinct kwargs set, all models in one pool
    anas = small_set(amaru)
    pool = P.Pool(max_members=8)
    pool.set_graphs(graphs)
    try:
        jobs, res = [], []
        for ana, kw in anas:
            args = P.solver_args(precision=precision, pcg_rtol=1e-10, **kw)
            jobs.append(P.Job(ana, args, pool))
        live = {}
        for i, j in enumerate(jobs):
            if not j.start():
                live[j.ana.dev.handle.value] = i
        assert pool.size() == len(live) >= 4
        steps = 0
        while live:
            steps += 1
            for h, si in pool.step():
                i = live.pop(h)
                if not jobs[i].resume(si):
                    live[jobs[i].ana.dev.handle.value] = i
        for j in jobs:
            assert j.result[0], j.result[1]
        captures, replays = pool.set_graphs()
        st = pool.stats()
        assert st["member_by_member_kernels"] > st["iteration_kernels"]
        if graphs:
            assert captures >= 1 and replays > captures
            # one graph launch per step queues all its iterations' kernels
            assert st["host_submissions"] == captures + replays
            assert st["iteration_kernels"] >= 32 * st["host_submissions"]
        else:
            assert captures == replays == 0
            assert st["host_submissions"] == st["iteration_kernels"] > 0
        for (ana, _), ref in zip(anas, alone):
            got = results(ana)
            assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
            assert np.array_equal(got[2], ref[2])
            assert got[3] == ref[3]
    finally:
        for ana, _ in anas:
            if ana.dev is not None:
                pool.forget(ana.dev)
                ana.dev.close()
        pool.close()


def test_run_analyses_vs_direct(amaru):
    """Pooled mixed-precision solves of Tet10 cantilevers against SuperLU."""
    from amaru_amd import pool as P, workloads as W
    params = W.cantilever_params(3, 6)
    for p in params:
        p["nelem"] = min(p["nelem"], 3000.0)
    anas = [W.cantilever(p) for p in params]
    res = P.run_analyses(anas, precision="mixed64acc", pcg_rtol=1e-10)
    assert all(ok for ok, _ in res), res
    for ana in anas:
        m = ana.model
        bcs = ana.stages[0].bcs
        ndofs, nu = m.configure_dofs(bcs)
        U, F = m.bc_vals(bcs)
        ana.dev.assemble()
        K = ana.dev.K_csc().tocsc()
        Uo = spla.splu(K[:nu, :nu].tocsc()).solve(F[:nu] - K[:nu, nu:] @ U[nu:])
        eq = m.eq.ravel()
        Ug = np.zeros(ndofs)
        Ug[eq[eq >= 0]] = m.uvals.ravel()[eq >= 0]
        err = np.linalg.norm(Ug[:nu] - Uo) / np.linalg.norm(Uo)
        assert err <= 1e-8, err
        ana.dev.close()


def test_run_analyses_two_level_bit_identical(amaru):
    """Pooled two-level solves: every model's coarse operator is factorised
    on its own stream while other models' solves and factorisations run, and
    the results are bit-identical to solving each model alone."""
    from amaru_amd import pool as P, workloads as W
    params = W.cantilever_params(5, 10)
    for p in params:
        p["nelem"] = min(max(p["nelem"], 4000.0), 12000.0)
    kw = dict(precision="mixed64acc", precond="two_level", pcg_rtol=1e-10)
    alone = []
    for p in params:
        ana = W.cantilever(p)
        ok, msg = amaru.solve(ana, **kw)
        assert ok, msg
        alone.append(ana.model.uvals.copy())
        ana.dev.close()
    anas = [W.cantilever(p) for p in params]
    res = P.run_analyses(anas, **kw)
    assert all(ok for ok, _ in res), res
    for ana, ref in zip(anas, alone):
        assert np.array_equal(ana.model.uvals, ref)
        ana.dev.close()


def test_pooled_two_level_is_reproducible(amaru):
    """The same pooled run three times gives bit-identical solutions and
    iteration counts, although models assemble, form their coarse operators
    and finish their NR iterations on their own streams while pool steps
    run (served-like mix: Hex8/Tet10 boxes, some von Mises)."""
    from amaru_amd import pool as P, workloads as W
    params = [p for p in W.request_params(1, 48) if p["dofs"] < 1.2e5][:16]
    runs = []
    for rep in range(3):
        anas = []
        for p in params:
            ana, _ = W.box_request(p)
            ana.model.configure_dofs(ana.stages[0].bcs)
            ana.dev = amaru.DeviceModel(ana.model)
            ana.dev.set_coarse()
            anas.append(ana)
        res = P.run_analyses(anas, precision="mixed64acc", precond="two_level", pcg_rtol=1e-10)
        assert all(ok for ok, _ in res), res
        runs.append([(a.model.uvals.copy(), a.stats["pcg_iterations"]) for a in anas])
        for a in anas:
            a.dev.close()
    for other in runs[1:]:
        for (u0, i0), (u1, i1) in zip(runs[0], other):
            assert i0 == i1 and np.array_equal(u0, u1)


def test_pool_count_does_not_change_results(amaru):
    """The same analyses pooled in one pool and spread over three pools that
    step concurrently: bit-identical results and statistics."""
    from amaru_amd import pool as P
    runs = []
    for npools in (1, 3):
        anas = [a for a, _ in small_set(amaru)]
        res = P.run_analyses(anas, npools=npools, precision="mixed64acc", precond="two_level", pcg_rtol=1e-10)
        assert all(ok for ok, _ in res), res
        runs.append([results(a) for a in anas])
        for a in anas:
            a.dev.close()
    for a, b in zip(*runs):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
        assert a[3] == b[3]


def test_join_between_steps_is_bit_identical(amaru):
    """A model that joins a running pool later gets the same iterates."""
    from amaru_amd import pool as P, workloads as W
    params = W.cantilever_params(5, 3)
    for p in params:
        p["nelem"] = 1500.0
    alone = []
    for p in params:
        ana = W.cantilever(p)
        ok, msg = amaru.solve(ana, precision="mixed64acc", pcg_rtol=1e-10)
        assert ok, msg
        alone.append(ana.model.uvals.copy())
        ana.dev.close()
    anas = [W.cantilever(p) for p in params]
    pool = P.Pool(max_members=4, check_every=32)   # the alone runs check every 32 iterations too
    args = P.solver_args(precision="mixed64acc", pcg_rtol=1e-10)
    try:
        jobs = [P.Job(a, args, pool) for a in anas]
        live = {}
        started = 0
        steps = 0
        while started < len(jobs) or live:
            if started < len(jobs) and steps % 3 == 0:   # one newcomer every third step
                j = jobs[started]
                if not j.start():
                    live[j.ana.dev.handle.value] = started
                started += 1
            steps += 1
            if pool.size():
                for h, si in pool.step():
                    i = live.pop(h)
                    if not jobs[i].resume(si):
                        live[jobs[i].ana.dev.handle.value] = i
        for a, ref in zip(anas, alone):
            assert np.array_equal(a.model.uvals, ref)
    finally:
        for a in anas:
            if a.dev is not None:
                pool.forget(a.dev)
                a.dev.close()
        pool.close()


def test_server_same_alone_and_in_mix(amaru):
    """Served requests with staggered arrivals: results equal the alone runs,
    memory admission respected, latencies recorded."""
    from amaru_amd import pool as P, workloads as W
    params = W.request_params(11, 8)
    for p in params:
        p["dofs"] = min(p["dofs"], 2.0e4)
    alone = []
    for p in params:
        ana, kw = W.box_request(p)
        # the pool decides every check_every iterations; alone with the same value
        ok, msg = amaru.solve(ana, precision="mixed64acc", pcg_rtol=1e-10, pcg_check_every=16, **kw)
        assert ok, msg
        alone.append((ana.model.uvals.copy(), ana.model.fvals.copy(), ana.stats["dof_solves"]))
        ana.dev.close()
    arr = W.arrival_times(11, len(params), rate=40.0)
    reqs = [P.Request(i, arr[i], (lambda p=p: (lambda a, k: (a, dict(k, precision="mixed64acc", pcg_rtol=1e-10)))(
        *W.box_request(p)))) for i, p in enumerate(params)]
    budget = 3 * max(P.estimate_device_bytes(W.box_request(p)[0].model) for p in params)
    srv = P.Server(max_members=4, hbm_budget=budget, workers=2, check_every=16)
    try:
        done = srv.serve(reqs)
    finally:
        srv.close()
    assert srv.peak_in_use <= budget + max(r.bytes for r in done)
    for r in done:
        assert r.ok, r.msg
        u, f, ds = alone[r.rid]
        assert np.array_equal(r.ana.model.uvals, u) and np.array_equal(r.ana.model.fvals, f)
        assert r.dof_solves == ds
        assert r.latency > 0 and r.t_admit >= r.t_arrival
        assert 0 < r.bytes <= P.estimate_device_bytes(r.ana.model)
    st = P.latency_stats(done)
    assert st["p99_s"] >= st["p50_s"] > 0


def test_pool_errors(amaru):
    from amaru_amd import pool as P, _lib
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1, 1, 1]], nx=2, ny=2, nz=2, cellshape="HEX8"))
    model = amaru.FEModel(mesh, [("all", amaru.LinearElastic(E=1.0, nu=0.3))])
    bcs = [("z==0", amaru.SurfaceBC(ux=0, uy=0, uz=0)), ("z==1", amaru.SurfaceBC(tz=-1.0))]
    model.configure_dofs(bcs)
    dev = amaru.DeviceModel(model)
    U, F = model.bc_vals(bcs)
    dev.stage_begin(U, F)
    dev.increment_begin(1.0, 1.0, False)
    dev.nr_prepare(1)
    pool = P.Pool(max_members=1)
    with pytest.raises(_lib.AmaruError):   # not attached
        pool.add(dev, dev.opts())
    pool.attach(dev)
    pool.add(dev, dev.opts())
    with pytest.raises(_lib.AmaruError):   # already a member
        pool.add(dev, dev.opts())
    with pytest.raises(_lib.AmaruError):   # solving alone while in the pool
        dev.solve_system(U.copy(), F.copy())
    with pytest.raises(_lib.AmaruError):   # its vectors are in use
        dev.nr_finish(None)
    fin = []
    for _ in range(1000):
        fin += pool.step()
        if fin:
            break
    (h, si), = fin
    assert h == dev.handle.value and si.converged
    r = dev.nr_finish(si)
    assert r["status"] == 0 and r["res"] < 1e-6
    dev.close()          # detaches itself
    assert pool.size() == 0
    pool.close()
