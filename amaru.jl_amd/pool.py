"""Batched and served analyses: many models' linear solves advanced together.

A ``Pool`` (``amr_pool_*`` in the C ABI) runs the PCG of every member model
with one launch per PCG phase for all of them.  ``run_analyses`` runs a set
of ``MechAnalysis`` objects to completion with their linear solves pooled
(batching); ``Server`` admits requests as they arrive, by free device
memory, and keeps each model resident across its load increments while
models join and leave the pool between steps (continuous batching).

The per-model control flow is the reference's ``mech_stage_solver!``
(src/mech/mech-solver.jl:186-492), unchanged: ``analysis_machine`` yields
at each Newton-Raphson iteration and is resumed with the iteration's result.
A model's PCG iterates do not depend on what else is in the pool, so every
result is bit-identical to solving that model alone with ``solve`` with the
same ``pcg_check_every`` as the pool's ``check_every`` (the iterations
between convergence and precision decisions).
"""
import collections
import concurrent.futures
import ctypes
import threading
import time

import numpy as np

from . import _lib
from .analysis import analysis_machine, solver_args
from .model import DeviceModel


class Pool:
    """amr_pool: up to ``max_members`` models solving at once on one device."""

    def __init__(self, device=0, max_members=256, check_every=32):
        self._h = ctypes.c_void_p()
        _lib.check(_lib.lib().amr_pool_create(ctypes.byref(self._h), int(device), int(max_members),
                                              int(check_every)))
        self.device = device
        self.cap = int(max_members)
        self._out_m = (ctypes.c_void_p * self.cap)()
        self._out_i = (_lib.SolveInfo * self.cap)()
        self._attached = set()

    def close(self):
        if self._h:
            _lib.lib().amr_pool_destroy(self._h)
            self._h = None
            self._attached.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def attach(self, dev):
        if dev.handle.value not in self._attached:
            _lib.check(_lib.lib().amr_pool_attach(self._h, dev.handle))
            self._attached.add(dev.handle.value)

    def detach(self, dev):
        if dev.handle and dev.handle.value in self._attached:
            _lib.check(_lib.lib().amr_pool_detach(dev.handle))
            self._attached.discard(dev.handle.value)

    def forget(self, dev):
        """The model is about to be destroyed (which detaches it)."""
        if dev.handle:
            self._attached.discard(dev.handle.value)

    def add(self, dev, opts):
        _lib.check(_lib.lib().amr_pool_add(self._h, dev.handle, ctypes.byref(opts)))

    def size(self):
        n = ctypes.c_int(0)
        _lib.check(_lib.lib().amr_pool_size(self._h, ctypes.byref(n)))
        return n.value

    def _results(self, n):
        return [(self._out_m[i], _lib.SolveInfo.from_buffer_copy(self._out_i[i])) for i in range(n)]

    def step(self):
        """One step; returns [(model handle value, SolveInfo)] of members that left."""
        n = ctypes.c_int(0)
        _lib.check(_lib.lib().amr_pool_step(self._h, self.cap, ctypes.byref(n), self._out_m, self._out_i))
        return self._results(n.value)

    def step_begin(self):
        """Queues a step and returns at once; members added before
        step_end start at the next step."""
        _lib.check(_lib.lib().amr_pool_step_begin(self._h))

    def release_left(self):
        """The models that left at the last step_end continue on their own
        streams from now on (else from the next step_begin); call it before
        continuing such a model on another thread."""
        _lib.check(_lib.lib().amr_pool_release_left(self._h))

    def step_done(self):
        """True when step_end would not wait."""
        d = ctypes.c_int(1)
        _lib.check(_lib.lib().amr_pool_step_done(self._h, ctypes.byref(d)))
        return bool(d.value)

    def step_end(self):
        """Waits for the step step_begin queued; as step()."""
        n = ctypes.c_int(0)
        _lib.check(_lib.lib().amr_pool_step_end(self._h, self.cap, ctypes.byref(n), self._out_m, self._out_i))
        return self._results(n.value)

    def set_graphs(self, enable=None):
        """Steps as one captured HIP graph launch (True) or kernel by kernel
        (False, the default: faster on MI355X); returns (captures, replays)
        so far."""
        c, r = ctypes.c_int64(0), ctypes.c_int64(0)
        flag = -1 if enable is None else (1 if enable else 0)
        _lib.check(_lib.lib().amr_pool_set_graphs(self._h, flag, ctypes.byref(c), ctypes.byref(r)))
        return c.value, r.value

    def stats(self):
        """Counters so far: batched coarse-operator formations and the models
        they formed; kernels of the steps' PCG iterations, the host
        submissions that queued them (graph launches when captured), and the
        kernels the same iterations would take launched member by member."""
        out = np.zeros(5, dtype=np.int64)
        _lib.check(_lib.lib().amr_pool_stats(self._h, _lib.ptr(out)))
        return {"coarse_batches": int(out[0]), "coarse_models": int(out[1]), "iteration_kernels": int(out[2]),
                "host_submissions": int(out[3]), "member_by_member_kernels": int(out[4])}

    def coarse_batches(self):
        """(batched coarse-operator formations, models they formed) so far."""
        st = self.stats()
        return st["coarse_batches"], st["coarse_models"]

    def timing(self, enable):
        ms = ctypes.c_double(0.0)
        n = ctypes.c_int64(0)
        _lib.check(_lib.lib().amr_pool_timing(self._h, 1 if enable else 0, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value


class Job:
    """One analysis driven through a pool: its machine is advanced until it
    needs a pooled linear solve, then parked until the pool returns it."""

    def __init__(self, ana, args, pool):
        self.ana = ana
        self.pool = pool
        self.gen = analysis_machine(ana, args)
        self.result = None
        self.done = False

    def start(self):
        return self._advance(None, first=True)

    def resume(self, solve_info):
        return self._advance(self.ana.dev.nr_finish(solve_info))

    def resume_begin(self, solve_info):
        """Queue this member's NR finish; resume_end completes it."""
        self.ana.dev.nr_finish_begin(solve_info)

    def resume_end(self):
        return self._advance(self.ana.dev.nr_finish_end())

    def prestart(self):
        """The part of start() that touches no pool, for a worker thread:
        the machine runs to its first pooled solve (stage set-up, boundary
        values) and that iteration's mount_K and right-hand side are queued
        on the model's own stream; returns True when the analysis has
        already ended.  join_started() completes the start."""
        it = self.next_solve(first=True, attach=False)
        if it is None:
            return True
        self.ana.dev.nr_prepare(it)
        return False

    def join_started(self):
        """After prestart(): the prepared solve joins the pool (on the
        thread that drives the pool)."""
        if self.done:
            return True
        self.pool.attach(self.ana.dev)
        self.join()
        return False

    def next_solve(self, r=None, first=False, attach=True):
        """Advances the machine (sending it r) to its next Newton-Raphson
        iteration that needs a pooled linear solve and returns that
        iteration's number, or None when the analysis has ended (result
        set).  The iterations of a model without unknowns are run here.
        With attach the model is attached to the job's pool."""
        try:
            it = next(self.gen) if first else self.gen.send(r)
            while True:
                dev = self.ana.dev
                if attach:
                    self.pool.attach(dev)
                if dev.model.nu == 0:
                    dev.nr_prepare(it)
                    it = self.gen.send(dev.nr_finish(None))
                    continue
                return it
        except StopIteration as stop:
            self.result = stop.value
            self.done = True
            return None

    def join(self):
        """The prepared iteration's linear solve joins the pool."""
        self.pool.add(self.ana.dev, self.ana.opts)

    def _advance(self, r, first=False):
        it = self.next_solve(r, first)
        if it is None:
            return True
        self.ana.dev.nr_prepare(it)
        self.join()
        return False


def prepare_and_join(pairs):
    """[(job, it)]: mount_K and the right-hand side of every job's iteration
    it, their element kernels batched across the models, then each job's
    solve joins its pool."""
    if not pairs:
        return
    devs = [j.ana.dev for j, _ in pairs]
    type(devs[0]).nr_prepare_many(devs, [it for _, it in pairs])
    for j, _ in pairs:
        j.join()


def start_all(jobs):
    """Job.start for every job, their first mount_K batched; returns each
    job's done flag."""
    pairs, done = [], []
    for j in jobs:
        it = j.next_solve(first=True)
        done.append(it is None)
        if it is not None:
            pairs.append((j, it))
    prepare_and_join(pairs)
    return done


def resume_all(finished, jobs_of, catch=False, between=None):
    """Resume the jobs of the members a step finished: all NR finishes are
    queued first (their update kernels batched across the models) and then
    completed, so they share one wait; ``between()`` runs after the finishes
    are queued (the next pool step is launched there, so that it overlaps
    the host work of completing them); the next iterations' mount_K are
    batched too.  Yields (handle, job, done), or with catch=True (handle,
    job, done, exception) so that one failing job does not stop the
    others."""
    jobs = [(h, jobs_of(h)) for h, _ in finished]
    errs = {}
    try:
        if jobs:
            devs = [job.ana.dev for _, job in jobs]
            type(devs[0]).nr_finish_begin_many(devs, [si for _, si in finished])
    except Exception as e:
        if not catch:
            raise
        errs = {h: e for h, _ in jobs}
    if between is not None:
        between()
    out, pairs = [], []
    for h, job in jobs:
        if h in errs:
            out.append((h, job, True, errs[h]))
            continue
        try:
            it = job.next_solve(job.ana.dev.nr_finish_end())
        except Exception as e:
            if not catch:
                raise
            out.append((h, job, True, e))
            continue
        if it is None:
            out.append((h, job, True, None))
        else:
            pairs.append((job, it))
            out.append((h, job, False, None))
    try:
        prepare_and_join(pairs)
    except Exception as e:
        if not catch:
            raise
        failed = {id(j) for j, _ in pairs}
        out = [(h, job, True, e) if id(job) in failed else r for (h, job, d, x), r in zip(out, out)]
    for h, job, done, exc in out:
        yield (h, job, done, exc) if catch else (h, job, done)


class _Pipeline:
    """Keeps one pool step in flight: ``end()`` collects it, and
    ``relaunch()`` (called by resume_all between queuing and completing the
    finished members' NR finishes) queues the next one for the members
    still solving.  Members added while a step is in flight start at the
    next step."""

    def __init__(self, pool):
        self.pool = pool
        self.in_flight = False

    def launch(self):
        if not self.in_flight and self.pool.size() > 0:
            self.pool.step_begin()
            self.in_flight = True

    relaunch = launch

    def end(self):
        if not self.in_flight:
            return []
        self.in_flight = False
        return self.pool.step_end()


class PoolSet:
    """Several pools on one device, each with its own stream, driven from one
    host thread.  A model belongs to one pool (the least loaded one, by a
    work weight, when its analysis starts) for its whole analysis.  Each
    pool's next step is queued as soon as its previous one is collected, so
    the pools' steps overlap on the device and a model's NR iterations are
    not held to the cadence of every other model.  A model's results do not
    depend on its pool or on the number of pools."""

    def __init__(self, npools=1, device=0, max_members=256, check_every=32, pools=None):
        self.pools = list(pools) if pools is not None else [
            Pool(device, max_members=max_members, check_every=check_every) for _ in range(max(1, int(npools)))]
        self.pipes = [_Pipeline(p) for p in self.pools]
        self.load = [0.0] * len(self.pools)
        self.count = [0] * len(self.pools)   # analyses assigned (each has at most one member at a time)
        self.t_launch = [0.0] * len(self.pools)
        self.launches = 0

    @property
    def cap(self):
        return sum(p.cap for p in self.pools)

    def assign(self, weight):
        free = [i for i in range(len(self.pools)) if self.count[i] < self.pools[i].cap]
        if not free:
            raise RuntimeError("every pool is full")
        k = min(free, key=lambda i: self.load[i])
        self.load[k] += weight
        self.count[k] += 1
        return k

    def release(self, k, weight):
        self.load[k] -= weight
        self.count[k] -= 1

    def size(self):
        return sum(p.size() for p in self.pools)

    def stats(self):
        """Pool.stats summed over the pools, with the steps launched."""
        out = {}
        for p in self.pools:
            for k, v in p.stats().items():
                out[k] = out.get(k, 0) + v
        out["steps"] = self.launches
        return out

    def in_flight(self):
        return any(p.in_flight for p in self.pipes)

    def _launch(self, k):
        if not self.pipes[k].in_flight and self.pools[k].size() > 0:
            self.pipes[k].launch()
            self.t_launch[k] = time.perf_counter()
            self.launches += 1

    def launch_idle(self):
        for k in range(len(self.pools)):
            self._launch(k)

    def relauncher(self, k):
        return lambda: self._launch(k)

    def collect(self, block=True):
        """[(pool index, finished members)] of the pools whose step has
        finished; with block and none finished, waits for the step queued
        first."""
        out = [(k, pipe.end()) for k, pipe in enumerate(self.pipes) if pipe.in_flight and self.pools[k].step_done()]
        if not out and block:
            fl = [k for k, pipe in enumerate(self.pipes) if pipe.in_flight]
            if fl:
                k = min(fl, key=lambda i: self.t_launch[i])
                out.append((k, self.pipes[k].end()))
        return out

    def end_all(self):
        for pipe in self.pipes:
            pipe.end()

    def forget(self, dev):
        for p in self.pools:
            p.forget(dev)

    def detach(self, dev):
        for p in self.pools:
            p.detach(dev)

    def timing(self, enable):
        """SpMV timing summed over the pools (as Pool.timing)."""
        ms = n = 0
        for p in self.pools:
            a, b = p.timing(enable)
            ms += a
            n += b
        return ms, n

    def close(self):
        for p in self.pools:
            p.close()


def work_weight(ana):
    """Relative device work of an analysis, for balancing pools: nodes times
    load increments."""
    return float(ana.model.nnodes) * max(1, sum(st.nincs for st in ana.stages))


def run_analyses(analyses, pool=None, device=0, check_every=32, npools=1, **kwargs):
    """solve!(ana; kwargs...) for every analysis, their linear solves pooled
    (in ``pool``: a Pool or a PoolSet; by default in ``npools`` pools of the
    library's own).  Returns the list of (success, message)."""
    args = solver_args(device=device, **kwargs)
    if args["scheme"] != "FE":
        raise NotImplementedError("pooled solves run scheme FE; solve() runs the others")
    own = pool is None
    if own:
        ps = PoolSet(npools, device, max_members=max(1, len(analyses)), check_every=check_every)
    else:
        ps = pool if isinstance(pool, PoolSet) else PoolSet(pools=[pool])
    try:
        jobs = {}
        results = [None] * len(analyses)
        order = sorted(range(len(analyses)), key=lambda i: -work_weight(analyses[i]))
        home = {i: ps.assign(work_weight(analyses[i])) for i in order}
        started = [Job(ana, args, ps.pools[home[i]]) for i, ana in enumerate(analyses)]
        for i, (job, done) in enumerate(zip(started, start_all(started))):
            if done:
                results[i] = job.result
            else:
                jobs[job.ana.dev.handle.value] = (i, job)
        while jobs:
            ps.launch_idle()
            for k, fin in ps.collect(block=True):
                idx = {h: jobs.pop(h) for h, _ in fin}
                for h, job, done in resume_all(fin, lambda h: idx[h][1], between=ps.relauncher(k)):
                    i = idx[h][0]
                    if done:
                        results[i] = job.result
                    else:
                        jobs[job.ana.dev.handle.value] = (i, job)
        return results
    finally:
        ps.end_all()
        if own:
            for ana in analyses:
                if ana.dev is not None:
                    ps.detach(ana.dev)
            ps.close()


# ------------------------------------------------------------------ serving

def estimate_device_bytes(model):
    """Upper estimate of the device memory amr_model_create and the stage
    driver allocate for ``model`` (admission before the pattern is built):
    node-block count bounded by the element type's neighbour count."""
    nn = model.nen
    nnodes, ne, nips = model.nnodes, model.ne, model.nips
    maxnb = {8: 27, 20: 81, 10: 65}[nn]
    nblk = min(nnodes * maxnb, ne * nn * nn)
    per_node = 888
    per_elem = 4 * nn + 4 + 72 * nn * nn + 24 * nn + 4 + 4 * nn + 224 * nips + 4 * nn * nn
    per_blk = 4 + 8 + 72 + 36 + 9
    # + the padding of the 32 vector slices (to 64 entries each)
    return int(per_node * nnodes + per_elem * ne + per_blk * (nblk + 64) + 46080 * 8 + 4096 + 24 * 64 * 8 + 8 * 64 * 4)


class Request:
    """A served request: ``build()`` returns (MechAnalysis, solver kwargs)
    on the host; the analysis runs on the device when admitted."""

    __slots__ = ("rid", "t_arrival", "build", "work", "ana", "kwargs", "reserve", "bytes", "t_admit", "t_start",
                 "t_done", "ok", "msg", "dof_solves", "job", "home", "weight")

    def __init__(self, rid, t_arrival, build, work=None):
        self.rid = rid
        self.t_arrival = float(t_arrival)
        self.build = build
        # optional client estimate of the request's device work (for example
        # DOFs x load increments), used by the largest-first order
        self.work = work
        self.ana = self.kwargs = self.job = self.home = None
        self.reserve = self.bytes = 0
        self.weight = 0.0
        self.t_admit = self.t_start = self.t_done = None
        self.ok = None
        self.msg = None
        self.dof_solves = 0

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_arrival


class Server:
    """Continuous-batching FE solve server on one device.

    Requests arrive at their ``t_arrival`` (seconds after ``serve`` starts).
    Host-side model construction and device upload run on worker threads; a
    request is admitted when its estimated device memory fits the HBM budget
    and the pool has a free member slot; its model stays resident until its
    analysis ends; its linear solves join the shared pool between steps.
    """

    def __init__(self, device=0, max_members=256, check_every=16, hbm_budget=None, workers=4, npools=3,
                 order="fifo", start_on_workers=True, resume_on_workers=True):
        if order not in ("fifo", "largest_first"):
            raise ValueError("order must be 'fifo' or 'largest_first'")
        self.order = order
        # an admitted request's analysis runs to its first solve (and queues
        # that iteration's mount_K) on the device worker; the serving loop
        # only joins it to its pool
        self.start_on_workers = bool(start_on_workers)
        # a member whose solve finished completes its NR iteration and
        # prepares the next one (or ends its analysis) on a worker; the
        # serving loop joins it to its pool again
        self.resume_on_workers = bool(resume_on_workers)
        self.device = device
        npools = max(1, int(npools))
        self.pools = PoolSet(npools, device, max_members=-(-int(max_members) // npools), check_every=check_every)
        if hbm_budget is None:
            free = ctypes.c_int64(0)
            total = ctypes.c_int64(0)
            _lib.check(_lib.lib().amr_mem_info(int(device), ctypes.byref(free), ctypes.byref(total)))
            hbm_budget = int(0.8 * free.value)
        self.budget = int(hbm_budget)
        self.workers = int(workers)
        self.in_use = 0
        self.peak_in_use = 0
        self.peak_held = 0      # admitted bytes plus finished models not freed yet
        # main-loop time accounting (seconds): pool steps, resuming finished
        # members, starting admitted ones, idle waits
        self.t_loop = {"step": 0.0, "resume": 0.0, "start": 0.0, "wait": 0.0, "free": 0.0}
        # worker time (seconds, summed over requests): host construction and
        # device set-up (upload, pattern, coarse space)
        self.t_build = {"host": 0.0, "device": 0.0, "resume": 0.0}
        # finished models whose device memory is released later (a release
        # can wait for the device), at idle points or when too much waits
        self._graveyard = []
        self._graveyard_bytes = 0
        self.graveyard_limit = 16 << 30

    def close(self):
        self.pools.close()

    @property
    def steps(self):
        return self.pools.launches

    def _finish(self, req, now, exc=None):
        """Results are on the host (the analysis read them back); free the
        model and its budget.  With exc the request failed."""
        req.t_done = now
        if req.home is not None:
            self.pools.release(req.home, req.weight)
            req.home = None
        if exc is not None or req.job is None or req.job.result is None:
            req.ok, req.msg = False, repr(exc) if exc is not None else "request failed"
        else:
            req.ok, req.msg = req.job.result
        req.dof_solves = req.ana.stats["dof_solves"] if req.ana is not None else 0
        dev = req.ana.dev if req.ana is not None else None
        if dev is not None and dev.handle:
            self._graveyard.append(dev)
            self._graveyard_bytes += req.bytes
            if self._graveyard_bytes > self.graveyard_limit:
                self._release()
        self.in_use -= req.bytes if req.bytes else req.reserve

    def _release(self):
        """Detach the finished models from the pools here; destroy them on
        the closing thread when serving (a destroy waits for the model's
        stream and copies its committed states to the host)."""
        ts = time.perf_counter()
        closer = getattr(self, "_closer", None)
        for dev in self._graveyard:
            self.pools.detach(dev)
            if closer is not None:
                closer.submit(dev.close)
            else:
                dev.close()
        self._graveyard.clear()
        self._graveyard_bytes = 0
        self.t_loop["free"] += time.perf_counter() - ts

    def serve(self, requests):
        """Runs every request; returns them with their timings filled in.
        A request whose construction or analysis raises fails alone
        (ok False, msg the exception)."""
        reqs = sorted(requests, key=lambda r: r.t_arrival)
        N = len(reqs)
        # host construction and device set-up have workers of their own, so
        # that admitted requests are not queued behind arrivals' host builds
        ex = concurrent.futures.ThreadPoolExecutor(max_workers=self.workers)
        ex_dev = concurrent.futures.ThreadPoolExecutor(max_workers=max(1, self.workers // 2))
        self._closer = concurrent.futures.ThreadPoolExecutor(max_workers=1)
        ex_res = concurrent.futures.ThreadPoolExecutor(max_workers=2)
        lock = threading.Lock()
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0  # noqa: E731

        def timed(kind, ts):
            with lock:
                self.t_build[kind] += time.perf_counter() - ts

        def host_build(req):
            ts = time.perf_counter()
            try:
                ana, kw = req.build()
                ana.model.configure_dofs(ana.stages[0].bcs)
                return req, ana, kw, estimate_device_bytes(ana.model), None
            except Exception as e:
                return req, None, None, 0, e
            finally:
                timed("host", ts)

        def device_build(req):
            ts = time.perf_counter()
            try:
                req.ana.dev = DeviceModel(req.ana.model, device=self.device)
                if req.kwargs.get("precond") == "two_level":
                    req.ana.dev.set_coarse()   # aggregates built here, off the serving loop
                if self.start_on_workers:
                    # stage set-up and the first mount_K, off the serving loop
                    req.job.prestart()
                return req, None
            except Exception as e:
                return req, e
            finally:
                timed("device", ts)

        def resume_prep(req):
            # the model left its pool (release_left), so its work runs on its
            # own stream, which no other thread uses meanwhile
            ts = time.perf_counter()
            try:
                it = req.job.next_solve(req.ana.dev.nr_finish_end(), attach=False)
                if it is not None:
                    req.ana.dev.nr_prepare(it)
                return req, None
            except Exception as e:
                return req, e
            finally:
                timed("resume", ts)

        nxt = 0
        building, creating, resuming = set(), set(), set()
        admit_q = collections.deque()
        arrived = []   # arrived, host build not started (largest_first only)
        running = {}
        pending = collections.deque()   # (request, error): NR update queued, not yet completed
        ndone = 0
        lpt = self.order == "largest_first"
        key = lambda r: (-(r.work or 0.0), r.t_arrival, r.rid)  # noqa: E731
        try:
            while ndone < N:
                now = clock()
                while nxt < N and reqs[nxt].t_arrival <= now:
                    if lpt:
                        arrived.append(reqs[nxt])
                    else:
                        building.add(ex.submit(host_build, reqs[nxt]))
                    nxt += 1
                if lpt and arrived:
                    # largest first among the arrived ones, one build per worker
                    # at a time so that a later, larger arrival can still pass
                    arrived.sort(key=key)
                    while arrived and len(building) < self.workers:
                        building.add(ex.submit(host_build, arrived.pop(0)))
                for f in [f for f in building if f.done()]:
                    building.discard(f)
                    req, ana, kw, est, exc = f.result()
                    if exc is not None:
                        self._finish(req, clock(), exc)
                        ndone += 1
                        continue
                    req.ana, req.kwargs, req.reserve = ana, kw, est
                    admit_q.append(req)
                if lpt and len(admit_q) > 1:
                    admit_q = collections.deque(sorted(admit_q, key=key))
                # admission: FIFO by free HBM budget and member slots
                while admit_q:
                    req = admit_q[0]
                    slots = self.pools.cap - len(running) - len(creating)
                    # finished models wait in the graveyard until an idle
                    # point frees them: their bytes still occupy the device
                    held = self.in_use + self._graveyard_bytes
                    fits = held + req.reserve <= self.budget or (not running and not creating)
                    if slots > 0 and not fits and self._graveyard and self.in_use + req.reserve <= self.budget:
                        self._release()
                        fits = True
                    if slots <= 0 or not fits:
                        break
                    admit_q.popleft()
                    with lock:
                        self.in_use += req.reserve
                        self.peak_in_use = max(self.peak_in_use, self.in_use)
                        self.peak_held = max(self.peak_held, self.in_use + self._graveyard_bytes)
                    req.t_admit = clock()
                    if self.start_on_workers:
                        req.weight = work_weight(req.ana)
                        req.home = self.pools.assign(req.weight)
                        req.job = Job(req.ana, solver_args(device=self.device, **req.kwargs),
                                      self.pools.pools[req.home])
                    creating.add(ex_dev.submit(device_build, req))
                for f in [f for f in creating if f.done()]:
                    creating.discard(f)
                    req, exc = f.result()
                    ts = time.perf_counter()
                    if exc is None:
                        req.bytes = req.ana.dev.device_bytes()
                        self.in_use += req.bytes - req.reserve
                        self.peak_in_use = max(self.peak_in_use, self.in_use)
                        req.t_start = clock()
                        try:
                            if self.start_on_workers:
                                done = req.job.join_started()
                            else:
                                req.weight = work_weight(req.ana)
                                req.home = self.pools.assign(req.weight)
                                req.job = Job(req.ana, solver_args(device=self.device, **req.kwargs),
                                              self.pools.pools[req.home])
                                done = req.job.start()
                        except Exception as e:
                            exc, done = e, True
                    else:
                        done = True
                    if done:
                        self._finish(req, clock(), exc)
                        ndone += 1
                    else:
                        running[req.ana.dev.handle.value] = req
                    self.t_loop["start"] += time.perf_counter() - ts
                if pending or resuming or self.pools.size() > 0 or self.pools.in_flight():
                    # collect every pool whose step has finished: queue its
                    # finished members' NR updates, queue its next step, and
                    # leave completing the updates to the passes below, one
                    # member per pass, so that a pool is relaunched as soon as
                    # its step ends, not after the host has completed another
                    # pool's members
                    self.pools.launch_idle()
                    ts = time.perf_counter()
                    fins = self.pools.collect(block=not pending and not resuming)
                    self.t_loop["step"] += time.perf_counter() - ts
                    ts = time.perf_counter()
                    for k, fin in fins:
                        # one batched NR finish (update kernels) per pool
                        done_reqs = [running.pop(h) for h, _ in fin]
                        try:
                            if done_reqs:
                                devs = [r.ana.dev for r in done_reqs]
                                type(devs[0]).nr_finish_begin_many(devs, [si for _, si in fin])
                                if self.resume_on_workers:
                                    self.pools.pools[k].release_left()
                                    resuming.update(ex_res.submit(resume_prep, r) for r in done_reqs)
                                else:
                                    pending.extend((r, None) for r in done_reqs)
                        except Exception as e:
                            pending.extend((r, e) for r in done_reqs)
                        self.pools.relauncher(k)()
                    ready = [f for f in resuming if f.done()]
                    if not fins and not ready and not pending and resuming:
                        # nothing to do until a worker or a pool step ends
                        concurrent.futures.wait(resuming, timeout=2e-4,
                                                return_when=concurrent.futures.FIRST_COMPLETED)
                        ready = [f for f in resuming if f.done()]
                    for f in ready:
                        resuming.discard(f)
                        req, exc = f.result()
                        done = True
                        if exc is None:
                            try:
                                done = req.job.join_started()
                            except Exception as e:
                                exc = e
                        if done:
                            self._finish(req, clock(), exc)
                            ndone += 1
                        else:
                            running[req.ana.dev.handle.value] = req
                    if pending:
                        req, exc = pending.popleft()
                        done = True
                        if exc is None:
                            try:
                                done = req.job.resume_end()
                            except Exception as e:
                                exc = e
                        if done:
                            self._finish(req, clock(), exc)
                            ndone += 1
                        else:
                            running[req.ana.dev.handle.value] = req
                    self.t_loop["resume"] += time.perf_counter() - ts
                else:
                    if self._graveyard:
                        self._release()
                    ts = time.perf_counter()
                    pend = building | creating
                    wait = (reqs[nxt].t_arrival - clock()) if nxt < N else None
                    if arrived:
                        wait = 0.0
                    if pend:
                        concurrent.futures.wait(pend, timeout=wait if wait is not None and wait > 0 else None,
                                                return_when=concurrent.futures.FIRST_COMPLETED)
                    elif wait is not None and wait > 0:
                        time.sleep(wait)
                    self.t_loop["wait"] += time.perf_counter() - ts
        finally:
            try:
                self.pools.end_all()
            finally:
                ex.shutdown(wait=True)
                ex_dev.shutdown(wait=True)
                ex_res.shutdown(wait=True)
                self._release()
                self._closer.shutdown(wait=True)
                self._closer = None
        return reqs


def latency_stats(reqs):
    lat = np.array([r.latency for r in reqs if r.latency is not None])
    if lat.size == 0:
        return {}
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max())}
