"""MechAnalysis / addstage! / solve! on the device (host control flow).

``solve`` is the reference's ``solve!(ana::MechAnalysis; args...)``
(src/mech/mech-solver.jl:155-181) → ``stage_iterator!`` (src/solver.jl:
82-198) → ``mech_stage_solver!`` (src/mech/mech-solver.jl:186-492), with
the hot calls going through the C ABI:

    mount_K(active_elems, ndofs)       -> DeviceModel.assemble      (K stays on the GPU)
    solve_system!(K, ΔUitr, Rtr, nu)   -> DeviceModel.solve_system  (PCG on the GPU)
    copyto!.(State, StateBk)           -> implicit: updates read the committed state
    update_state!(active_elems, ΔUt)   -> DeviceModel.update_state
    copyto!.(StateBk, State)           -> DeviceModel.commit

The increment / Newton-Raphson / auto-increment logic is kept as the
reference has it, with its predictor-corrector schemes :FE (default), :ME,
:BE and :Ralston (src/mech/mech-solver.jl:280-349; the device runs both
solves and updates of an iteration, ``amr_set_scheme``).
"""
import math

import numpy as np

from .model import DeviceModel

EPS = 2.220446049250313e-16

SCHEMES = ("FE", "ME", "BE", "Ralston")   # src/mech/mech-solver.jl:280-288

SOLVER_DEFAULTS = dict(tol=0.01, rspan=0.01, rtol=0.01, dT0=0.01, dTmin=1e-7, dTmax=0.1, scheme="FE", maxits=5,
                       autoinc=False, quiet=True,
                       # device options (additions of this implementation)
                       device=0, precision="fp64", precond="jacobi", pcg_rtol=1e-10, pcg_maxiter=200000,
                       pcg_replace_ratio=0.0, pcg_check_every=32)


class Stage:
    """Stage (src/stage.jl:1-23)."""

    def __init__(self, bcs, nincs=1, nouts=0):
        self.bcs = list(bcs)
        self.nincs = int(nincs)
        self.nouts = int(nouts)
        self.status = "pending"


class MechAnalysis:
    """MechAnalysis(model) (src/mech/mech-solver.jl:43-74)."""

    def __init__(self, model):
        self.model = model
        self.stages = []
        self.loggers = []      # callables f(model, T) after each converged increment
        self.log = []          # (stage, inc, it, residue) like solve.log
        # T: pseudo-time reached in the last stage; cutbacks: increments
        # retried with a smaller step (auto-increment)
        self.stats = dict(nr_iterations=0, pcg_iterations=0, increments=0, dof_solves=0, T=0.0, cutbacks=0)
        self.dev = None
        self.opts = None       # amr_solve_opts of the running stage


def addstage(ana, bcs, nincs=1, nouts=0):
    """addstage!(ana, bcs; nincs, nouts) (src/analysis.jl:27-32)."""
    st = Stage(bcs, nincs, nouts)
    ana.stages.append(st)
    return st


def solve(ana, **kwargs):
    """solve!(ana; args...) → (success, message)."""
    args = solver_args(**kwargs)
    status = None
    gen = analysis_machine(ana, args)
    try:
        it = next(gen)
        while True:
            it = gen.send(ana.dev.nr_iteration(it, ana.opts))
    except StopIteration as stop:
        status = stop.value
    return status


def solver_args(**kwargs):
    unknown = set(kwargs) - set(SOLVER_DEFAULTS)
    if unknown:
        raise TypeError(f"solve!: unknown arguments {sorted(unknown)}")
    args = dict(SOLVER_DEFAULTS, **kwargs)
    scheme = str(args["scheme"]).lstrip(":")
    if scheme not in SCHEMES:
        raise ValueError(f"solve!: scheme must be one of {sorted(SCHEMES)}")
    args["scheme"] = scheme
    if not 1 <= args["maxits"] <= 10:
        raise ValueError("maxits must be in 1..10")
    return args


def analysis_machine(ana, args):
    """stage_iterator! (src/solver.jl:82-198) as a generator.

    It yields the number ``it`` of each Newton-Raphson iteration it needs
    and expects the ``DeviceModel.nr_iteration``-style result of that
    iteration to be sent back; it returns (success, message).  ``solve``
    answers each request at once; a pool runner answers them when the
    model's linear solve, advanced together with other models', is done.
    """
    todo = [s for s in ana.stages if s.status != "done"]
    if not todo:
        raise RuntimeError("stage_iterator!: No stages have been set")
    ok, msg = True, None
    for stage in todo:
        stage.status = "solving"
        nincs = stage.nincs
        if stage.nouts > 0 and not args["autoinc"] and stage.nouts > nincs:
            nincs = stage.nouts
        stage.nincs = nincs
        ok, msg = yield from _stage_solver(ana, stage, args)
        stage.status = "done" if ok else "failed"
        if not ok:
            break
    return ok, msg


def _stage_solver(ana, stage, a):
    model = ana.model
    tol, rtol, rspan = a["tol"], a["rtol"], a["rspan"]
    dTmin, dTmax, maxits, autoinc = a["dTmin"], a["dTmax"], a["maxits"], a["autoinc"]
    bcs = stage.bcs
    ndofs, nu = model.configure_dofs(bcs)
    if ana.dev is None:
        ana.dev = DeviceModel(model, device=a["device"])
    else:
        ana.dev.set_dofs()
    dev = ana.dev
    ana.stats["T"] = 0.0
    dev.set_scheme(a["scheme"])
    Uex, Fex = model.bc_vals(bcs)
    if model.__dict__.get("_state_dev") is dev:
        model.state  # noqa: B018  (the previous stage's states, before this one commits)
    dev.stage_begin(Uex, Fex)
    ana.opts = DeviceModel.opts(precision=a["precision"], rtol=a["pcg_rtol"], maxiter=a["pcg_maxiter"],
                                precond=a["precond"], replace_ratio=a["pcg_replace_ratio"],
                                check_every=a["pcg_check_every"])
    # eq id -> flat (node, dir) index; dof.vals start values of this stage
    flat = np.zeros(ndofs, dtype=np.int64)
    live = model.eq.ravel() >= 0
    flat[model.eq.ravel()[live]] = np.nonzero(live)[0]
    u0, f0 = model.uvals.copy(), model.fvals.copy()

    def sync_nodal():
        U, F = dev.stage_get()
        model.uvals = u0.copy()
        model.fvals = f0.copy()
        model.uvals.ravel()[flat] += U
        model.fvals.ravel()[flat] += F

    nouts = stage.nouts
    dTcheck = 1.0 / nouts if nouts > 0 else 1.0
    Tcheck = dTcheck
    T = 0.0
    dT = 1.0 / stage.nincs
    if autoinc:
        dT = min(dT, dTmax, dTcheck, a["dT0"])
    dTbk = 0.0
    inc = 0
    status = None
    while T < 1.0 - dTmin:
        inc += 1
        dTcr = min(rspan, 1 - T)
        acr = min(dT / dTcr, 1.0)
        dev.increment_begin(dT, acr, T < 1 - rspan)
        res = res1 = 0.0
        converged = syserror = False
        sysmsg = None
        for it in range(1, maxits + 1):
            lastres = res
            r = yield it   # mount_K, solve_system!, update_state!, R, norms
            ana.stats["pcg_iterations"] += r["pcg_iterations"]
            if r["status"] != 0:
                syserror = True
                sysmsg = r["message"]
                break
            ana.stats["nr_iterations"] += 1
            res, err = r["res"], r["err"]
            ana.log.append((len(ana.stages), inc, it, res))
            if it == 1:
                res1 = res
            if res < tol or err < rtol:
                converged = True
                break
            if math.isnan(res):
                break
            if it > 1 and res > lastres:
                break
        if syserror:
            converged = False
        if converged:
            dev.increment_commit()
            ana.stats["increments"] += 1
            ana.stats["dof_solves"] += nu
            T += dT
            ana.stats["T"] = T
            if T > Tcheck - dTmin:
                Tcheck += dTcheck
            if ana.loggers:
                sync_nodal()
                model.state = dev.state_get(committed=True)
                for lg in ana.loggers:
                    lg(model, T)
            if autoinc:
                if dTbk > 0.0:
                    dT = min(dTbk, Tcheck - T)
                    dTbk = 0.0
                else:
                    q = max(1 + math.tanh(math.log10(tol / (res1 + EPS))), 1.1)
                    dTtr = min(q * dT, dTmax, 1 - T)
                    if T + dTtr > Tcheck - dTmin:
                        dTbk = dT
                        dT = Tcheck - T
                    else:
                        dT = dTtr
                        dTbk = 0.0
        else:
            inc -= 1
            dev.increment_reject()
            if autoinc:
                ana.stats["cutbacks"] += 1
                q = 1 + math.tanh(math.log10(tol / (res1 + EPS)))
                q = min(max(q, 0.2), 0.9)
                if syserror:
                    q = 0.7
                dT = float(f"{q * dT:.3g}")
                if dT < dTmin:
                    status = "Solver did not converge."
                    break
            else:
                status = "Solver did not converge. Try `autoinc=true`. " + (sysmsg or "")
                break
    sync_nodal()
    model.state_on_device(dev)
    return status is None, status
