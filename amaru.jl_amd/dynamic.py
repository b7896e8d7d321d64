"""DynamicAnalysis / solve! for the dynamic mechanical analysis, on the device.

``dyn_solve`` is the reference's ``solve!(ana::DynamicAnalysis; args...)``
(src/mech/dyn-solver.jl:191-200) → ``dyn_stage_solver!`` (:203-539) with
fixed increments (Newmark's average acceleration with Rayleigh damping
C = alpha M + beta K) and, optionally, auto-increment; the hot calls
through the C ABI:

    mount_K(elems, ndofs)            -> DeviceModel.assemble
    mount_M(elems, ndofs)            -> DeviceModel.assemble_M
    K + (4/dt^2) M + (2/dt) C        -> DeviceModel.combine_KM (in K's storage)
    M * v, C * v                     -> DeviceModel.matvec ("M", "K0")
    solve_system!(Kp, dUi, dFp, nu)  -> DeviceModel.solve_system
    update_state!(elems, dUt, dt)    -> DeviceModel.update_state

Supported as the reference runs them: the first stage, no seismic input.
The reference's stage solver assigns its velocity and acceleration vectors
A and V only inside ``if stage.id == 1`` (src/mech/dyn-solver.jl:262-282)
and reads them in every increment (:378), so a second stage stops there with
an undefined-variable error; a second stage is refused here with that
reason.  With
auto-increment a failed increment is retried from the last converged state
with dT = round(clamp(1 + tanh(log10(tol/res1)), 0.2, 0.9) dT, 3 digits);
after a converged one the reference's growth factor is 1 (its iteration
counter ``nits`` is never advanced), so dT stays, capped by 1/nincs and the
remaining span.

A failed constitutive update (a material failure or NaN internal forces)
does not stop the stage: the reference ignores update_state!'s status here
(src/mech/dyn-solver.jl:389) and goes on with the internal forces its
threads had gathered before the failing element, which depends on the
thread count.  Here such an iteration counts as an unbounded residue: the
iteration loop ends, the increment is not converged, the committed states
are restored and, with auto-increment, the step is cut (by the smallest
factor, 0.2, when it happens at the first iteration).
"""
import math

import numpy as np

from ._lib import AmaruError
from .analysis import Stage
from .model import DeviceModel

EPS = 2.220446049250313e-16

DYN_DEFAULTS = dict(tol=0.01, dTmin=1e-7, dTmax=0.1, maxits=5, alpha=0.0, beta=0.0, autoinc=False, sism=False,
                    # device options (additions of this implementation)
                    device=0, precision="fp64", precond="jacobi", pcg_rtol=1e-12)


class DynamicAnalysis:
    """DynamicAnalysis(model) (src/mech/dyn-solver.jl:35-64)."""

    def __init__(self, model):
        self.model = model
        self.stages = []
        self.log = []          # (inc, it, residue)
        self.dev = None
        self.t = 0.0
        self.V = None          # nodal velocities and accelerations (nnodes x 3)
        self.A = None


def addstage(ana, bcs, nincs=1, nouts=0, tspan=0.0):
    """addstage!(ana, bcs; nincs, nouts, tspan) (src/analysis.jl:27-32)."""
    st = Stage(bcs, nincs, nouts)
    st.tspan = float(tspan)
    ana.stages.append(st)
    return st


def dyn_solve(ana, **kwargs):
    """Runs the stages; returns (ok, message)."""
    a = dict(DYN_DEFAULTS)
    unknown = set(kwargs) - set(a)
    if unknown:
        raise TypeError(f"dyn_solve: unknown arguments {sorted(unknown)}")
    a.update(kwargs)
    if a["sism"]:
        # the reference's dyn_stage_solver! reads sism_file, sism_dir, tss and
        # tds without binding them (src/mech/dyn-solver.jl:272-275, 286-287, 349-351):
        # its seismic path stops with an undefined-variable error
        raise NotImplementedError("dyn_solve: seismic input is not implemented")
    if len(ana.stages) != 1:
        raise NotImplementedError("dyn_solve: one stage (the reference's dyn_stage_solver! binds its velocities and "
                                  "accelerations only in stage 1 and fails on a second stage)")
    st = ana.stages[0]
    st.status = "solving"
    ok, msg = _stage(ana, st, a)
    st.status = "done" if ok else "error"
    return ok, msg


def _stage(ana, stage, a):
    model = ana.model
    bcs = stage.bcs
    ndofs, nu = model.configure_dofs(bcs)
    if ana.dev is None:
        ana.dev = DeviceModel(model, device=a["device"])
    dev = ana.dev
    dev.set_dofs()
    sopt = dict(precision=a["precision"], precond=a["precond"], rtol=a["pcg_rtol"])
    tol, maxits, alpha, beta = a["tol"], a["maxits"], a["alpha"], a["beta"]

    # initial accelerations: M A = Fex at t = 0 (dyn-solver.jl:279-290)
    dev.assemble()
    dev.assemble_M()
    A = np.zeros(ndofs)
    V = np.zeros(ndofs)
    Uex, Fex = model.bc_vals(bcs, 0.0)
    dev.combine_KM(0.0, 1.0)                     # K's storage holds M for this solve
    dev.solve_system(A, Fex, **sopt)

    autoinc = a["autoinc"]
    T, dT, t = 0.0, 1.0 / stage.nincs, ana.t
    dTcheck = 1.0 / stage.nouts if stage.nouts > 0 else 1.0
    Tcheck, dTbk = dTcheck, 0.0
    if autoinc:
        dT = min(dT, a["dTmax"], dTcheck)
    U = np.zeros(ndofs)
    Fin = np.zeros(ndofs)
    Fina = np.zeros(ndofs)
    TFin = np.zeros(ndofs)
    inc = 0
    while T < 1.0 - a["dTmin"]:
        dt = stage.tspan * dT
        inc += 1
        Uex, Fex = model.bc_vals(bcs, t + dt)
        Fex_Fin = Fex - Fina
        dUa = np.zeros(ndofs)
        dUi = Uex.copy()
        residue, nfails, maxfails, converged, res1 = 0.0, 0, 3, False, 0.0
        for it in range(1, maxits + 1):
            if it > 1:
                dUi[:] = 0.0
            lastres = residue
            dev.assemble()
            dev.assemble_M()
            # dFp = Fex_Fin + M (A + 4V/dt - 4 dUa/dt^2) + C (V - 2 dUa/dt)
            y1 = A + 4 * V / dt - 4 * dUa / dt ** 2
            y2 = V - 2 * dUa / dt
            dFp = Fex_Fin + dev.matvec("M", y1) + _cmul(dev, alpha, beta, y2, "K")
            dev.combine_KM(1.0 + 2.0 * beta / dt, 4.0 / dt ** 2 + 2.0 * alpha / dt)
            dev.solve_system(dUi, dFp, **sopt)
            dUt = dUa + dUi
            try:
                dFin = dev.update_state(dUt)
            except AmaruError:
                ana.log.append((inc, it, math.inf))
                if it == 1:
                    res1 = math.inf
                break
            Fina = Fin + dFin
            Va = -V + 2 * dUt / dt
            Aa = -A + 4 * (dUt - V * dt) / dt ** 2
            TFin = Fina + _cmul(dev, alpha, beta, Va, "K0") + dev.matvec("M", Aa)
            residue = float(np.max(np.abs((Fex - TFin)[:nu]))) if nu else 0.0
            dUa += dUi
            Fex_Fin = Fex - Fina
            Fex_Fin[nu:] = 0.0
            ana.log.append((inc, it, residue))
            if it == 1:
                res1 = residue
            if residue > tol:
                Fina = Fina - dFin
            if residue < tol:
                converged = True
                break
            if np.isnan(residue):
                break
            if it > 1 and residue > lastres:
                break
            if residue > 0.9 * lastres:
                nfails += 1
            if nfails == maxfails:
                break
        if not converged:
            inc -= 1
            dev.restore()                            # copyto!.(State, StateBk)
            if autoinc:
                dT = float(f"{cut_factor(tol, res1) * dT:.3g}")
                if dT < a["dTmin"]:
                    return False, "solver did not converge"
                continue
            return False, "solver did not converge"
        Fin = Fina
        U += dUa
        dev.commit()
        A, V = Aa, Va
        T += dT
        t += dt
        _store(model, U, TFin, V, A)
        if T > Tcheck - a["dTmin"]:
            Tcheck += dTcheck
        if autoinc:
            if dTbk > 0.0:
                dT = min(dTbk, Tcheck - T)
                dTbk = 0.0
            else:
                dTtr = min(dT, 1.0 / stage.nincs, 1.0 - T)
                if T + dTtr > Tcheck - a["dTmin"]:
                    dTbk = dT
                    dT = Tcheck - T
                else:
                    dT = dTtr
                    dTbk = 0.0
    ana.t = t
    return True, None


def cut_factor(tol, res1):
    """clamp(1 + tanh(log10(tol/res1)), 0.2, 0.9) (src/mech/dyn-solver.jl:
    515-516); an unbounded first residue gives the smallest factor."""
    if math.isinf(res1):
        return 0.2
    return min(max(1 + math.tanh(math.log10(tol / (res1 + EPS))), 0.2), 0.9)


def _cmul(dev, alpha, beta, x, kname):
    """C x with C = alpha M + beta K (K as assembled this iteration)."""
    y = np.zeros_like(x)
    if alpha != 0.0:
        y += alpha * dev.matvec("M", x)
    if beta != 0.0:
        y += beta * dev.matvec(kname, x)
    return y


def _store(model, U, F, V, A):
    """dof.vals[u, f, v, a] (dyn-solver.jl:432-438), node order."""
    eq = model.eq
    has = eq >= 0
    for name, vec in (("uvals", U), ("fvals", F), ("vvals", V), ("avals", A)):
        out = np.zeros((model.nnodes, 3))
        out[has] = vec[eq[has]]
        setattr(model, name, out)
