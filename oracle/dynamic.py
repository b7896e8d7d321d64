"""Python restatement of Amaru.jl's dynamic stage solver.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

``dyn_stage_solver`` follows src/mech/dyn-solver.jl:203-539 for the first
stage, no seismic input, with fixed increments or auto-increment (a failed
increment is retried from the last converged state with
dT = round(clamp(1 + tanh(log10(tol/res1)), 0.2, 0.9) dT, 3 digits); after a
converged one the reference's growth factor is 1, because its iteration
counter ``nits`` is never advanced, so dT stays, capped by 1/nincs and the
remaining span): initial accelerations from
``solve_system!(M, A, Fex, nu)``, then per increment Newton-Raphson
iterations on the pseudo-stiffness ``Kp = K + (4/dt^2) M + (2/dt) C`` with
``C = alpha M + beta K``, velocities and accelerations by the average
acceleration rule, and the residual ``Fex - (Fin + C V + M A)`` over the
unknowns.  K and M come from the oracle's mount_K / mount_M, solves from
SuperLU.
"""
import math

import numpy as np

from .fem import EPS, solve_system


def dyn_stage_solver(model, bcs, rho, nincs, tspan, tol=0.01, maxits=5, alpha=0.0, beta=0.0, dTmin=1e-7,
                     dTmax=0.1, autoinc=False, log=None):
    """Returns (ok, message); fills model.uvals, fvals (TFin at the dofs),
    vvals, avals and model.state like the reference's dof values."""
    ndofs, nu = model.configure_dofs(bcs)
    state_bk = model.state.copy()
    state = model.state.copy()
    K = model.mount_K(state).tocsc()
    M = model.mount_M(rho).tocsc()
    A = np.zeros(ndofs)
    V = np.zeros(ndofs)
    Uex, Fex = model.bc_vals(bcs, 0.0)
    msg = solve_system(M, A, Fex, nu)
    if msg:
        return False, msg
    T, dT, t = 0.0, 1.0 / nincs, 0.0
    Tcheck, dTcheck, dTbk = 1.0, 1.0, 0.0
    if autoinc:
        dT = min(dT, dTmax, dTcheck)
    U = np.zeros(ndofs)
    Fin = np.zeros(ndofs)
    Fina = np.zeros(ndofs)
    TFin = np.zeros(ndofs)
    while T < 1.0 - dTmin:
        dt = tspan * dT
        Uex, Fex = model.bc_vals(bcs, t + dt)
        Fex_Fin = Fex - Fina
        dUa = np.zeros(ndofs)
        dUi = Uex.copy()
        residue, nfails, maxfails, converged, res1 = 0.0, 0, 3, False, 0.0
        for it in range(1, maxits + 1):
            if it > 1:
                dUi[:] = 0.0
            lastres = residue
            K = model.mount_K(state).tocsc()
            M = model.mount_M(rho).tocsc()
            C = alpha * M + beta * K
            Kp = K + (4 / dt ** 2) * M + (2 / dt) * C
            dFp = Fex_Fin + M @ (A + 4 * V / dt - 4 * dUa / dt ** 2) + C @ (V - 2 * dUa / dt)
            msg = solve_system(Kp.tocsc(), dUi, dFp, nu)
            if msg:
                return False, msg
            dUt = dUa + dUi
            dFin, state, msg = model.update_state(dUt, state_bk)
            if msg:
                # the reference ignores the status (dyn-solver.jl:389) and
                # goes on with a thread-dependent partial internal force; a
                # failed update counts as an unbounded residue here
                if log is not None:
                    log.append((it, math.inf))
                if it == 1:
                    res1 = math.inf
                break
            Fina = Fin + dFin
            Va = -V + 2 * dUt / dt
            Aa = -A + 4 * (dUt - V * dt) / dt ** 2
            TFin = Fina + C @ Va + M @ Aa
            residue = float(np.max(np.abs((Fex - TFin)[:nu]))) if nu else 0.0
            dUa += dUi
            Fex_Fin = Fex - Fina
            Fex_Fin[nu:] = 0.0
            if log is not None:
                log.append((it, residue))
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
            state = state_bk.copy()
            if autoinc:
                # clamp(1 + tanh(log10(tol/res1)), 0.2, 0.9), dyn-solver.jl:515-516
                q = 0.2 if math.isinf(res1) else min(max(1 + math.tanh(math.log10(tol / (res1 + EPS))), 0.2), 0.9)
                dT = float(f"{q * dT:.3g}")
                if dT < dTmin:
                    model.state = state_bk.copy()
                    return False, "solver did not converge"
                continue
            model.state = state_bk.copy()
            return False, "solver did not converge"
        Fin = Fina
        U += dUa
        state_bk = state.copy()
        A, V = Aa, Va
        T += dT
        t += dt
        if T > Tcheck - dTmin:
            Tcheck += dTcheck
        if autoinc:
            if dTbk > 0.0:
                dT = min(dTbk, Tcheck - T)
                dTbk = 0.0
            else:
                dTtr = min(dT, 1.0 / nincs, 1.0 - T)
                if T + dTtr > Tcheck - dTmin:
                    dTbk = dT
                    dT = Tcheck - T
                else:
                    dT = dTtr
                    dTbk = 0.0
    model.state = state_bk.copy()
    eq = model.eq
    for name, vec in (("uvals", U), ("fvals", TFin), ("vvals", V), ("avals", A)):
        out = np.zeros((model.nnodes, 3))
        has = eq >= 0
        out[has] = vec[eq[has]]
        setattr(model, name, out)
    return True, None
