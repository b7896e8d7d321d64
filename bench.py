"""Benchmark: DOF-solves/s of the FE assembly + solve path on one MI355X.

Workload (one GPU): linear elastic block 1x1x1 m, 40x40x40 HEX20 with the
3x3x3 rule, E = 30 GPa, nu = 0.2, base clamped, 1 MPa pressure on the top
face (811,923 DOFs, 797,040 unknowns).  One step = one linear static solve:
mount_K (element stiffness on the FP64 matrix cores + assembly) +
solve_system! (PCG with FP32-stored K, FP64 gather and accumulation and
FP64 residual replacement, to relative residual 1e-10, preconditioned by
Jacobi plus a rigid-body-mode coarse space whose operator is formed from K
in every step) + update_state! (IP stresses and internal forces), all
device resident.
value = unknown DOFs solved per second (DOF-solves/s).  The plain FP64
path (FP64 Jacobi-PCG, same tolerance) is timed beside it on the same
model, and the oracle (CPU restatement: mount_K rule + SuperLU) on a
bounded sample of the same problem family.

Two more legs are reported beside the headline:
* "batched": independent Tet10 cantilevers (seeded; element counts
  log-uniform 500-20000), each one linear static analysis; the plain FP64
  path solves them one after another, the batched path solves all of them
  with their PCGs pooled (one launch per PCG phase for every model) in
  mixed precision.  Models are resident on the device before timing.
* "served": a seeded request stream (Hex8/Tet10 boxes of 1e3-3e5 DOFs, 70 %
  elastic, 30 % von Mises in 5 increments) with Poisson arrivals; the plain
  path serves one request at a time (its throughput fixes the arrival
  rate, a multiple of it); the server admits requests by HBM budget and
  batches their solves continuously.  Latency = arrival to result on host,
  host-side model construction included.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
SPMV_KERNEL = {"mixed": "pcg_spmv_kernel<L, float, float, float>",
               "mixed64acc": "pcg_spmv_kernel<L, float, double, double>",
               "fp64": "pcg_spmv_kernel<L, double, double, double>"}


def load_pkg():
    import importlib.util
    if "amaru_amd" in sys.modules:
        return sys.modules["amaru_amd"]
    pkg = os.path.join(ROOT, "amaru.jl_amd")
    spec = importlib.util.spec_from_file_location("amaru_amd", os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["amaru_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def block_problem(am, n, shape="HEX20", nips=27):
    mesh = am.Mesh(am.Block([[0, 0, 0], [1, 1, 1]], nx=n, ny=n, nz=n, cellshape=shape))
    model = am.FEModel(mesh, [("all", am.MechSolid, am.LinearElastic(E=30e9, nu=0.2))], nips=nips)
    bcs = [("z==0", am.SurfaceBC(ux=0, uy=0, uz=0)), ("z==1", am.SurfaceBC(tz=-1e6))]
    ndofs, nu = model.configure_dofs(bcs)
    U, F = model.bc_vals(bcs)
    return mesh, model, bcs, U, F


def measured_traffic(precision, n):
    """HBM bytes per SpMV launch from the committed rocprofv3 --pmc summary of
    the same kernel and workload (FETCH_SIZE + WRITE_SIZE, calibrated on a
    copy kernel), or None."""
    path = os.path.join(ROOT, "profiles", "pcg_spmv_hex20_40cube_pmc_traffic.md")
    if precision != "mixed64acc" or n != 40 or not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1]), os.path.relpath(path, ROOT)
    return None, None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_steps(dev, opts, k):
    lat = []
    its = []
    r = None
    for _ in range(k):
        t0 = time.perf_counter()
        dev.increment_begin(1.0, 1.0, False)
        r = dev.nr_iteration(1, opts)
        lat.append(time.perf_counter() - t0)
        if r["status"] != 0:
            raise RuntimeError(r["message"])
        its.append((r["pcg_iterations"], r["replacements"], int(r["fallback"])))
    return lat, its, r


def cpu_baseline(n_cpu, threads):
    """The oracle (C mount_K rule with OpenMP + SuperLU on K11) on a bounded
    sample: the same problem family at n_cpu^3 HEX20 elements."""
    sys.path.insert(0, ROOT)
    from oracle import HEX20 as O20, fem
    import scipy.sparse.linalg as spla
    mesh_X, mesh_C = fem.block_mesh((0, 0, 0), (1, 1, 1), n_cpu, n_cpu, n_cpu, O20)
    m = fem.Model(mesh_X, mesh_C, O20, [("LinearElastic", (30e9, 0.2))], nips=27)
    obcs = [("z==0", fem.SurfaceBC(ux=0, uy=0, uz=0)), ("z==1", fem.SurfaceBC(tz=-1e6))]
    ndofs, nu = m.configure_dofs(obcs)
    U, F = m.bc_vals(obcs)
    t0 = time.perf_counter()
    K = m.mount_K(nthreads=threads)
    t1 = time.perf_counter()
    lu = spla.splu(K[:nu, :nu].tocsc())
    x = lu.solve(F[:nu] - K[:nu, nu:] @ U[nu:])
    t2 = time.perf_counter()
    U[:nu] = x
    m.update_state(U, m.state)
    t3 = time.perf_counter()
    R = K[nu:, :] @ U   # reactions K21 U1 + K22 U2
    return dict(nu=nu, ndofs=ndofs, t_assemble=t1 - t0, t_solve=t2 - t1, t_update=t3 - t2, total=t3 - t0,
                U=U, R=R)


def ke_probe(am, args):
    """mount_K of an NR iteration on the block with each element-stiffness
    path, five times each after one untimed call (for a kernel trace), and
    the largest relative Frobenius difference of the two paths' Ke."""
    import ctypes
    mesh, model, bcs, Uex, Fex = block_problem(am, args.n)
    dev = am.DeviceModel(model, device=0)
    dev.stage_begin(Uex * 0, Fex)
    dev.increment_begin(1.0, 1.0, False)
    am.lib()
    hip = ctypes.CDLL("libamdhip64.so.7")   # the runtime the library loaded
    out, Ke = {}, {}
    for path in ("exact", "mfma"):
        dev.set_ke_path(path)
        dev.nr_prepare(1)
        assert hip.hipDeviceSynchronize() == 0
        t = time.perf_counter()
        for _ in range(5):
            dev.nr_prepare(1)
        assert hip.hipDeviceSynchronize() == 0
        out[path] = (time.perf_counter() - t) / 5 * 1e3
        Ke[path] = dev.Ke()
        log(f"ke-probe: {path}: {out[path]:.2f} ms per mount_K")
    d = np.linalg.norm((Ke["mfma"] - Ke["exact"]).reshape(len(Ke["exact"]), -1), axis=1)
    d /= np.linalg.norm(Ke["exact"].reshape(len(Ke["exact"]), -1), axis=1)
    log(f"ke-probe: max relative Frobenius difference {d.max():.3e}")
    print(json.dumps({"mount_K_ms": out, "max_rel_frobenius": float(d.max())}))
    dev.close()


def direct_check(am, args, cb):
    """The headline step's modes on the CPU sample's block (the largest one
    the bench's SuperLU factorises in bounded time), against that FP64 direct
    solve: relative L2 error of the displacements and of the reactions."""
    mesh, model, bcs, Uex, Fex = block_problem(am, args.cpu_n)
    nu = model.nu
    assert nu == cb["nu"]
    dev = am.DeviceModel(model, device=0)
    dev.stage_begin(Uex, Fex)
    out = {"block": f"HEX20 {args.cpu_n}^3, 27 IP", "unknowns": nu, "oracle": "SuperLU (FP64 direct) on K11"}
    modes = {"headline": dict(precision=args.precision, precond=args.precond),
             "plain_fp64": dict(precision="fp64", precond="jacobi"),
             "mixed_f64acc_two_level": dict(precision="mixed64acc", precond="two_level")}
    for name, kw in modes.items():
        # the element stiffness each mode has in the timed steps
        dev.set_ke_path("exact" if name == "plain_fp64" else args.ke)
        r = run_steps(dev, dev.opts(rtol=args.rtol, **kw), 1)[2]
        dU, dF = dev.stage_get(last=True)
        out[name] = {"mode": dict(kw, element_stiffness="exact" if name == "plain_fp64" else args.ke),
                     "rel_err_U": rel_err(dU[:nu], cb["U"][:nu]),
                     "rel_err_reactions": rel_err(dF[nu:], cb["R"]), "pcg_iterations": r["pcg_iterations"]}
    dev.close()
    return out


def dof_solves(anas):
    return sum(a.stats["dof_solves"] for a in anas)


class phase:
    """A roctx range around a leg of the bench (a no-op without a profiler),
    so that a kernel trace can be split by leg."""
    _lib = None

    def __init__(self, name):
        self.name = name.encode()

    def __enter__(self):
        import ctypes
        if phase._lib is None:
            phase._lib = ctypes.CDLL("librocprofiler-sdk-roctx.so.1")
        phase._lib.roctxRangePushA(ctypes.c_char_p(self.name))

    def __exit__(self, *exc):
        phase._lib.roctxRangePop()
        return False


def rel_err(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def batched_leg(am, args):
    """Tet10 cantilever batch: plain FP64 one at a time vs pooled mixed precision."""
    from amaru_amd import pool as P, workloads as W
    params = W.cantilever_params(args.batch_seed, args.batch)
    t0 = time.perf_counter()

    def build(two_level=False):
        """Models resident on the device: mesh, DOFs, sparsity pattern and, for
        the two-level preconditioner, its aggregates (all mesh-only data)."""
        anas = []
        for p in params:
            ana = W.cantilever(p)
            ana.model.configure_dofs(ana.stages[0].bcs)
            ana.dev = am.DeviceModel(ana.model)
            if two_level:
                ana.dev.set_coarse()
            anas.append(ana)
        return anas

    plain_set = build()
    log(f"batched: {len(params)} models built in {time.perf_counter() - t0:.1f}s, "
        f"{sum(a.model.nu for a in plain_set)} unknowns in total")
    nplain = min(args.batch_plain, len(plain_set)) if args.batch_plain > 0 else len(plain_set)
    t1 = time.perf_counter()
    for ana in plain_set[:nplain]:
        ok, msg = am.solve(ana, precision="fp64", precond="jacobi", pcg_rtol=args.rtol)
        if not ok:
            raise RuntimeError(msg)
    t_plain = time.perf_counter() - t1
    plain_rate = dof_solves(plain_set[:nplain]) / t_plain
    Uplain = [a.model.uvals.copy() for a in plain_set[:nplain]]
    for a in plain_set:
        a.dev.close()
    log(f"batched: plain fp64, {nplain} models one at a time: {t_plain:.2f}s, {plain_rate:.4g} DOF-solves/s")
    res = {}
    for prec, pc in (("fp64", "jacobi"), ("mixed64acc", "jacobi"), ("fp64", "two_level"), ("mixed64acc", "two_level")):
        key = f"{prec}_{pc}"
        warm = [W.cantilever(params[0])]
        P.run_analyses(warm, precision=prec, precond=pc, pcg_rtol=args.rtol)  # load the kernels
        warm[0].dev.close()
        anas = build(two_level=pc == "two_level")
        pool = P.PoolSet(args.batch_pools, 0, max_members=len(anas), check_every=32)
        t2 = time.perf_counter()
        with phase(f"batched {key}"):
            out = P.run_analyses(anas, pool=pool, precision=prec, precond=pc, pcg_rtol=args.rtol)
        t_b = time.perf_counter() - t2
        if not all(ok for ok, _ in out):
            raise RuntimeError(f"batched {key}: {[m for ok, m in out if not ok][:3]}")
        errs = [rel_err(a.model.uvals, u) for a, u in zip(anas, Uplain)]
        res[key] = {"value": dof_solves(anas) / t_b, "seconds": t_b, "max_rel_err_vs_plain": max(errs),
                    "pcg_iterations_total": int(sum(a.stats["pcg_iterations"] for a in anas)),
                    "pcg_iterations_max": max(a.stats["pcg_iterations"] for a in anas),
                    "pools": pool.stats()}
        log(f"batched {key}: {t_b:.2f}s, {res[key]['value']:.4g} DOF-solves/s, max err vs plain {max(errs):.2e}, "
            f"PCG its total {res[key]['pcg_iterations_total']}")
        # error against the FP64 direct solve on a bounded sample of the models
        if prec == "mixed64acc":
            import scipy.sparse.linalg as spla
            derr = []
            order = np.argsort([a.model.nu for a in anas])
            for i in order[:: max(1, len(order) // args.batch_direct)][:args.batch_direct]:
                a = anas[int(i)]
                m = a.model
                ndofs, nu = m.configure_dofs(a.stages[0].bcs)
                U, F = m.bc_vals(a.stages[0].bcs)
                pool.detach(a.dev)
                a.dev.assemble()
                K = a.dev.K_csc().tocsc()
                x = spla.splu(K[:nu, :nu].tocsc()).solve(F[:nu] - K[:nu, nu:] @ U[nu:])
                eq = m.eq.ravel()
                Ug = np.zeros(ndofs)
                Ug[eq[eq >= 0]] = m.uvals.ravel()[eq >= 0]
                derr.append(rel_err(Ug[:nu], x))
            res[key]["max_rel_err_vs_fp64_direct"] = max(derr)
            res[key]["direct_sample"] = len(derr)
        for a in anas:
            pool.forget(a.dev)
            a.dev.close()
        pool.close()
    best_key = max(res, key=lambda k: res[k]["value"])
    best = res[best_key]
    return {"workload": f"{len(params)} Tet10 cantilevers 1x0.1x0.1 m (seed {args.batch_seed}), 500-20000 "
                        "elements log-uniform, E 10-200 GPa, nu 0.2-0.35, tip load +-1 kN, 4 IP",
            "unknowns_total": int(sum(a.model.nu for a in plain_set)),
            "plain_fp64": {"value": plain_rate, "seconds": t_plain, "models": nplain,
                           "pcg_iterations_total": int(sum(a.stats["pcg_iterations"] for a in plain_set[:nplain]))},
            "pooled": res, "best": best_key, "value": best["value"],
            "speedup_vs_plain_fp64": best["value"] / plain_rate,
            "speedup_breakdown": {k: v["value"] / plain_rate for k, v in res.items()}}


def footing_leg(am, args):
    """Elastoplastic footing: the full Newton-Raphson analysis (prescribed
    settlement in 20 increments, auto-increment by default as the reference's
    remedy for non-convergence, tangent assembly and return mapping every NR
    iteration) on the plain FP64 path and on the mixed two-level path.  The
    perfectly plastic soil may stop converging before the full settlement:
    then the pseudo-time reached and the cut-backs are reported."""
    from amaru_amd import workloads as W
    out = {}
    autoinc = not args.footing_fixed
    for name, prec, pc in (("plain_fp64", "fp64", "jacobi"), ("mixed64acc_two_level", "mixed64acc", "two_level")):
        ana, kw = W.footing(n=args.footing, settlement=args.footing_settlement)
        if autoinc:
            kw = dict(kw, autoinc=True)
        ana.model.configure_dofs(ana.stages[0].bcs)
        ana.dev = am.DeviceModel(ana.model)
        if pc == "two_level":
            ana.dev.set_coarse()
        t0 = time.perf_counter()
        ok, msg = am.solve(ana, precision=prec, precond=pc, pcg_rtol=args.rtol, **kw)
        dt = time.perf_counter() - t0
        st = ana.stats
        out[name] = {"value": st["dof_solves"] / dt, "seconds": dt, "completed": bool(ok),
                     "T_reached": st["T"], "settlement_reached_m": st["T"] * args.footing_settlement,
                     "increments": st["increments"], "cutbacks": st["cutbacks"],
                     "nr_iterations": st["nr_iterations"], "pcg_iterations": st["pcg_iterations"],
                     "plastic_ips": int((ana.model.state[:, 13] != 0).sum())}
        if not ok:
            out[name]["message"] = msg.strip()
        out[name + "_U"] = ana.model.uvals.copy()
        log(f"footing {name}: {dt:.2f}s, {'done' if ok else 'stopped'} at T={st['T']:.4f}, "
            f"{out[name]['value']:.4g} DOF-solves/s, increments {st['increments']}, cut-backs {st['cutbacks']}, "
            f"NR {st['nr_iterations']}, PCG {st['pcg_iterations']}")
        ana.dev.close()
    err = rel_err(out.pop("mixed64acc_two_level_U"), out.pop("plain_fp64_U"))
    nu = ana.model.nu
    return {"workload": f"Drucker-Prager footing, soil 10x10x5 m, {2 * args.footing}x{2 * args.footing}x"
                        f"{args.footing} Hex8, 2x2 m footing settled {args.footing_settlement} m in 20 increments"
                        + (", auto-increment" if autoinc else ", fixed increments"),
            "unknowns": nu, **out, "rel_err_mixed_vs_plain": err,
            "speedup_vs_plain_fp64": out["mixed64acc_two_level"]["value"] / out["plain_fp64"]["value"]}


def post_leg(am, dev, model, args):
    """Output post-processing on the solved block: ip_state_vals of every IP
    and nodal_patch_recovery of the 19 output fields.  The first call also
    builds the patches (host, once per mesh); the timed calls are the
    per-output cost.  The oracle's restatement is timed on a smaller block."""
    t0 = time.perf_counter()
    vals, reps = dev.nodal_values()
    first = time.perf_counter() - t0
    ts = []
    for _ in range(args.post):
        t0 = time.perf_counter()
        dev.nodal_values()
        ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    out = {"workload": f"nodal_patch_recovery of 19 IP output fields, HEX20 {args.n}^3 block (27 IP): "
                       f"{model.ne * model.nips} IPs, {model.nnodes} nodes",
           "first_call_s": first, "seconds": t, "nodes_per_s": model.nnodes / t,
           "patch_reach_max": int(reps.max()), "calls": len(ts)}
    log(f"post-processing: first call {first:.3f}s, then {t * 1e3:.2f} ms per call")
    if not args.no_cpu:
        from oracle import HEX20 as O20, recovery as R
        n = args.cpu_n
        mesh_c, model_c, _, _, _ = block_problem(am, n)
        st = np.random.default_rng(0).standard_normal((model_c.ne * model_c.nips, 14))
        t0 = time.perf_counter()
        V = R.ip_values(st, np.zeros(len(st), dtype=int))
        R.nodal_patch_recovery(model_c.coords, model_c.conn, O20, V, 27)
        tc = time.perf_counter() - t0
        out["cpu_oracle"] = {"sample": f"HEX20 {n}^3 (27 IP), {model_c.nnodes} nodes, numpy, 1 process",
                             "seconds": tc, "nodes_per_s": model_c.nnodes / tc}
        log(f"post-processing oracle on {n}^3: {tc:.2f}s")
    return out


def mass_leg(am, dev, model, args):
    """mount_M on the block (rho = 2400): element mass matrices and the
    eps-filtered assembly into the device matrix, timed per call; the
    oracle's restatement is timed on a smaller block."""
    rho = np.full(model.ne, 2400.0)
    dev.assemble_M(rho)           # first call also allocates the mass buffers
    ts = []
    for _ in range(args.mass):
        t0 = time.perf_counter()
        dev.assemble_M(rho)
        ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    out = {"workload": f"mount_M, HEX20 {args.n}^3 block (27 IP), rho 2400: {model.ne} elements",
           "seconds": t, "elements_per_s": model.ne / t, "calls": len(ts)}
    log(f"mass matrix: {t * 1e3:.2f} ms per call")
    if not args.no_cpu:
        from oracle import HEX20 as O20, fem as OF
        _, model_c, bcs_c, _, _ = block_problem(am, args.cpu_n)
        om = OF.Model(model_c.coords, model_c.conn, O20, [("LinearElastic", (30e9, 0.2))], nips=27)
        om.configure_dofs([("z==0", OF.SurfaceBC(ux=0, uy=0, uz=0))])
        t0 = time.perf_counter()
        om.mount_M(2400.0)
        tc = time.perf_counter() - t0
        out["cpu_oracle"] = {"sample": f"HEX20 {args.cpu_n}^3 (27 IP), {model_c.ne} elements, C (1 thread) + "
                                       "triplet assembly", "seconds": tc, "elements_per_s": model_c.ne / tc}
        log(f"mass matrix oracle on {args.cpu_n}^3: {tc:.2f}s")
    return out


def meshgen_leg(am, args):
    """The block's mesh (split_block!) and DOF map (configure_dofs!) built on
    the device against the host restatement, timed per call (arrays returned
    to the host in both)."""
    bl = am.Block([[0, 0, 0], [1, 1, 1]], nx=args.n, ny=args.n, nz=args.n, cellshape="HEX20")
    bcs = [("z==0", am.SurfaceBC(ux=0, uy=0, uz=0)), ("z==1", am.SurfaceBC(tz=-1e6))]
    am.Mesh(bl, device=0)
    td, th = [], []
    for _ in range(args.meshgen):
        t0 = time.perf_counter()
        md = am.Mesh(bl, device=0)
        td.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        mh = am.Mesh(bl)
        th.append(time.perf_counter() - t0)
    same = bool(np.array_equal(md.coords, mh.coords) and np.array_equal(md.conn, mh.conn))
    a = am.FEModel(mh, [("all", am.LinearElastic(E=30e9, nu=0.2))], nips=27)
    b = am.FEModel(md, [("all", am.LinearElastic(E=30e9, nu=0.2))], nips=27)
    t0 = time.perf_counter()
    a.configure_dofs(bcs)
    t_dof_h = time.perf_counter() - t0
    t0 = time.perf_counter()
    b.configure_dofs(bcs, device=0)
    t_dof_d = time.perf_counter() - t0
    out = {"workload": f"HEX20 {args.n}^3 block: {mh.nnodes} nodes, {mh.ne} elements",
           "mesh_device_s": float(np.median(td)), "mesh_host_s": float(np.median(th)),
           "dofs_device_s": t_dof_d, "dofs_host_s": t_dof_h,
           "bit_identical": same and bool(np.array_equal(a.eq, b.eq))}
    log(f"mesh generation: device {out['mesh_device_s'] * 1e3:.1f} ms, host {out['mesh_host_s'] * 1e3:.1f} ms; "
        f"dof map device {t_dof_d * 1e3:.1f} ms, host {t_dof_h * 1e3:.1f} ms; identical {out['bit_identical']}")
    return out


def loads_leg(am, model, args):
    """Body force with an expression value over every element of the block
    and a traction on its top facets, on the device against the host
    restatement (force vectors returned to the host in both)."""
    from amaru_amd.expr import compile_expr
    val = "-2400*9.81*(1 + 0.1*sin(3*x)*cos(2*y))"
    elems = np.arange(model.ne)
    f = compile_expr(val)
    model._body_forces(elems, "wz", f, val, device=0)
    td = []
    for _ in range(args.loads):
        t0 = time.perf_counter()
        Fd = model._body_forces(elems, "wz", f, val, device=0)
        td.append(time.perf_counter() - t0)
    t0 = time.perf_counter()
    Fh = model._body_forces(elems, "wz", f)
    th = time.perf_counter() - t0
    err = float(np.abs(Fd - Fh).max() / np.abs(Fh).max())
    _, fn = model.mesh.facets_where(model.coords[:, 2] == 1.0)
    from amaru_amd import model as M
    C = model.coords[fn]
    t0 = time.perf_counter()
    Td = M.distributed_forces_device(0, C, "tn", "-1e6*(1 + x*y)")
    t_tr_d = time.perf_counter() - t0
    t0 = time.perf_counter()
    Th = M.distributed_forces(C, "tn", "-1e6*(1 + x*y)")
    t_tr_h = time.perf_counter() - t0
    out = {"workload": f"body force wz = {val} on {model.ne} HEX20 elements (27 IP); traction tn on {len(fn)} "
                       "QUAD8 top facets",
           "body_device_s": float(np.median(td)), "body_host_s": th, "body_rel_diff": err,
           "traction_device_s": t_tr_d, "traction_host_s": t_tr_h,
           "traction_rel_diff": float(np.abs(Td - Th).max() / np.abs(Th).max())}
    log(f"loads: body device {out['body_device_s'] * 1e3:.1f} ms, host {th * 1e3:.0f} ms (rel diff {err:.1e}); "
        f"traction device {t_tr_d * 1e3:.1f} ms, host {t_tr_h * 1e3:.1f} ms")
    return out


def dynamic_leg(am, args):
    """The dynamic stage solver (Newmark, Rayleigh damping) on a HEX20
    cantilever, on the device; the oracle's restatement (scipy, SuperLU) on a
    smaller one beside it."""
    from amaru_amd import dynamic as D

    def build(n):
        mesh = am.Mesh(am.Block([[0, 0, 0], [2.0, 0.25, 0.25]], nx=8 * n, ny=n, nz=n, cellshape="HEX20"))
        model = am.FEModel(mesh, [("all", am.MechSolid(rho=7850.0), am.LinearElastic(E=200e9, nu=0.3))])
        bcs = [("x==0", am.SurfaceBC(ux=0, uy=0, uz=0)), ("x==2", am.SurfaceBC(tz=-1e4))]
        return mesh, model, bcs
    n, nincs = args.dynamic, 10
    mesh, model, bcs = build(n)
    ana = D.DynamicAnalysis(model)
    D.addstage(ana, bcs, nincs=nincs, tspan=5e-3)
    t0 = time.perf_counter()
    ok, msg = D.dyn_solve(ana, alpha=20.0, beta=1e-5, pcg_rtol=1e-10)
    dt = time.perf_counter() - t0
    if not ok:
        raise RuntimeError(f"dynamic: {msg}")
    out = {"workload": f"HEX20 cantilever 2 x 0.25 x 0.25 m, {8 * n}x{n}x{n} cells (8 IP), {model.nu} unknowns, "
                       f"{nincs} Newmark increments, Rayleigh damping", "seconds": dt,
           "dof_solves_per_s": model.nu * nincs / dt, "nr_iterations": len(ana.log)}
    ana.dev.close()
    log(f"dynamic: {dt:.2f}s for {nincs} increments, {model.nu} unknowns")
    if not args.no_cpu:
        from oracle import HEX20 as O20, fem as OF
        from oracle.dynamic import dyn_stage_solver
        nc = max(1, n // 4)
        mesh_c, model_c, _ = build(nc)
        om = OF.Model(mesh_c.coords, mesh_c.conn, O20, [("LinearElastic", (200e9, 0.3))])
        obcs = [("x==0", OF.SurfaceBC(ux=0, uy=0, uz=0)), ("x==2", OF.SurfaceBC(tz=-1e4))]
        t0 = time.perf_counter()
        ok, msg = dyn_stage_solver(om, obcs, 7850.0, nincs=nincs, tspan=5e-3, alpha=20.0, beta=1e-5)
        tc = time.perf_counter() - t0
        nu_c = model_c.configure_dofs([("x==0", am.SurfaceBC(ux=0, uy=0, uz=0))])[1]
        out["cpu_oracle"] = {"sample": f"HEX20 {8 * nc}x{nc}x{nc} cells, {nu_c} unknowns, {nincs} increments "
                                       "(scipy + SuperLU, 1 thread)", "seconds": tc,
                             "dof_solves_per_s": nu_c * nincs / tc}
        log(f"dynamic oracle: {tc:.2f}s")
    return out


def served_leg(am, args):
    """Request stream: plain one-at-a-time capacity, then the server at a
    Poisson rate that is a multiple of it."""
    from amaru_amd import pool as P, workloads as W
    params = W.request_params(args.serve_seed, args.serve)

    def builder(p, prec, pc="jacobi"):
        def build():
            ana, kw = W.box_request(p)
            return ana, dict(kw, precision=prec, precond=pc, pcg_rtol=args.rtol)
        return build

    # plain: one request at a time, arrival to result (host construction included)
    t0 = time.perf_counter()
    lat_plain, ds_plain, t_build_plain, U_plain = [], 0, 0.0, []
    for p in params:
        t1 = time.perf_counter()
        ana, kw = builder(p, "fp64")()
        ana.model.configure_dofs(ana.stages[0].bcs)
        ana.dev = am.DeviceModel(ana.model)
        t_build_plain += time.perf_counter() - t1
        ok, msg = am.solve(ana, **kw)
        if not ok:
            raise RuntimeError(msg)
        lat_plain.append(time.perf_counter() - t1)
        ds_plain += ana.stats["dof_solves"]
        U_plain.append(ana.model.uvals.copy())
        ana.dev.close()
    t_plain = time.perf_counter() - t0
    plain_rate = ds_plain / t_plain
    req_rate = len(params) / t_plain
    def serve(arr, order):
        # work hint: DOFs x load increments, as a client would state it
        reqs = [P.Request(i, arr[i], builder(p, args.serve_precision, args.serve_precond),
                          work=p["dofs"] * (5 if p["vm"] else 1)) for i, p in enumerate(params)]
        srv = P.Server(max_members=args.serve_members, hbm_budget=int(args.serve_hbm_gb * 2 ** 30), workers=4,
                       npools=args.serve_pools, order=order)
        try:
            done = srv.serve(reqs)
            pool_stats = srv.pools.stats()
        finally:
            srv.close()
        bad = [r.msg for r in done if not r.ok]
        if bad:
            raise RuntimeError(f"served: {bad[:3]}")
        ds = sum(r.dof_solves for r in done)
        makespan = max(r.t_done for r in done) - min(r.t_arrival for r in done)
        err = {"elastic": 0.0, "von_mises": 0.0}
        for r in done:
            k = "von_mises" if params[r.rid]["vm"] else "elastic"
            err[k] = max(err[k], rel_err(r.ana.model.uvals, U_plain[r.rid]))
        return {"value": ds / makespan, "makespan_s": makespan, "dof_solves": int(ds),
                "max_rel_err_vs_plain": err,
                "latency_s": P.latency_stats(done), "peak_admitted_gb": srv.peak_in_use / 2 ** 30,
                "pool_steps": srv.steps, "pools": pool_stats,
                "main_loop_s": {k: round(v, 4) for k, v in srv.t_loop.items()},
                "worker_s": {k: round(v, 4) for k, v in srv.t_build.items()},
                "phases_mean_s": {"arrival_to_admit": float(np.mean([r.t_admit - r.t_arrival for r in done])),
                                  "admit_to_device_ready": float(np.mean([r.t_start - r.t_admit for r in done])),
                                  "device_ready_to_done": float(np.mean([r.t_done - r.t_start for r in done]))}}

    # a server runs warm: serve the 4 smallest requests once (untimed) so that
    # one-time library initialisation (rocSOLVER's kernels) is not in the run
    small = sorted(range(len(params)), key=lambda i: params[i]["dofs"])[:4]
    srv = P.Server(max_members=args.serve_members, hbm_budget=int(args.serve_hbm_gb * 2 ** 30), workers=4,
                   npools=args.serve_pools)
    try:
        srv.serve([P.Request(k, 0.0, builder(params[i], args.serve_precision, args.serve_precond))
                   for k, i in enumerate(small)])
    finally:
        srv.close()
    # a batch submitted at once is scheduled largest first; a stream in arrival order
    with phase("served capacity"):
        cap = serve(np.zeros(len(params)), "largest_first")
    log(f"served capacity (all requests at t=0): {cap['value']:.4g} DOF-solves/s, "
        f"{cap['value'] / plain_rate:.2f}x plain")
    rate = args.serve_load * req_rate
    with phase("served poisson"):
        lat = serve(W.arrival_times(args.serve_seed, len(params), rate), "fifo")
    log(f"served at {rate:.3g} requests/s: {lat['value']:.4g} DOF-solves/s, p50 {lat['latency_s']['p50_s']:.3f}s "
        f"p99 {lat['latency_s']['p99_s']:.3f}s")
    out = {"workload": f"{len(params)} requests (seed {args.serve_seed}): Hex8/Tet10 boxes 1e3-3e5 DOFs, 70% "
                       "elastic, 30% von Mises 5 increments",
           "precision": args.serve_precision, "precond": args.serve_precond, "hbm_budget_gb": args.serve_hbm_gb,
           "warmup": "the 4 smallest requests served once, untimed, before the runs",
           "server": {"pools": args.serve_pools, "check_every": 16, "host_workers": 4,
                      "order": {"capacity_all_at_t0": "largest_first", "poisson": "fifo"}},
           "value": cap["value"], "speedup_vs_plain_fp64": cap["value"] / plain_rate,
           "capacity_all_at_t0": cap,
           "poisson": dict(lat, arrival_rate_per_s=rate, load_vs_plain_request_rate=args.serve_load),
           "plain_fp64": {"value": plain_rate, "seconds": t_plain, "host_build_s": t_build_plain,
                          "latency_s": {"p50_s": float(np.percentile(lat_plain, 50)),
                                        "p99_s": float(np.percentile(lat_plain, 99))}}}
    return out


def pmc_probe(am, args):
    """One mixed-precision step on the block (its SpMV launches are what a
    counter pass reads) and a plain device copy of 1 GiB, the calibration
    kernel of known traffic (1 GiB read, 1 GiB written)."""
    import ctypes
    mesh, model, bcs, Uex, Fex = block_problem(am, args.n)
    dev = am.DeviceModel(model, device=0)
    dev.set_ke_path(args.ke)
    dev.stage_begin(Uex, Fex)
    run_steps(dev, dev.opts(precision=args.precision, rtol=args.rtol, precond=args.precond), 1)
    dev.close()
    # the copy through the HIP runtime the library uses (a device-to-device
    # hipMemcpy runs the runtime's copy kernel)
    hip = ctypes.CDLL("libamdhip64.so.7")
    n = 1 << 30
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipFree(a) == 0 and hip.hipFree(b) == 0
    log("pmc probe done: one step, one 1 GiB copy")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=40, help="elements per edge (40 = the 811,923-DOF block)")
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--precision", default="mixed64acc", choices=["mixed", "mixed64acc", "fp64"])
    ap.add_argument("--precond", default="two_level", choices=["jacobi", "block_jacobi", "two_level"],
                    help="headline preconditioner (two_level: its coarse operator formed again in every step)")
    ap.add_argument("--ke", default="mfma", choices=["mfma", "exact"],
                    help="element stiffness of the headline step: FP64 matrix cores (Gram form for the linear "
                         "elastic block) or the bit-exact kernel")
    ap.add_argument("--ke-probe", action="store_true",
                    help="time mount_K with each element-stiffness path on the block and exit (for a kernel trace)")
    ap.add_argument("--variant-steps", type=int, default=1)
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--cpu-n", type=int, default=12, help="HEX20 elements per edge of the CPU sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch", type=int, default=256, help="cantilevers in the batched leg (0 = skip)")
    ap.add_argument("--batch-seed", type=int, default=0)
    ap.add_argument("--batch-plain", type=int, default=0, help="models timed on the plain path (0 = all)")
    ap.add_argument("--batch-direct", type=int, default=6, help="models checked against the FP64 direct solve")
    ap.add_argument("--batch-pools", type=int, default=3, help="pools stepping concurrently in the batched leg")
    ap.add_argument("--serve-pools", type=int, default=3, help="pools stepping concurrently in the server")
    ap.add_argument("--serve", type=int, default=48, help="requests in the served leg (0 = skip)")
    ap.add_argument("--footing", type=int, default=30, help="footing leg: n for the (2n x 2n x n) block (0 = skip)")
    ap.add_argument("--footing-settlement", type=float, default=0.1, help="footing settlement in m")
    ap.add_argument("--footing-fixed", action="store_true", help="footing leg with fixed increments (no autoinc)")
    ap.add_argument("--post", type=int, default=3, help="timed nodal-recovery calls on the block (0 = skip)")
    ap.add_argument("--mass", type=int, default=3, help="timed mount_M calls on the block (0 = skip)")
    ap.add_argument("--meshgen", type=int, default=3, help="timed device mesh generations of the block (0 = skip)")
    ap.add_argument("--loads", type=int, default=3, help="timed device body-force integrations on the block (0 = skip)")
    ap.add_argument("--dynamic", type=int, default=8, help="dynamic leg: cross-section cells of the HEX20 cantilever (0 = skip)")
    ap.add_argument("--serve-seed", type=int, default=1)
    ap.add_argument("--serve-load", type=float, default=4.0,
                    help="Poisson arrival rate as a multiple of the plain path's request rate")
    ap.add_argument("--serve-members", type=int, default=64)
    ap.add_argument("--serve-hbm-gb", type=float, default=64.0)
    ap.add_argument("--serve-precond", default="two_level", choices=["jacobi", "block_jacobi", "two_level"])
    ap.add_argument("--serve-precision", default="mixed64acc", choices=["fp64", "mixed", "mixed64acc"])
    ap.add_argument("--pmc-probe", action="store_true",
                    help="run one step and a 1 GiB copy only (for rocprofv3 --pmc passes), print nothing else")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (models would be sharded whole, one GPU "
              "per model, or a single model by node ranges with halo exchange). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    am = load_pkg()
    if args.pmc_probe:
        pmc_probe(am, args)
        return
    if args.ke_probe:
        ke_probe(am, args)
        return
    t0 = time.perf_counter()
    mesh, model, bcs, Uex, Fex = block_problem(am, args.n)
    dev = am.DeviceModel(model, device=0)
    info = dev.info()
    # element tiles timed once (model set-up, untimed); the stiffness of the
    # NR iterations on the FP64 matrix cores unless --ke exact
    tiles = dev.tune_elem_tiles()
    dev.set_ke_path(args.ke)
    if args.precond == "two_level":
        # the coarse operator is formed from K in every step's solve, inside
        # the timed region (K does not change between steps here)
        dev.set_coarse_refresh("assembly")
    dev.stage_begin(Uex, Fex)
    log(f"model: {info['nnodes']} nodes, {info['ne']} elements, ndofs {model.ndofs}, nu {model.nu}, "
        f"{info['nblk']} blocks; setup {time.perf_counter() - t0:.1f}s")
    # headline: FP32-stored K with FP64 gather/accumulation, FP64 residual
    # replacement, two-level preconditioner (--precond); the Jacobi, the
    # literal FP32-accumulation and the nodal block-Jacobi variants are
    # measured beside it
    mixed = dev.opts(precision=args.precision, rtol=args.rtol, precond=args.precond)
    plain = dev.opts(precision="fp64", rtol=args.rtol, precond="jacobi")
    variants = {"mixed_f64acc_jacobi": dev.opts(precision="mixed64acc", rtol=args.rtol, precond="jacobi"),
                "mixed_f32acc_jacobi": dev.opts(precision="mixed", rtol=args.rtol, precond="jacobi"),
                "mixed_f64acc_block_jacobi": dev.opts(precision="mixed64acc", rtol=args.rtol, precond="block_jacobi"),
                "fp64_block_jacobi": dev.opts(precision="fp64", rtol=args.rtol, precond="block_jacobi"),
                "fp64_two_level": dev.opts(precision="fp64", rtol=args.rtol, precond="two_level"),
                "mixed_f64acc_two_level": dev.opts(precision="mixed64acc", rtol=args.rtol, precond="two_level")}

    # accuracy reference: FP64 PCG to 1e-13 (a direct factorisation of this
    # K11, 134M nonzeros, is out of reach for a CPU run)
    r = run_steps(dev, dev.opts(precision="fp64", rtol=1e-13), 1)[2]
    Uref, _ = dev.stage_get(last=True)
    log(f"reference FP64 PCG(1e-13): {r['pcg_iterations']} iterations, relres {r['relres']:.2e}")

    run_steps(dev, mixed, args.warmup)
    dev.timing(True)
    with phase("headline steps"):
        lat, its, r = run_steps(dev, mixed, args.steps)
    spmv_ms, spmv_n = dev.timing(False)
    Umix, _ = dev.stage_get(last=True)
    # the same steps without the per-SpMV events (reported beside the value)
    lat_g, _, _ = run_steps(dev, mixed, args.steps)
    err_mixed = float(np.linalg.norm(Umix - Uref) / np.linalg.norm(Uref))
    log(f"mixed: {np.mean(lat):.3f}s/step, (PCG its, replacements, fallback) {its}, relres {r['relres']:.2e}, "
        f"err {err_mixed:.2e}; {np.mean(lat_g):.3f}s/step without the SpMV events")

    # the plain FP64 path is the drop-in one: bit-exact element stiffness too
    dev.set_ke_path("exact")
    lat_p, its_p, rp = run_steps(dev, plain, args.plain_steps)
    dev.set_ke_path(args.ke)
    Upl, _ = dev.stage_get(last=True)
    err_plain = float(np.linalg.norm(Upl - Uref) / np.linalg.norm(Uref))
    log(f"plain fp64: {np.mean(lat_p):.3f}s/step, PCG its {its_p}, err {err_plain:.2e}")

    extra = {}
    for name, o in variants.items():
        if args.variant_steps <= 0:
            break
        if name == f"mixed_f64acc_{args.precond}" and args.precision == "mixed64acc":
            continue   # the headline itself
        if name.endswith("two_level") and args.precond != "two_level":
            # the coarse operator is formed once per DOF map (first solve),
            # not per step: form it untimed
            run_steps(dev, o, 1)
        lv, iv, rv = run_steps(dev, o, args.variant_steps)
        Uv, _ = dev.stage_get(last=True)
        extra[name] = {"value": model.nu * len(lv) / float(np.sum(lv)), "ms_per_step": 1e3 * float(np.mean(lv)),
                       "pcg": iv, "rel_err": float(np.linalg.norm(Uv - Uref) / np.linalg.norm(Uref))}
        log(f"{name}: {np.mean(lv):.3f}s/step, {iv}")
    lanes = dev.spmv_lanes()
    spmv_col_bytes = dev.spmv_col_bytes()

    nu = model.nu
    total = float(np.sum(lat))
    value = nu * args.steps / total
    plain_value = nu * len(lat_p) / float(np.sum(lat_p))
    nblk, nnodes = info["nblk"], info["nnodes"]
    # 3x3 blocks (FP32 or FP64) + column indices (16-bit offsets or int32) +
    # rowptr + p gather (once, FP32 or FP64) + FP64 q write + FP64 p read for p.q
    vb = 8 if args.precision == "fp64" else 4
    xb = 4 if args.precision == "mixed" else 8
    cb = spmv_col_bytes
    spmv_bytes = (9 * vb + cb) * nblk + 4 * (nnodes + 1) + 3 * xb * nnodes + 24 * nnodes + 24 * nnodes
    avg_s = spmv_ms / 1e3 / max(spmv_n, 1)
    achieved = spmv_bytes / avg_s / 1e9 if spmv_n else None
    traffic, traffic_src = measured_traffic(args.precision, args.n)
    out = {
        "metric": "DOF-solves/s + p50/p99 latency vs plain FP64 GPU path at 1e-8 rel error",
        "value": value,
        "unit": "DOF-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"mixed": "f32 SpMV (f32 accumulation), f32 precond, f64 vectors/dots/residual",
                  "mixed64acc": "f32-stored K gathered/accumulated in f64, f32 precond, f64 vectors/dots/residual",
                  "fp64": "f64"}[args.precision] + "; f64 assembly and IP state",
        "data": "synthetic",
        "config": {"workload": f"HEX20 block {args.n}^3, 27 IP, E=30GPa nu=0.2, base clamped, top pressure 1MPa",
                   "ndofs": model.ndofs, "unknowns": nu, "rtol": args.rtol, "precond": args.precond,
                   "coarse_operator": "formed in every step" if args.precond == "two_level" else None,
                   "step": "mount_K+solve_system!+update_state!", "spmv_lanes": lanes,
                   "element_stiffness": args.ke, "element_tiles": {"ach": tiles[0], "epw": tiles[1]}},
        "variants": extra,
        "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "ms_per_step_without_spmv_events": 1e3 * float(np.mean(lat_g)),
        "plain_fp64": {"value": plain_value, "ms_per_step": 1e3 * float(np.mean(lat_p)),
                       "pcg_iterations": its_p, "rel_err": err_plain},
        "speedup_vs_plain_fp64": value / plain_value,
        "variant_speedups_vs_plain_fp64": {k: v["value"] / plain_value for k, v in extra.items()},
        "pcg_iterations": its,
        "rel_err_vs_fp64_pcg_1e-13": err_mixed,
        "roofline": {"bound": "hbm", "kernel": SPMV_KERNEL[args.precision].replace("L", str(
                         lanes["fp64"] if args.precision == "fp64" else
                         lanes["fp32"] if args.precision == "mixed" else lanes["fp32_f64acc"])), "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                     "frac": achieved / 8000.0 if achieved else None, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "bytes_per_launch": spmv_bytes, "col_index_bytes": cb, "avg_launch_us": avg_s * 1e6,
                     "launches": spmv_n},
    }
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        cb = cpu_baseline(args.cpu_n, threads)
        out["cpu_baseline"] = {"value": cb["nu"] / cb["total"], "unit": "DOF-solves/s", "cores": threads,
                               "kind": "port",
                               "sample": (f"oracle on HEX20 {args.cpu_n}^3 (27 IP) of the same block problem, "
                                          f"{cb['nu']} unknowns: mount_K {cb['t_assemble']:.2f}s "
                                          f"({threads} OpenMP threads), SuperLU factor+solve "
                                          f"{cb['t_solve']:.2f}s (1 thread), update {cb['t_update']:.2f}s")}
        out["rel_err_vs_fp64_direct"] = direct_check(am, args, cb)
        log(f"direct oracle check: {out['rel_err_vs_fp64_direct']}")
    if args.post:
        out["post_processing"] = post_leg(am, dev, model, args)
    if args.mass:
        out["mass_matrix"] = mass_leg(am, dev, model, args)
    if args.meshgen:
        out["mesh_generation"] = meshgen_leg(am, args)
    if args.loads:
        out["distributed_loads"] = loads_leg(am, model, args)
    if args.dynamic:
        out["dynamic"] = dynamic_leg(am, args)
    dev.close()
    for name, leg, n in (("batched", batched_leg, args.batch), ("served", served_leg, args.serve),
                         ("footing", footing_leg, args.footing)):
        if n <= 0:
            continue
        try:
            out[name] = leg(am, args)
        except Exception as e:  # report, keep the headline line
            log(f"{name} leg failed: {e!r}")
            out[name] = {"error": repr(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
