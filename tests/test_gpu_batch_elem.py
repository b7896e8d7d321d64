"""GPU: the element phase of several models batched into one launch per
kernel family (amr_nr_prepare_many, amr_nr_finish_begin_many) gives the same
bits as each model's own launches.

The set mixes every element kernel the batched launches dispatch to: HEX8
(8 IP), HEX20 (27 and 8 IP) and TET10 (4 IP), linear elastic, von Mises and
Drucker-Prager, the exact stiffness kernel in each of its three tilings and
the matrix-core paths (Gram form for elastic models, HEX20 for plastic
ones), and each update tiling.  The models are on their own streams (the
batch runs on the device's batch stream between events) for the first
mount_K, and on their pool's stream for the finishes and later mount_Ks of
a pooled run (run_analyses), which must equal solving each model alone.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def mixed_set(am):
    le = am.LinearElastic(E=200e9, nu=0.3)
    vm = am.VonMises(E=200e9, nu=0.3, fy=250e6, H=2e9)
    dp = am.DruckerPrager(E=50e9, nu=0.3, alpha=0.2, kappa=2e6, H=1e8)
    clamp = ("x==0", am.SurfaceBC(ux=0, uy=0, uz=0))
    # (shape, cells, nips, material, tip load or deflection, increments, Ke path, (ach, epw) choice)
    spec = [("HEX8", (8, 2, 2), 0, le, am.SurfaceBC(tz=-1e5), 1, "exact", 0),
            ("HEX8", (6, 2, 2), 0, vm, am.SurfaceBC(uz=-0.02), 3, "exact", 1),
            ("HEX20", (4, 2, 2), 27, le, am.SurfaceBC(tz=-1e5), 1, "mfma", 2),
            ("HEX20", (4, 1, 1), 27, vm, am.SurfaceBC(uz=-0.02), 2, "mfma", 0),
            ("HEX20", (5, 2, 1), 8, dp, am.SurfaceBC(uz=-0.002), 2, "exact", 1),
            ("HEX20", (3, 2, 2), 8, le, am.SurfaceBC(ty=1e5), 1, "mfma", 0),
            ("TET10", (6, 2, 2), 0, le, am.SurfaceBC(tz=-2e5), 1, "mfma", 1),
            ("TET10", (5, 2, 2), 0, vm, am.SurfaceBC(uz=-0.015), 3, "exact", 2),
            ("TET10", (4, 1, 1), 0, le, am.SurfaceBC(tz=-1e5), 1, "exact", 0)]
    out = []
    for shape, n, nips, mat, load, nincs, path, tile in spec:
        mesh = am.Mesh(am.Block([[0, 0, 0], [1.0, 0.2, 0.2]], nx=n[0], ny=n[1], nz=n[2], cellshape=shape))
        model = am.FEModel(mesh, [("all", am.MechSolid, mat)], nips=nips)
        ana = am.MechAnalysis(model)
        am.addstage(ana, [clamp, ("x==1", load)], nincs=nincs)
        model.configure_dofs(ana.stages[0].bcs)
        ana.dev = am.DeviceModel(model)
        ana.dev.set_ke_path(path)
        ach_opts, epw_opts = TILES[(shape, model.nips)]
        ana.dev.set_elem_tiles(ach_opts[tile], epw_opts[(tile + 1) % 3])
        out.append(ana)
    return out


# the tilings each element shape offers (stiffness row chunks, update elements per workgroup)
TILES = {("HEX8", 8): ((1, 2, 4), (32, 16, 8)), ("HEX20", 27): ((7, 10, 20), (9, 4, 2)),
         ("HEX20", 8): ((7, 10, 20), (16, 8, 32)), ("TET10", 4): ((2, 5, 10), (32, 16, 64))}


def results(ana):
    m = ana.model
    return m.uvals.copy(), m.fvals.copy(), m.state.copy(), dict(ana.stats)


def test_first_mount_k_batched_equals_alone(amaru):
    """Element stiffness of the first NR iteration: the batched launch (the
    models on their own streams) equals each model's own launch."""
    one, many = mixed_set(amaru), mixed_set(amaru)
    for group in (one, many):
        for ana in group:
            U, F = ana.model.bc_vals(ana.stages[0].bcs)
            ana.dev.stage_begin(U, F)
            ana.dev.increment_begin(1.0 / ana.stages[0].nincs, 1.0, False)
    for ana in one:
        ana.dev.nr_prepare(1)
    amaru.DeviceModel.nr_prepare_many([a.dev for a in many], [1] * len(many))
    for a, b in zip(one, many):
        assert np.array_equal(a.dev.Ke(), b.dev.Ke())
    for ana in one + many:
        ana.dev.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed64acc"])
def test_pooled_mixed_set_equals_alone(amaru, precision):
    """Full analyses (plastic ones with auto-increment): run_analyses
    batches every model's first mount_K, and per pool step the finishing
    models' updates and their next mount_K; U, F, IP states and the
    iteration counts equal solving each model alone."""
    from amaru_amd import pool as P
    alone = []
    for ana in mixed_set(amaru):
        ok, msg = amaru.solve(ana, precision=precision, pcg_rtol=1e-10, pcg_check_every=16, autoinc=True)
        assert ok, msg
        alone.append(results(ana))
        ana.dev.close()
    anas = mixed_set(amaru)
    ps = P.PoolSet(2, 0, max_members=len(anas), check_every=16)
    try:
        out = P.run_analyses(anas, pool=ps, precision=precision, pcg_rtol=1e-10, autoinc=True)
        assert all(ok for ok, _ in out), out
        for ana, ref in zip(anas, alone):
            got = results(ana)
            for g, r in zip(got[:3], ref[:3]):
                assert np.array_equal(g, r)
            assert got[3] == ref[3]
    finally:
        for ana in anas:
            ps.forget(ana.dev)
            ana.dev.close()
        ps.close()
