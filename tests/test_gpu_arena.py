"""GPU: model memory from the per-device block cache.

A destroyed model's blocks go back to the cache and the next models take
them: a model built on reused blocks must give the same K, solution and
states as the first (no buffer may rely on fresh memory), the cached bytes
count as free in amr_mem_info, and amr_mem_trim returns them.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run(amaru, shape, n):
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1.0, 0.3, 0.3]], nx=n, ny=2, nz=2, cellshape=shape))
    model = amaru.FEModel(mesh, [("all", amaru.MechSolid, amaru.VonMises(E=200e9, nu=0.3, fy=250e6, H=1e9))])
    bcs = [("x==0", amaru.SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", amaru.SurfaceBC(tz=-2e7))]
    model.configure_dofs(bcs)
    U, F = model.bc_vals(bcs)
    dev = amaru.DeviceModel(model)
    dev.set_coarse(target_nodes=32)
    dev.stage_begin(U, F)
    dev.increment_begin(1.0, 1.0, False)
    r = dev.nr_iteration(1, dev.opts(precision="mixed64acc", precond="two_level", rtol=1e-12))
    assert r["status"] == 0
    dev.assemble()   # mount_K with the reference's pattern at the new state
    K = dev.K_csc()
    dU, dF = dev.stage_get(last=True)
    return dev, K, dU, dF, dev.state_get()


def mem_free(L, _lib):
    f, t = ctypes.c_int64(), ctypes.c_int64()
    _lib.check(L.amr_mem_info(0, ctypes.byref(f), ctypes.byref(t)))
    return f.value


def test_reused_blocks_give_the_same_results(amaru):
    from amaru_amd import _lib
    L = _lib.lib()
    _lib.check(L.amr_mem_trim(0))
    first = {}
    for shape, n in (("HEX20", 8), ("TET10", 6)):
        dev, *out = run(amaru, shape, n)
        first[shape] = out
        dev.close()
    free0 = mem_free(L, _lib)
    # the same models again, in the other order, on the cached blocks
    for shape, n in (("TET10", 6), ("HEX20", 8)):
        dev, K, dU, dF, st = run(amaru, shape, n)
        K0, dU0, dF0, st0 = first[shape]
        assert np.array_equal(K.indptr, K0.indptr) and np.array_equal(K.indices, K0.indices)
        assert np.array_equal(K.data, K0.data)
        assert np.array_equal(dU, dU0) and np.array_equal(dF, dF0) and np.array_equal(st, st0)
        dev.close()
    # cached blocks count as free; trimming keeps them free
    assert abs(mem_free(L, _lib) - free0) <= 64 << 20
    _lib.check(L.amr_mem_trim(0))
    assert abs(mem_free(L, _lib) - free0) <= 64 << 20
