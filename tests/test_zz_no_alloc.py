"""The DOF map entry point that allocates nothing (amr_dof_map_noalloc)
against amr_dof_map, through the C ABI.

Its work space is a static device array shared by every call on the GPU,
holding the last call's contents; two calls in a row (the second smaller)
and a call without the active-node flags must each give amr_dof_map's eq ids
and sizes exactly.  Sizes beyond its limit are refused before any device
work (checked on the CPU in test_abi.py).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def dof_map_host_api(L, _lib, active, presc):
    nnodes = presc.size // 3
    eq = np.empty(nnodes * 3, dtype=np.int32)
    nd, nu = ctypes.c_int64(), ctypes.c_int64()
    _lib.check(L.amr_dof_map(0, nnodes, None if active is None else _lib.ptr(active), _lib.ptr(presc),
                             _lib.ptr(eq), ctypes.byref(nd), ctypes.byref(nu)))
    return eq, nd.value, nu.value


def dof_map_noalloc(L, _lib, torch, active, presc):
    nnodes = presc.size // 3
    da = None if active is None else torch.from_numpy(active).cuda()
    dp = torch.from_numpy(presc).cuda()
    deq = torch.full((nnodes * 3,), -7, dtype=torch.int32, device="cuda")
    dsz = torch.full((2,), -7, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    _lib.check(L.amr_dof_map_noalloc(0, nnodes, None if da is None else ctypes.c_void_p(da.data_ptr()),
                                     ctypes.c_void_p(dp.data_ptr()), ctypes.c_void_p(deq.data_ptr()),
                                     ctypes.c_void_p(dsz.data_ptr()), ctypes.c_void_p(stream.cuda_stream)))
    stream.synchronize()
    return deq.cpu().numpy(), dsz.cpu().tolist()


@pytest.mark.parametrize("with_active", [True, False])
def test_dof_map_noalloc_equals_dof_map(amaru, with_active):
    import torch
    from amaru_amd import _lib
    L = _lib.lib()
    # two calls in a row on the shared work space, the second smaller; the
    # first spans several hundred tiles of 256 dofs with ragged ends
    for nnodes, seed in ((400_003, 0), (1_001, 1)):
        rng = np.random.default_rng(seed)
        active = (rng.random(nnodes) < 0.9).astype(np.uint8) if with_active else None
        presc = (rng.random(nnodes * 3) < 0.2).astype(np.uint8)
        eq_ref, nd, nu = dof_map_host_api(L, _lib, active, presc)
        eq, sizes = dof_map_noalloc(L, _lib, torch, active, presc)
        assert sizes == [nd, nu]
        assert np.array_equal(eq, eq_ref)
        if with_active:
            assert (eq_ref == -1).any()
        assert nu > 0 and nd > nu


def test_dof_map_noalloc_of_an_analysis(amaru):
    """The eq ids of a configured model (configure_dofs!, src/bc.jl:198-233)."""
    import torch
    from amaru_amd import _lib
    L = _lib.lib()
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1, 1, 1]], nx=6, ny=5, nz=4, cellshape="HEX20"))
    model = amaru.FEModel(mesh, [("all", amaru.MechSolid, amaru.LinearElastic(E=1e9, nu=0.3))])
    bcs = [("z==0", amaru.SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", amaru.SurfaceBC(ux=0))]
    ndofs, nu = model.configure_dofs(bcs)
    presc = np.zeros((mesh.nnodes, 3), dtype=np.uint8)
    presc[model.eq >= nu] = 1
    eq, sizes = dof_map_noalloc(L, _lib, torch, None, presc.ravel())
    assert sizes == [ndofs, nu]
    assert np.array_equal(eq.reshape(-1, 3), model.eq)
