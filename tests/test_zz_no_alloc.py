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


class Hip:
    """Device buffers from the HIP runtime the library itself uses (the
    library is loaded first, so its libamdhip64 is the one found)."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.bufs = []

    def check(self, rc):
        assert rc == 0, f"HIP error {rc}"

    def upload(self, a, fill=None):
        p = ctypes.c_void_p()
        self.check(self.h.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 1))))
        self.bufs.append(p)
        if fill is not None:
            # a byte pattern no result holds, to see that every element is written
            self.check(self.h.hipMemset(p, ctypes.c_int(fill), ctypes.c_size_t(a.nbytes)))
        else:
            self.check(self.h.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1))
        return p

    def download(self, p, a):
        self.check(self.h.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(a.nbytes), 2))
        return a

    def free(self):
        for p in self.bufs:
            self.check(self.h.hipFree(p))
        self.bufs = []


def dof_map_noalloc(L, _lib, active, presc):
    nnodes = presc.size // 3
    hip = Hip()
    try:
        da = None if active is None else hip.upload(active)
        dp = hip.upload(presc)
        eq = np.empty(nnodes * 3, dtype=np.int32)
        sz = np.empty(2, dtype=np.int64)
        deq, dsz = hip.upload(eq, fill=0x7F), hip.upload(sz, fill=0x7F)
        hip.check(hip.h.hipDeviceSynchronize())
        _lib.check(L.amr_dof_map_noalloc(0, nnodes, da, dp, deq, dsz, None))
        hip.check(hip.h.hipDeviceSynchronize())
        return hip.download(deq, eq), hip.download(dsz, sz).tolist()
    finally:
        hip.free()


@pytest.mark.parametrize("with_active", [True, False])
def test_dof_map_noalloc_equals_dof_map(amaru, with_active):
    from amaru_amd import _lib
    L = _lib.lib()
    # two calls in a row on the shared work space, the second smaller; the
    # first spans several hundred tiles of 256 dofs with ragged ends
    for nnodes, seed in ((400_003, 0), (1_001, 1)):
        rng = np.random.default_rng(seed)
        active = (rng.random(nnodes) < 0.9).astype(np.uint8) if with_active else None
        presc = (rng.random(nnodes * 3) < 0.2).astype(np.uint8)
        eq_ref, nd, nu = dof_map_host_api(L, _lib, active, presc)
        eq, sizes = dof_map_noalloc(L, _lib, active, presc)
        assert sizes == [nd, nu]
        assert np.array_equal(eq, eq_ref)
        if with_active:
            assert (eq_ref == -1).any()
        assert nu > 0 and nd > nu


def test_dof_map_noalloc_of_an_analysis(amaru):
    """The eq ids of a configured model (configure_dofs!, src/bc.jl:198-233)."""
    from amaru_amd import _lib
    L = _lib.lib()
    mesh = amaru.Mesh(amaru.Block([[0, 0, 0], [1, 1, 1]], nx=6, ny=5, nz=4, cellshape="HEX20"))
    model = amaru.FEModel(mesh, [("all", amaru.MechSolid, amaru.LinearElastic(E=1e9, nu=0.3))])
    bcs = [("z==0", amaru.SurfaceBC(ux=0, uy=0, uz=0)), ("x==1", amaru.SurfaceBC(ux=0))]
    ndofs, nu = model.configure_dofs(bcs)
    presc = np.zeros((mesh.nnodes, 3), dtype=np.uint8)
    presc[model.eq >= nu] = 1
    eq, sizes = dof_map_noalloc(L, _lib, None, presc.ravel())
    assert sizes == [ndofs, nu]
    assert np.array_equal(eq.reshape(-1, 3), model.eq)
