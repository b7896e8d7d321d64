"""The C ABI library loads and exports every function include/amaru_hip.h
declares (CPU: no compute calls)."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(amr_\w+)\s*\(", src, flags=re.M)
    return names


def test_library_exports_every_declared_symbol(amaru):
    from amaru_amd import _lib
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), f"{n} not exported"
    assert set(names) == set(_lib.SIGNATURES), "ctypes signature table out of sync with the header"


def test_error_path_without_gpu_work(amaru):
    import ctypes
    from amaru_amd import _lib
    L = _lib.lib()
    # argument validation happens before any device call
    rc = L.amr_model_create(ctypes.byref(ctypes.c_void_p()), 0, 99, 0, 1, None, 0, None, 0, None, None, None, None,
                            0, 0)
    assert rc == -1
    assert "null pointer" in L.amr_last_error().decode() or "element type" in L.amr_last_error().decode()


def test_load_programs_are_checked_without_gpu(amaru):
    """amr_check_expr validates a postfix load program on the host; the
    compiler emits programs it accepts."""
    from amaru_amd import _lib
    from amaru_amd.expr import compile_program
    L = _lib.lib()
    for e in (2.0, "3*z", "x^2 + sin(y) - 0.5*t", "-(x+y)*z/2 - div(x, 0.7)", "sqrt(abs(x))*exp(-t)+log(2+y)",
              "max(tan(x), log10(1+y)) - min(z, t)"):
        ops, consts = compile_program(e)
        assert L.amr_check_expr(_lib.ptr(ops), len(ops), len(consts)) == 0
    bad = [np.array([[14, 0]], dtype=np.int32),                 # binary operation on an empty stack
           np.array([[1, 0], [2, 0]], dtype=np.int32),          # leaves two values
           np.array([[0, 3]], dtype=np.int32),                  # constant index out of range
           np.array([[99, 0]], dtype=np.int32),                 # unknown operation
           np.array([[1, 0]] * 17 + [[14, 0]] * 16, dtype=np.int32)]   # deeper than 16 values
    for ops in bad:
        assert L.amr_check_expr(_lib.ptr(np.ascontiguousarray(ops)), len(ops), 1) == -1
    with pytest.raises(ValueError):
        compile_program("x==0 && y>1")


def test_dof_map_noalloc_refuses_oversize_without_gpu(amaru):
    """The allocation-free DOF map checks its work-space limit before any
    device call."""
    from amaru_amd import _lib
    L = _lib.lib()
    n = (1 << 24) // 3 + 1   # 3n dofs: one node over the limit
    rc = L.amr_dof_map_noalloc(0, n, None, 1, 1, 1, None)
    assert rc == -1 and "scratch limit" in L.amr_last_error().decode()
    assert L.amr_dof_map_noalloc(0, 10, None, None, 1, 1, None) == -1
