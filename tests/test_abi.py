"""The C ABI library loads and exports every function include/amaru_hip.h
declares (CPU: no compute calls)."""
import os
import re

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
