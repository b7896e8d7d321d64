"""CPU oracle for the Amaru.jl FE assembly and solve path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg may import this package, and only as the
checker.  The product (``amaru.jl_amd``) never imports it.

Contents
--------
``amaru_oracle.c``  C restatement of the element/material arithmetic
                    (stiffness, mass, IP updates) and the mount_K triplet rule (built by ``Makefile`` into
                    ``libamaru_oracle.so``).
``fem.py``          Python restatement of the structured mesher, DOF
                    numbering, boundary conditions, solve_system! (SuperLU
                    direct solve of K11) and mech_stage_solver!.
``recovery.py``     ip_state_vals of the solid materials and
                    nodal_patch_recovery (numpy, SVD pseudo-inverse).
``dynamic.py``      dyn_stage_solver! (Newmark, Rayleigh damping) over the
                    restated mount_K / mount_M / update_state and SuperLU.

Every function cites the reference file:line it restates.  Pins: see
``tests/test_oracle_*.py`` (reference golden values, closed forms and
invariants).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    """Compile libamaru_oracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    """ctypes handle of the oracle library (built on first use)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "libamaru_oracle.so")
    src = os.path.join(_HERE, "amaru_oracle.c")
    if not os.path.exists(path) or os.path.getmtime(path) < os.path.getmtime(src):
        build()
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    i64, i32 = ctypes.c_int64, ctypes.c_int
    L.ora_nnodes.argtypes = [i32]
    L.ora_ip_rule.argtypes = [i32, i32, P, P]
    L.ora_shape_deriv.argtypes = [i32, P, P]
    L.ora_shape_func.argtypes = [i32, P, P]
    L.ora_rule_derivs.argtypes = [i32, i32, P]
    L.ora_calcDe.argtypes = [ctypes.c_double, ctypes.c_double, P]
    L.ora_calcD.argtypes = [i32, P, P, P]
    L.ora_update_state.argtypes = [i32, P, P, P, P]
    L.ora_elem_stiffness.argtypes = [i32, i32, i64, P, P, P, P, P, i32, P, P, i32]
    L.ora_elem_stiffness.restype = i64
    L.ora_update_elems.argtypes = [i32, i32, i64, P, P, P, P, P, P, i32, P, P, P, P, i32]
    L.ora_gather_forces.argtypes = [i64, i32, P, P, i64, P]
    L.ora_count_triplets.argtypes = [i64, i32, P]
    L.ora_elem_mass.argtypes = [i32, i32, i64, P, P, P, P]
    L.ora_elem_mass.restype = i64
    L.ora_count_triplets.restype = i64
    L.ora_mount_K.argtypes = [i64, i32, P, P, i64, P, P, P]
    L.ora_mount_K.restype = i64
    _LIB = L
    return L


def ptr(a):
    """Raw data pointer of a C-contiguous numpy array."""
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


HEX8, HEX20, TET10 = 1, 2, 3
NSTATE = 14


def rule(etype, nips=0):
    """IP natural coordinates (n×3) and weights of a rule."""
    R = np.zeros((27, 3))
    w = np.zeros(27)
    n = lib().ora_ip_rule(etype, nips, ptr(R), ptr(w))
    if n < 0:
        raise ValueError(f"no {nips}-point rule for element type {etype}")
    return R[:n].copy(), w[:n].copy()


def shape_deriv(etype, R):
    nn = lib().ora_nnodes(etype)
    D = np.zeros((nn, 3))
    lib().ora_shape_deriv(etype, ptr(np.ascontiguousarray(R, dtype=float)), ptr(D))
    return D


def shape_func(etype, R):
    nn = lib().ora_nnodes(etype)
    N = np.zeros(nn)
    lib().ora_shape_func(etype, ptr(np.ascontiguousarray(R, dtype=float)), ptr(N))
    return N
