"""ctypes wrapper of ``oracle/liboracle_cpu.so`` (C/OpenMP oracle) - TEST INFRASTRUCTURE ONLY.

Only the tests, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` may use this.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle_cpu.so"

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_f64 = ctypes.c_double
_lib = None


def build():
    """Compile the C oracle (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        h = ctypes.CDLL(str(LIB))
        h.oracle_assemble_K.argtypes = [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int]
        h.oracle_spmv.argtypes = [_i64, _i64, _vp, _vp, _vp, _vp, _vp, ctypes.c_int]
        h.oracle_step.argtypes = [_i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _f64, _vp, _vp,
                                  _f64, _f64, _f64, _vp, _i64, _f64, _vp, ctypes.c_int]
        h.oracle_run.argtypes = [_i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp,
                                 _vp, _f64, _f64, _f64, _vp, _i64, _f64, _vp, _i64, _vp,
                                 ctypes.c_int]
        h.oracle_max_threads.restype = ctypes.c_int
        _lib = h
    return _lib


def _p(a):
    return a.ctypes.data_as(_vp)


def assemble_K(nodes, elems, D, coefK, indptr, indices, threads=0):
    """Reference-layout stiffness ``K [nnzb, S, 3, 3]``."""
    nodes = np.ascontiguousarray(nodes, np.float64)
    elems = np.ascontiguousarray(elems, np.int64)
    D = np.ascontiguousarray(D, np.float64)
    coefK = np.ascontiguousarray(coefK, np.float64)
    indptr = np.ascontiguousarray(indptr, np.int64)
    indices = np.ascontiguousarray(indices, np.int64)
    S = coefK.shape[1]
    K = np.zeros((indices.size, S, 3, 3))
    lib().oracle_assemble_K(elems.shape[0], S, _p(nodes), _p(elems), _p(D), _p(coefK), _p(indptr),
                            _p(indices), _p(K), threads)
    return K


def spmv(indptr, indices, K, v, threads=0):
    S = K.shape[1]
    n = indptr.size - 1
    res = np.empty((3 * n, S))
    v = np.ascontiguousarray(v, np.float64)
    lib().oracle_spmv(n, S, _p(indptr), _p(indices), _p(K), _p(v), _p(res), threads)
    return res


def step(indptr, indices, K, LM, LHS, f1, p, u, up, dt, beta=0.0, damp=0.0, bdy_dofs=None,
         bdy_u=0.0, threads=0, out=None):
    S = K.shape[1]
    n = indptr.size - 1
    unew = np.empty((3 * n, S)) if out is None else out
    bd = np.ascontiguousarray(np.zeros(0, np.int64) if bdy_dofs is None else bdy_dofs, np.int64)
    rc = lib().oracle_step(n, S, indices.size, _p(indptr), _p(indices), _p(K),
                           _p(np.ascontiguousarray(LM)), _p(np.ascontiguousarray(LHS)),
                           _p(np.ascontiguousarray(f1, np.float64)), float(p),
                           _p(np.ascontiguousarray(u)), _p(np.ascontiguousarray(up)), float(dt),
                           float(beta), float(damp), _p(bd), bd.size, float(bdy_u), _p(unew), threads)
    if rc != 0:
        raise MemoryError("oracle_step: allocation failed")
    return unew


def run(indptr, indices, K, LM, LHS, f1, p, u, up, dt, beta=0.0, damp=0.0, bdy_dofs=None,
        bdy_u=0.0, snap_steps=(), threads=0):
    """``len(p)`` steps of :func:`step` (bit-identical results) in one native call.

    Returns ``(u_n, u_{n-1}, snaps)`` with ``snaps[i]`` the state after ``snap_steps[i]`` steps.
    """
    S = K.shape[1]
    n = indptr.size - 1
    p = np.ascontiguousarray(np.atleast_1d(p), np.float64)
    u = np.array(u, np.float64, order="C", copy=True)
    up = np.array(up, np.float64, order="C", copy=True)
    bd = np.ascontiguousarray(np.zeros(0, np.int64) if bdy_dofs is None else bdy_dofs, np.int64)
    snap = np.ascontiguousarray(snap_steps, np.int64)
    if snap.size and (snap.min() < 1 or snap.max() > p.size):
        raise ValueError("snapshot steps must lie in 1..len(p)")
    snaps = np.empty((snap.size, 3 * n, S))
    rc = lib().oracle_run(n, S, indices.size, _p(indptr), _p(indices), _p(np.ascontiguousarray(K)),
                          _p(np.ascontiguousarray(LM)), _p(np.ascontiguousarray(LHS)),
                          _p(np.ascontiguousarray(f1, np.float64)), _p(p), p.size, _p(u), _p(up),
                          float(dt), float(beta), float(damp), _p(bd), bd.size, float(bdy_u),
                          _p(snap), snap.size, _p(snaps), threads)
    if rc != 0:
        raise MemoryError("oracle_run: allocation failed")
    return u, up, snaps


def max_threads():
    return lib().oracle_max_threads()
