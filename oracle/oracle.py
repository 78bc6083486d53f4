"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the device path. Imported by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg; never by the
product package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.join(os.path.dirname(_HERE), "pulsar.jl_amd")
if _PKG not in sys.path:
    sys.path.insert(0, _PKG)

from pulsar_rocm.problem import (PR_GRAD_EXACT, PR_GRAD_REFERENCE, Problem, controls_cm,  # noqa: E402
                                 grad_from_cm, to_c)

_LIB = None


def build() -> str:
    """Compile liboracle.so with the committed Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        L.or_eigh.argtypes = [ctypes.c_int, dp, dp, dp]
        L.or_expm_neg_i.argtypes = [ctypes.c_int, dp, ctypes.c_double, dp]
        L.or_expm.argtypes = [ctypes.c_int, dp, dp]
        L.or_expm_frechet.argtypes = [ctypes.c_int, dp, dp, dp]
        L.or_eval.argtypes = [ctypes.c_void_p, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              dp, dp, dp, dp]
        L.or_total_propagator.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _cm(m):
    """numpy complex matrix -> column-major interleaved doubles."""
    return np.ascontiguousarray(np.asarray(m, np.complex128).T).view(np.float64).copy()


def _from_cm(buf, d):
    return buf.view(np.complex128).reshape(d, d).T.copy()


def expm_neg_i(H, dt):
    d = H.shape[0]
    h = _cm(H)
    out = np.zeros(2 * d * d)
    rc = lib().or_expm_neg_i(d, _p(h), float(dt), _p(out))
    assert rc == 0, rc
    return _from_cm(out, d)


def expm(A):
    n = A.shape[0]
    a = _cm(A)
    out = np.zeros(2 * n * n)
    rc = lib().or_expm(n, _p(a), _p(out))
    assert rc == 0, rc
    return _from_cm(out, n)


def expm_frechet(X, E):
    n = X.shape[0]
    x, e = _cm(X), _cm(E)
    out = np.zeros(2 * n * n)
    rc = lib().or_expm_frechet(n, _p(x), _p(e), _p(out))
    assert rc == 0, rc
    return _from_cm(out, n)


def eigh(H):
    d = H.shape[0]
    h = _cm(H)
    w = np.zeros(d)
    V = np.zeros(2 * d * d)
    rc = lib().or_eigh(d, _p(h), _p(w), _p(V))
    assert rc == 0, rc
    return w, _from_cm(V, d)


def evaluate(p: Problem, u: np.ndarray, grad: bool = True, grad_mode: int = PR_GRAD_REFERENCE,
             separate_fg: bool = False, n_threads: int = 0, per_member: bool = False):
    """Ensemble F and G[n_ctrl, n_t] of one problem (FP64 oracle)."""
    c = to_c(p)
    ub = controls_cm(u)
    F = np.zeros(1)
    G = np.zeros(p.n_ctrl * p.n_t) if grad else None
    Fm = np.zeros(p.n_members) if per_member else None
    Gm = np.zeros(p.n_members * p.n_ctrl * p.n_t) if (per_member and grad) else None
    null = ctypes.POINTER(ctypes.c_double)()
    rc = lib().or_eval(ctypes.byref(c), _p(ub), int(grad_mode), int(separate_fg), int(n_threads),
                       _p(F), _p(G) if grad else null, _p(Fm) if Fm is not None else null,
                       _p(Gm) if Gm is not None else null)
    if rc != 0:
        raise RuntimeError(f"or_eval failed with status {rc}")
    out_G = grad_from_cm(G, p.n_ctrl, p.n_t) if grad else None
    if per_member:
        gm = None
        if Gm is not None:
            gm = np.stack([grad_from_cm(Gm[m * p.n_ctrl * p.n_t:(m + 1) * p.n_ctrl * p.n_t],
                                        p.n_ctrl, p.n_t) for m in range(p.n_members)])
        return float(F[0]), out_G, Fm, gm
    return float(F[0]), out_G


def total_propagator(p: Problem, u: np.ndarray, member: int = 0):
    c = to_c(p)
    ub = controls_cm(u)
    out = np.zeros(2 * p.dim * p.dim)
    rc = lib().or_total_propagator(ctypes.byref(c), int(member), _p(ub), _p(out))
    if rc != 0:
        raise RuntimeError(f"or_total_propagator failed with status {rc}")
    return _from_cm(out, p.dim)


__all__ = ["build", "lib", "expm_neg_i", "expm", "expm_frechet", "eigh", "evaluate",
           "total_propagator", "PR_GRAD_EXACT", "PR_GRAD_REFERENCE"]
