"""Problem description shared by the device binding and its callers.

`Problem` is the Python image of the C-ABI struct `pr_problem`
(include/pulsar_rocm.h): one GRAPE problem = drift + controls + ensemble +
target, with ComplexF64 matrices stored column-major exactly as Julia stores
them. It corresponds to the reference's QuantumSystem / ControlSequence /
QuantumTarget triple (src/Types/QuantumSystem.jl:23-130,
src/Types/ControlSequence.jl:37-59, src/Types/Targets.jl:37-144) plus the
ensemble of perturbed systems built by build_ensemble_from_systems
(src/Optimization/Ensemble/SystemBuilder.jl:49-87).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

PR_CLOSED, PR_LINDBLAD = 0, 1
PR_GATE, PR_PAIRS_SQUARE, PR_PAIRS_REAL, PR_TRACKING, PR_GATE_PROPAGATORS, PR_STATE = 0, 1, 2, 3, 4, 5
PR_STATE_PROPAGATORS = 6
PR_PAIRS_MODULUS = 7
PR_FP64, PR_MIXED, PR_SPLIT16 = 0, 1, 2
PR_SEGMENT_AUTOTUNE = -1  # pr_options.segment: time candidate segment lengths at plan creation
PR_GRAD_REFERENCE, PR_GRAD_EXACT = 0, 1


def cmat(a) -> np.ndarray:
    """A complex matrix (or stack of matrices) in column-major ComplexF64 order.

    Returns a C-contiguous complex128 array whose memory is the column-major
    image of each matrix (i.e. the last two axes are stored transposed).
    """
    a = np.asarray(a, dtype=np.complex128)
    return np.ascontiguousarray(np.swapaxes(a, -1, -2))


def from_cmat(buf: np.ndarray, d: int) -> np.ndarray:
    """Inverse of `cmat` for one d x d matrix held as interleaved doubles."""
    z = np.asarray(buf).view(np.complex128) if np.asarray(buf).dtype != np.complex128 else np.asarray(buf)
    return np.ascontiguousarray(z.reshape(d, d).T)


@dataclass
class Problem:
    """One GRAPE problem in the layout of `pr_problem`.

    H0: (d, d) or (M, d, d) complex; Hc: (n_ctrl, d, d) complex (row-major
    numpy matrices; converted to column-major when passed to C).
    """

    dim: int
    n_ctrl: int
    n_t: int
    H0: np.ndarray
    Hc: np.ndarray
    dt: float = 0.0
    dt_slices: Optional[np.ndarray] = None
    dynamics: int = PR_CLOSED
    objective: int = PR_GATE
    n_members: int = 1
    ctrl_scale: Optional[np.ndarray] = None
    weight: Optional[np.ndarray] = None
    U_target: Optional[np.ndarray] = None
    psi_init: Optional[np.ndarray] = None
    psi_targ: Optional[np.ndarray] = None
    # PR_TRACKING checkpoints: 1-based steps, weights, (n_track, d, d) density matrices
    track_step: Optional[np.ndarray] = None
    track_weight: Optional[np.ndarray] = None
    track_target: Optional[np.ndarray] = None
    _keep: list = field(default_factory=list, repr=False)

    @property
    def H0_per_member(self) -> int:
        return 1 if np.asarray(self.H0).ndim == 3 else 0

    @property
    def n_pairs(self) -> int:
        return 0 if self.psi_init is None else int(np.asarray(self.psi_init).shape[0])

    @property
    def n_track(self) -> int:
        return 0 if self.track_step is None else int(np.asarray(self.track_step).shape[0])

    def validate_shapes(self) -> None:
        d = self.dim
        H0 = np.asarray(self.H0)
        if H0.shape not in ((d, d), (self.n_members, d, d)):
            raise ValueError(f"H0 shape {H0.shape} does not match dim={d}, n_members={self.n_members}")
        if np.asarray(self.Hc).shape != (self.n_ctrl, d, d):
            raise ValueError(f"Hc shape {np.asarray(self.Hc).shape} != ({self.n_ctrl}, {d}, {d})")
        if self.ctrl_scale is not None and np.asarray(self.ctrl_scale).shape != (self.n_members, self.n_ctrl):
            raise ValueError("ctrl_scale must be [n_members, n_ctrl]")
        if self.weight is not None and np.asarray(self.weight).shape != (self.n_members,):
            raise ValueError("weight must be [n_members]")
        if self.dt_slices is not None and np.asarray(self.dt_slices).shape != (self.n_t,):
            raise ValueError("dt_slices must be [n_t]")
        if self.objective in (PR_GATE, PR_GATE_PROPAGATORS):
            if self.U_target is None or np.asarray(self.U_target).shape != (d, d):
                raise ValueError("gate objective needs U_target (dim x dim)")
        elif self.objective == PR_TRACKING:
            if self.psi_init is None or np.asarray(self.psi_init).ndim != 2 or np.asarray(self.psi_init).shape[1] != d:
                raise ValueError("the tracking objective needs psi_init [n_pairs, dim]")
            n = self.n_track
            if (self.track_weight is None or self.track_target is None or np.asarray(self.track_weight).shape != (n,)
                    or np.asarray(self.track_target).shape != (n, d, d)):
                raise ValueError("tracking needs track_step [n], track_weight [n] and track_target [n, dim, dim]")
        else:
            if self.psi_init is None or self.psi_targ is None:
                raise ValueError("pair objectives need psi_init and psi_targ")
            if np.asarray(self.psi_init).shape != np.asarray(self.psi_targ).shape or np.asarray(self.psi_init).shape[1] != d:
                raise ValueError("psi_init / psi_targ must both be [n_pairs, dim]")


class CProblem(ctypes.Structure):
    """ctypes image of `pr_problem` (include/pulsar_rocm.h)."""

    _fields_ = [
        ("dim", ctypes.c_int),
        ("n_ctrl", ctypes.c_int),
        ("n_t", ctypes.c_int),
        ("dt", ctypes.c_double),
        ("dt_slices", ctypes.POINTER(ctypes.c_double)),
        ("dynamics", ctypes.c_int),
        ("objective", ctypes.c_int),
        ("n_members", ctypes.c_int),
        ("H0", ctypes.POINTER(ctypes.c_double)),
        ("H0_per_member", ctypes.c_int),
        ("Hc", ctypes.POINTER(ctypes.c_double)),
        ("ctrl_scale", ctypes.POINTER(ctypes.c_double)),
        ("weight", ctypes.POINTER(ctypes.c_double)),
        ("U_target", ctypes.POINTER(ctypes.c_double)),
        ("n_pairs", ctypes.c_int),
        ("psi_init", ctypes.POINTER(ctypes.c_double)),
        ("psi_targ", ctypes.POINTER(ctypes.c_double)),
        ("n_track", ctypes.c_int),
        ("track_step", ctypes.POINTER(ctypes.c_int)),
        ("track_weight", ctypes.POINTER(ctypes.c_double)),
        ("track_target", ctypes.POINTER(ctypes.c_double)),
    ]


class COptions(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int),
        ("grad_mode", ctypes.c_int),
        ("segment", ctypes.c_int),
        ("store_props", ctypes.c_int),
        ("pack", ctypes.c_int),
    ]


def _dptr(arr: Optional[np.ndarray], keep: list):
    if arr is None:
        return ctypes.POINTER(ctypes.c_double)()
    a = np.ascontiguousarray(arr)
    if a.dtype == np.complex128:
        a = a.view(np.float64)
    a = np.ascontiguousarray(a, dtype=np.float64)
    keep.append(a)
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def to_c(p: Problem) -> CProblem:
    """Build the ctypes struct; the arrays it points to live in p._keep."""
    p.validate_shapes()
    keep: list = []
    c = CProblem()
    c.dim, c.n_ctrl, c.n_t = p.dim, p.n_ctrl, p.n_t
    c.dt = float(p.dt)
    c.dt_slices = _dptr(None if p.dt_slices is None else np.asarray(p.dt_slices, np.float64), keep)
    c.dynamics, c.objective, c.n_members = p.dynamics, p.objective, p.n_members
    c.H0 = _dptr(cmat(p.H0), keep)
    c.H0_per_member = p.H0_per_member
    c.Hc = _dptr(cmat(p.Hc), keep)
    c.ctrl_scale = _dptr(None if p.ctrl_scale is None else np.asarray(p.ctrl_scale, np.float64), keep)
    c.weight = _dptr(None if p.weight is None else np.asarray(p.weight, np.float64), keep)
    c.U_target = _dptr(None if p.U_target is None else cmat(p.U_target), keep)
    c.n_pairs = p.n_pairs
    c.psi_init = _dptr(None if p.psi_init is None else np.asarray(p.psi_init, np.complex128), keep)
    c.psi_targ = _dptr(None if p.psi_targ is None else np.asarray(p.psi_targ, np.complex128), keep)
    c.n_track = p.n_track
    if p.track_step is not None:
        steps = np.ascontiguousarray(p.track_step, dtype=np.int32)
        keep.append(steps)
        c.track_step = steps.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    c.track_weight = _dptr(None if p.track_weight is None else np.asarray(p.track_weight, np.float64), keep)
    c.track_target = _dptr(None if p.track_target is None else cmat(p.track_target), keep)
    p._keep = keep
    return c


def controls_cm(u: np.ndarray) -> np.ndarray:
    """[n_ctrl, n_t] numpy controls -> the column-major Julia buffer (u_jk at j + k*n_ctrl)."""
    u = np.asarray(u, dtype=np.float64)
    return np.ascontiguousarray(u.T).reshape(-1)


def grad_from_cm(buf: np.ndarray, n_ctrl: int, n_t: int) -> np.ndarray:
    """Column-major [n_ctrl x n_t] buffer -> numpy [n_ctrl, n_t]."""
    return np.asarray(buf).reshape(n_t, n_ctrl).T.copy()
