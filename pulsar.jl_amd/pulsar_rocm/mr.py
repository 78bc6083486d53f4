"""Batched MR kernels on the :rocm device.

  MRControl / grape_state_kernel          src/Application/MR/OptControl.jl:137-259,
                                          src/Application/MR/GRAPEState.jl:73-250
  LindbladMRControl / grape_lindblad_kernel
                                          OptControl.jl:733-885, GRAPELindblad.jl:83-215
  build_drift_liouvillian, build_control_liouvillian, pure_state_to_vec_rho
                                          src/Physics/Lindblad.jl:89-197
  TrackingPoint / grape_tracking_kernel   src/Application/MR/GRAPETracking.jl:89-157
  fidelity_forward                        src/Application/MR/GRAPEState.jl:406-566
  optimcon (:lbfgs, :grape; tracking)     src/Application/MR/OptControl.jl:310-611, 1040-1236

Ensemble members are the (drift, power level) pairs in the reference's order
(`for H in drifts for p in pwr_levels`, GRAPEState.jl:146); every member
propagates every (rho_init[s], rho_targ[s]) pair; F and G are the mean over
N_ens = n_drifts * n_pwr * n_pairs (GRAPEState.jl:239-244). Penalties other
than :none are not part of this path.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import native
from .problem import (PR_CLOSED, PR_GRAD_REFERENCE, PR_LINDBLAD, PR_PAIRS_MODULUS, PR_PAIRS_REAL, PR_PAIRS_SQUARE,
                      PR_TRACKING, Problem)


def build_drift_liouvillian(H_drift, jump_ops=(), decay_rates=()):
    """-i(I (x) H - H^T (x) I) + sum_k g_k (conj(L) (x) L - 1/2 I (x) L^dag L - 1/2 conj(L^dag L) (x) I)
    on column-stacked vec(rho) (Lindblad.jl:126-157)."""
    H = np.asarray(H_drift, dtype=np.complex128)
    N = H.shape[0]
    I = np.eye(N)
    L = -1j * np.kron(I, H) + 1j * np.kron(H.T, I)
    for Lk, g in zip(jump_ops, decay_rates):
        if g == 0.0 or not np.isfinite(g):
            continue
        if g < 0:
            raise ValueError(f"Lindblad decay rate must be non-negative; got {g}")
        Lk = np.asarray(Lk, dtype=np.complex128)
        A = Lk.conj().T @ Lk
        L = L + g * (np.kron(Lk.conj(), Lk) - 0.5 * np.kron(I, A) - 0.5 * np.kron(A.conj(), I))
    return L


def build_control_liouvillian(H_ctrl):
    """-i(I (x) H - H^T (x) I) (Lindblad.jl:189-197)."""
    H = np.asarray(H_ctrl, dtype=np.complex128)
    I = np.eye(H.shape[0])
    return -1j * np.kron(I, H) + 1j * np.kron(H.T, I)


def vec_rho(rho):
    return np.asarray(rho, dtype=np.complex128).reshape(-1, order="F")


def pure_state_to_vec_rho(psi):
    c = np.asarray(psi, dtype=np.complex128).reshape(-1)
    return vec_rho(np.outer(c, c.conj()))


def _fid_code(fidelity: str, lindblad: bool = False) -> int:
    """The device objective of an MR fidelity symbol: the GRAPE-compatible
    types of _validate_ctrl (OptControl.jl:646-650) for closed dynamics, the
    two of lindblad_grad_prefactor (Lindblad.jl:220-232) for Lindblad."""
    f = fidelity.lstrip(":")
    if f == "square":
        return PR_PAIRS_SQUARE
    if f == "real":
        return PR_PAIRS_REAL
    if f == "modulus" and not lindblad:
        return PR_PAIRS_MODULUS
    if lindblad:
        raise ValueError(f"lindblad_grad_prefactor: ':{f}' not supported. Use :square or :real.")
    raise ValueError(f"ctrl.fidelity ':{f}' is not GRAPE-compatible. Use one of: real, square, modulus")


@dataclass
class TrackingPoint:
    """TrackingPoint (GRAPETracking.jl:89-97): the density matrix after slice
    `step` (1-based, 1..n_t) is compared with `target_dm` at weight `weight`."""

    step: int
    target_dm: np.ndarray
    weight: float


@dataclass
class MRControl:
    """The fields of MRControl (OptControl.jl:137-259) that the state and
    tracking kernels read."""

    drifts: List[np.ndarray]
    operators: List[np.ndarray]
    rho_init: List[np.ndarray]
    rho_targ: List[np.ndarray]
    pulse_dt: np.ndarray
    pwr_levels: List[float] = field(default_factory=lambda: [1.0])
    fidelity: str = "real"  # the MRControl keyword default (OptControl.jl:217)
    backend: str = "rocm"
    precision: str = "f64"
    tracking: List[TrackingPoint] = field(default_factory=list)
    # optimcon settings (OptControl.jl:210-220 defaults)
    l_bound: float = -1.0
    u_bound: float = 1.0
    method: str = "lbfgs"
    max_iter: int = 500
    grad_tol: float = 1e-7
    lbfgs_memory: int = 10

    def tracking_problem(self) -> Problem:
        """The tracking objective over the same ensemble: initial states
        rho_init (pure), checkpoints from `tracking` (GRAPETracking.jl:162-186)."""
        if not self.tracking:
            raise ValueError("grape_tracking_kernel: ctrl.tracking is empty. "
                             "Add TrackingPoint entries or call grape_state_kernel instead.")
        base = self.problem()
        d = base.dim
        base.objective = PR_TRACKING
        base.psi_targ = None
        base.track_step = np.array([int(t.step) for t in self.tracking], dtype=np.int32)
        base.track_weight = np.array([float(t.weight) for t in self.tracking], dtype=np.float64)
        base.track_target = np.stack([np.asarray(t.target_dm, np.complex128).reshape(d, d) for t in self.tracking])
        return base

    def problem(self) -> Problem:
        drifts = [np.asarray(h, np.complex128) for h in self.drifts]
        ops = np.stack([np.asarray(o, np.complex128) for o in self.operators])
        M = len(drifts) * len(self.pwr_levels)
        H0 = np.stack([h for h in drifts for _ in self.pwr_levels])
        scale = np.array([[p] * len(ops) for _ in drifts for p in self.pwr_levels], dtype=np.float64)
        psi0 = np.stack([np.asarray(r, np.complex128).reshape(-1) for r in self.rho_init])
        psit = np.stack([np.asarray(r, np.complex128).reshape(-1) for r in self.rho_targ])
        dts = np.asarray(self.pulse_dt, dtype=np.float64)
        return Problem(dim=ops.shape[1], n_ctrl=len(ops), n_t=len(dts), H0=H0, Hc=ops, dt_slices=dts,
                       dynamics=PR_CLOSED, objective=_fid_code(self.fidelity), n_members=M, ctrl_scale=scale,
                       psi_init=psi0, psi_targ=psit)


@dataclass
class LindbladMRControl:
    """The fields of LindbladMRControl (OptControl.jl:733-885) read by the Lindblad kernel;
    the Liouvillians are built as the reference precomputes _L_drifts/_L_controls."""

    drifts: List[np.ndarray]
    operators: List[np.ndarray]
    rho_init: List[np.ndarray]
    rho_targ: List[np.ndarray]
    pulse_dt: np.ndarray
    jump_ops: List[np.ndarray] = field(default_factory=list)
    decay_rates: List[float] = field(default_factory=list)
    pwr_levels: List[float] = field(default_factory=lambda: [1.0])
    fidelity: str = "square"  # the LindbladMRControl keyword default (OptControl.jl:818)
    backend: str = "rocm"
    l_bound: float = -1.0
    u_bound: float = 1.0
    method: str = "lbfgs"
    max_iter: int = 500
    grad_tol: float = 1e-7
    lbfgs_memory: int = 10
    tracking: List[TrackingPoint] = field(default_factory=list)  # always empty (Hilbert-space feature)

    def problem(self) -> Problem:
        Ld = [build_drift_liouvillian(h, self.jump_ops, self.decay_rates) for h in self.drifts]
        Lc = np.stack([build_control_liouvillian(o) for o in self.operators])

        def as_vec(r):
            r = np.asarray(r, np.complex128)
            return vec_rho(r) if r.ndim == 2 else pure_state_to_vec_rho(r) if r.size ** 2 == Lc.shape[1] else r

        s0 = np.stack([as_vec(r) for r in self.rho_init])
        st = np.stack([as_vec(r) for r in self.rho_targ])
        M = len(Ld) * len(self.pwr_levels)
        L0 = np.stack([L for L in Ld for _ in self.pwr_levels])
        scale = np.array([[p] * len(Lc) for _ in Ld for p in self.pwr_levels], dtype=np.float64)
        dts = np.asarray(self.pulse_dt, dtype=np.float64)
        return Problem(dim=Lc.shape[1], n_ctrl=len(Lc), n_t=len(dts), H0=L0, Hc=Lc, dt_slices=dts,
                       dynamics=PR_LINDBLAD, objective=_fid_code(self.fidelity, lindblad=True), n_members=M, ctrl_scale=scale,
                       psi_init=s0, psi_targ=st)


def _blockdiag_at(i, f, A):
    d = A.shape[0]
    out = np.zeros((f * d, f * d), np.complex128)
    out[i * d:(i + 1) * d, i * d:(i + 1) * d] = A
    return out


class PackedMembers:
    """Small-dimension MR ensembles packed into 16 x 16 blocks.

    A state-pair problem with d <= 8 (closed, or Lindblad with a 2-level
    system, d = 4) spends a whole 16 x 16 block per member. With f the largest divisor of M with f d <= 16 (and f n_pairs <= 16),
    members m f .. m f + f - 1 become the diagonal
    blocks of one packed member: block-diagonal drift, and per control j and
    block i a channel blockdiag(0, .., A_j, .., 0) with the member's scale
    s_{mf+i, j} and the waveform row u_j; the pairs of every block are pairs
    of the packed member. Exp, chains and traces of a block-diagonal matrix
    act block by block, so F and the per-block gradient rows are those of the
    members (up to rounding: a pack shares one scaling exponent); the gradient
    of control j is the sum of its f channel rows. Tracking checkpoints
    become blockdiag(C, .., C). (The pairs belong to the problem, not to a
    member, so every pack is full: f divides M.)"""

    @staticmethod
    def factor(p: Problem) -> int:
        top = min(16 // p.dim, 16 // max(p.n_pairs, 1), p.n_members)
        return max([f for f in range(1, top + 1) if p.n_members % f == 0])

    def __init__(self, p: Problem):
        d, M, nc, npairs = p.dim, p.n_members, p.n_ctrl, p.n_pairs
        f = self.factor(p)
        self.f, self.nc, self.original = f, nc, p
        Mp = M // f
        H0 = np.asarray(p.H0, np.complex128)
        H0s = H0 if H0.ndim == 3 else np.broadcast_to(H0, (M, d, d))
        sc = np.ones((M, nc)) if p.ctrl_scale is None else np.asarray(p.ctrl_scale, np.float64)
        H0p = np.zeros((Mp, f * d, f * d), np.complex128)
        scp = np.zeros((Mp, f * nc))
        for m in range(M):
            q, i = divmod(m, f)
            H0p[q, i * d:(i + 1) * d, i * d:(i + 1) * d] = H0s[m]
            scp[q, i * nc:(i + 1) * nc] = sc[m]
        Hc = np.asarray(p.Hc, np.complex128)
        Hcp = np.stack([_blockdiag_at(i, f, Hc[j]) for i in range(f) for j in range(nc)])
        def pack_states(v):
            if v is None:
                return None
            v = np.asarray(v, np.complex128)
            out = np.zeros((f * npairs, f * d), np.complex128)
            for i in range(f):
                out[i * npairs:(i + 1) * npairs, i * d:(i + 1) * d] = v
            return out

        track = {}
        if p.objective == PR_TRACKING:
            # every block is compared with the same checkpoint density matrices
            tt = np.asarray(p.track_target, np.complex128)
            track = dict(track_step=np.asarray(p.track_step), track_weight=np.asarray(p.track_weight),
                         track_target=np.stack([np.kron(np.eye(f), t) for t in tt]))
        self.problem = Problem(dim=f * d, n_ctrl=f * nc, n_t=p.n_t, H0=H0p, Hc=Hcp, dt=p.dt,
                               dt_slices=p.dt_slices, dynamics=p.dynamics, objective=p.objective, n_members=Mp,
                               ctrl_scale=scp, weight=np.full(Mp, 1.0 / (M * npairs)),
                               psi_init=pack_states(p.psi_init), psi_targ=pack_states(p.psi_targ), **track)

    @staticmethod
    def applies(p: Problem) -> bool:
        return (p.objective in (PR_PAIRS_SQUARE, PR_PAIRS_REAL, PR_TRACKING) and p.dim <= 8
                and p.weight is None and p.n_members >= 2 and p.n_pairs <= 8 and PackedMembers.factor(p) >= 2)

    def controls(self, u):
        return np.tile(np.asarray(u, np.float64), (self.f, 1))

    def gradient(self, G):
        return np.asarray(G).reshape(self.f, self.nc, -1).sum(axis=0)


class KernelPlan:
    """A reusable device plan for repeated kernel calls with one control context
    (tracking=True: the tracking objective of an MRControl). Small closed
    state-pair ensembles run packed (PackedMembers) unless pack=False."""

    def __init__(self, ctrl, grad_mode: int = PR_GRAD_REFERENCE, tracking: bool = False, pack: bool = True):
        self.problem = ctrl.tracking_problem() if tracking else ctrl.problem()
        self.packed = PackedMembers(self.problem) if pack and PackedMembers.applies(self.problem) else None
        run = self.packed.problem if self.packed else self.problem
        # (pack=False: no packing in the library either)
        self.plan = native.Plan([run], grad_mode=grad_mode, pack=1 if pack else 0)

    def eval(self, waveform: np.ndarray, grad: bool = True):
        w = np.asarray(waveform, dtype=np.float64)
        if self.packed:
            F, G = self.plan.eval([self.packed.controls(w)], grad=grad)
            return float(F[0]), (self.packed.gradient(G[0]) if grad else None)
        F, G = self.plan.eval([w], grad=grad)
        return float(F[0]), (G[0] if grad else None)

    def __call__(self, waveform: np.ndarray):
        return self.eval(waveform, grad=True)

    def close(self):
        self.plan.close()


def grape_state_kernel(waveform: np.ndarray, ctrl: MRControl):
    """(fidelity, grad[n_ctrl x n_t]) of GRAPEState.jl:73-85 on the device."""
    kp = KernelPlan(ctrl)
    try:
        return kp(waveform)
    finally:
        kp.close()


def grape_lindblad_kernel(waveform: np.ndarray, ctrl: LindbladMRControl):
    """(fidelity, grad[n_ctrl x n_t]) of GRAPELindblad.jl:83-92 on the device."""
    kp = KernelPlan(ctrl)
    try:
        return kp(waveform)
    finally:
        kp.close()


def grape_tracking_kernel(waveform: np.ndarray, ctrl: MRControl):
    """(fidelity, grad[n_ctrl x n_t]) of GRAPETracking.jl:145-157 on the device
    (ctrl.tracking must be non-empty, as there)."""
    kp = KernelPlan(ctrl, tracking=True)
    try:
        return kp(waveform)
    finally:
        kp.close()


class FidelityForward:
    """fidelity_forward (GRAPEState.jl:406-566) on the device with a reusable
    plan: the ensemble-averaged state-transfer fidelity of one waveform
    (F only, no gradient walk), or of a population of waveforms in one launch
    (`batch`), as derivative-free optimizers evaluate them. Members are
    [(H, p) for H in drifts for p in pwr_levels]; F averages over members and
    pairs (N_ens = n_drifts n_pwr n_pairs, :512)."""

    def __init__(self, drifts, pwr_levels, operators, rho_init, rho_targ, pulse_dt,
                 fidelity_type: str = "real", population: int = 1):
        import copy
        ctrl = MRControl(drifts=list(drifts), operators=list(operators), rho_init=list(rho_init),
                         rho_targ=list(rho_targ), pulse_dt=np.asarray(pulse_dt, np.float64),
                         pwr_levels=[float(p) for p in pwr_levels], fidelity=str(fidelity_type))
        self.problem = ctrl.problem()
        self.population = int(population)
        if self.population < 1:
            raise ValueError("population must be >= 1")
        self.packed = PackedMembers(self.problem) if PackedMembers.applies(self.problem) else None
        run = self.packed.problem if self.packed else self.problem
        self.plan = native.Plan([copy.copy(run) for _ in range(self.population)])

    def batch(self, waveforms) -> np.ndarray:
        """F of each of `population` waveforms [n_ctrl x n_t]."""
        ws = [np.asarray(w, dtype=np.float64) for w in waveforms]
        if len(ws) != self.population:
            raise ValueError(f"expected {self.population} waveforms, got {len(ws)}")
        if self.packed:
            ws = [self.packed.controls(w) for w in ws]
        F, _ = self.plan.eval(ws, grad=False)
        return np.asarray(F, dtype=np.float64)

    def __call__(self, waveform) -> float:
        if self.population != 1:
            raise ValueError("a population plan evaluates with batch()")
        return float(self.batch([waveform])[0])

    def close(self):
        self.plan.close()


def fidelity_forward(waveform, drifts, pwr_levels, operators, rho_init, rho_targ, pulse_dt,
                     fidelity_type: str = "real", backend: str = "rocm") -> float:
    """fidelity_forward (GRAPEState.jl:434-466) with backend :rocm: one
    F-only device evaluation (for repeated calls keep a FidelityForward)."""
    if str(backend).lstrip(":") != "rocm":
        raise ValueError(f"fidelity_forward: backend :{str(backend).lstrip(':')} is not this library's (use :rocm)")
    ff = FidelityForward(drifts, pwr_levels, operators, rho_init, rho_targ, pulse_dt, fidelity_type)
    try:
        return ff(waveform)
    finally:
        ff.close()


# ------------------------------------------------------------------ optimcon
def _lbfgs_direction(g, s_list, y_list):
    """Two-loop recursion, ascent direction (OptControl.jl:565-590)."""
    k = len(s_list)
    if k == 0:
        return g.copy()
    q = g.copy()
    a = np.zeros(k)
    rho = [1.0 / np.dot(y_list[i], s_list[i]) for i in range(k)]
    for i in range(k - 1, -1, -1):
        a[i] = rho[i] * np.dot(s_list[i], q)
        q = q - a[i] * y_list[i]
    gamma = np.dot(s_list[-1], y_list[-1]) / np.dot(y_list[-1], y_list[-1])
    r = gamma * q
    for i in range(k):
        b = rho[i] * np.dot(y_list[i], r)
        r = r + (a[i] - b) * s_list[i]
    return r


def _line_search(w, d, F0, a0, f_only, lo, hi):
    """Armijo backtracking (OptControl.jl:594-611): c1 = 1e-4, up to 20 halvings;
    returns (alpha, w_new, F_new, evaluations counted as the reference counts them)."""
    a = a0
    for halving in range(1, 21):
        w_new = np.clip(w + a * d, lo, hi)
        F_new = f_only(w_new)
        if F_new >= F0 - 1e-4 * abs(a):
            return a, w_new, F_new, halving + 1
        a *= 0.5
    w_new = np.clip(w + a * d, lo, hi)
    F_new = f_only(w_new)
    return a, (w_new if F_new > F0 else w.copy()), max(F_new, F0), 21


def _vec(m):
    return np.asarray(m, np.float64).reshape(-1, order="F")  # Julia's vec (column-major)


def optimcon_lbfgs_loop(fg, f_only, guess, l_bound, u_bound, max_iter, grad_tol, lbfgs_memory, metadata=None):
    """_optimcon_lbfgs / _optimcon_lbfgs_tracking (OptControl.jl:382-484, 1040-1139)
    with fg(w) -> (F, G) and f_only(w) -> F."""
    from .api import OptimizationResult
    import time
    t0 = time.perf_counter()
    w = np.clip(np.array(guess, np.float64), l_bound, u_bound)
    F_curr, G_curr = fg(w)
    best_w, best_F = w.copy(), F_curr
    n_fid = n_grad = 1
    f_hist, g_hist = [F_curr], [float(np.linalg.norm(G_curr))]
    s_list, y_list = [], []
    converged, reason = False, "max iterations reached"
    for _ in range(max_iter):
        g_flat = _vec(G_curr)
        d_flat = _lbfgs_direction(g_flat, s_list, y_list)
        d_max = np.max(np.abs(d_flat))
        if d_max < 1e-14:
            converged, reason = True, "zero gradient direction"
            break
        d_flat = d_flat / d_max
        d_mat = d_flat.reshape(w.shape, order="F")
        a0 = 0.05 * (u_bound - l_bound)
        _, w_new, _, n_ls = _line_search(w, d_mat, F_curr, a0, f_only, l_bound, u_bound)
        n_fid += n_ls
        F_new2, G_new = fg(w_new)
        n_fid += 1
        n_grad += 1
        s_vec = _vec(w_new - w)
        y_vec = _vec(G_new) - g_flat
        if np.dot(y_vec, s_vec) > 1e-14 * np.linalg.norm(s_vec) ** 2:
            s_list.append(s_vec)
            y_list.append(y_vec)
            if len(s_list) > lbfgs_memory:
                s_list.pop(0)
                y_list.pop(0)
        w, G_curr = w_new, G_new
        F_prev, F_curr = F_curr, F_new2
        f_hist.append(F_curr)
        gnorm = float(np.linalg.norm(G_curr))
        g_hist.append(gnorm)
        if F_curr > best_F:
            best_F, best_w = F_curr, w.copy()
        if max_iter > 0 and F_curr >= 1.0 - 1e-8:
            converged, reason = True, "near-perfect fidelity"
            break
        if gnorm < grad_tol:
            converged, reason = True, f"gradient norm < {grad_tol}"
            break
        if len(f_hist) >= 2 and abs(F_curr - F_prev) < 1e-9:
            converged, reason = True, "fidelity stalled"
            break
    return OptimizationResult(best_w, best_F, f_hist, g_hist, len(f_hist) - 1, converged, reason,
                              time.perf_counter() - t0, n_fid, n_grad, dict(metadata or {}))


def optimcon_grape_loop(fg, guess, l_bound, u_bound, max_iter, grad_tol, metadata=None):
    """_optimcon_grape / _optimcon_grape_tracking (OptControl.jl:488-561, 1141-1216):
    steps of 0.05 (u_bound - l_bound) / max|G| along G, clipped."""
    from .api import OptimizationResult
    import time
    t0 = time.perf_counter()
    w = np.clip(np.array(guess, np.float64), l_bound, u_bound)
    F_curr, G_curr = fg(w)
    best_w, best_F = w.copy(), F_curr
    n_fid = n_grad = 1
    f_hist, g_hist = [F_curr], [float(np.linalg.norm(G_curr))]
    converged, reason = False, "max iterations reached"
    for _ in range(max_iter):
        g_max = np.max(np.abs(G_curr))
        if g_max > 1e-14:
            w = np.clip(w + (0.05 * (u_bound - l_bound) / g_max) * G_curr, l_bound, u_bound)
        F_curr, G_curr = fg(w)
        n_fid += 1
        n_grad += 1
        f_hist.append(F_curr)
        gnorm = float(np.linalg.norm(G_curr))
        g_hist.append(gnorm)
        if F_curr > best_F:
            best_F, best_w = F_curr, w.copy()
        if F_curr >= 1.0 - 1e-8:
            converged, reason = True, "near-perfect fidelity"
            break
        if gnorm < grad_tol:
            converged, reason = True, f"gradient norm < {grad_tol}"
            break
    return OptimizationResult(best_w, best_F, f_hist, g_hist, len(f_hist) - 1, converged, reason,
                              time.perf_counter() - t0, n_fid, n_grad, dict(metadata or {}))


def _validate(ctrl, guess):
    """_validate_mr_control (OptControl.jl:615-650), the shape checks."""
    for name in ("drifts", "operators", "rho_init", "pwr_levels", "pulse_dt"):
        if len(getattr(ctrl, name)) == 0:
            raise ValueError(f"ctrl.{name} must be non-empty")
    n_ctrl, n_t = np.asarray(guess).shape
    if n_ctrl != len(ctrl.operators):
        raise ValueError(f"guess has {n_ctrl} rows but ctrl.operators has {len(ctrl.operators)} entries")
    if n_t != len(ctrl.pulse_dt):
        raise ValueError(f"guess has {n_t} columns but ctrl.pulse_dt has {len(ctrl.pulse_dt)} entries")


_STOP_REASON = {0: "max iterations reached", 1: "near-perfect fidelity", 3: "fidelity stalled",
                4: "zero gradient direction"}


def optimcon(ctrl, guess, on_device: bool = True):
    """optimcon(ctrl, guess) for :lbfgs and :grape (OptControl.jl:310-367).
    Non-empty ctrl.tracking selects the tracking kernel, as there.

    on_device=True: the whole loop runs in the library (pr_mr_optimcon), with
    the waveform, gradient and curvature pairs in device memory; the line
    search evaluates F only. on_device=False: the same loop on the host
    (optimcon_lbfgs_loop / optimcon_grape_loop) over the same device plan."""
    _validate(ctrl, guess)
    method = str(ctrl.method).lstrip(":")
    if method not in ("lbfgs", "grape"):
        raise ValueError(f"method :{method} is not part of this library (use :lbfgs or :grape)")
    tracking = bool(getattr(ctrl, "tracking", []))
    kp = KernelPlan(ctrl, tracking=tracking)
    try:
        n_ens = kp.problem.n_members * kp.problem.n_pairs
        if on_device:
            from .api import OptimizationResult
            import time
            cfg = native.MrConfig(native.PR_MR_LBFGS if method == "lbfgs" else native.PR_MR_GRAPE,
                                  int(ctrl.max_iter), float(ctrl.grad_tol), int(ctrl.lbfgs_memory),
                                  float(ctrl.l_bound), float(ctrl.u_bound), kp.packed.f if kp.packed else 1)
            t0 = time.perf_counter()
            best, fh, gh, res = native.mr_optimcon_device(kp.plan, guess, cfg)
            reason = (f"gradient norm < {ctrl.grad_tol}" if res.reason == 2 else _STOP_REASON[res.reason])
            algo = ("MR Tracking " if tracking else
                    ("Lindblad " if isinstance(ctrl, LindbladMRControl) else "MR ")) + \
                ("L-BFGS-GRAPE" if method == "lbfgs" else "GRAPE")
            meta = {"algorithm": algo, "on_device": True}
            if tracking:
                meta["n_checkpoints"] = len(ctrl.tracking)
            else:
                meta["n_ensemble"] = n_ens
            if method == "lbfgs":
                meta["lbfgs_memory"] = ctrl.lbfgs_memory
            return OptimizationResult(best, res.fidelity, fh, gh, res.n_iterations, bool(res.converged), reason,
                                      time.perf_counter() - t0, res.n_fidelity_evaluations,
                                      res.n_gradient_evaluations, meta)
        def fg(w):
            return kp.eval(w, grad=True)

        def f_only(w):
            return kp.eval(w, grad=False)[0]

        if method == "lbfgs":
            meta = ({"algorithm": "MR Tracking L-BFGS-GRAPE", "n_checkpoints": len(ctrl.tracking)} if tracking else
                    {"algorithm": ("Lindblad" if isinstance(ctrl, LindbladMRControl) else "MR") + " L-BFGS-GRAPE",
                     "n_ensemble": n_ens})
            meta["lbfgs_memory"] = ctrl.lbfgs_memory
            return optimcon_lbfgs_loop(fg, f_only, guess, ctrl.l_bound, ctrl.u_bound, ctrl.max_iter,
                                       ctrl.grad_tol, ctrl.lbfgs_memory, meta)
        meta = {"algorithm": ("MR Tracking GRAPE" if tracking else "MR GRAPE"), "n_ensemble": n_ens}
        return optimcon_grape_loop(fg, guess, ctrl.l_bound, ctrl.u_bound, ctrl.max_iter, ctrl.grad_tol, meta)
    finally:
        kp.close()
