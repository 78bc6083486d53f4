"""Host-side mirror of Pulsar.jl's GRAPE interface, executed on the :rocm device.

Same names, argument meaning and error behaviour as the reference's Julia API
for this path; every fidelity and gradient is computed by libpulsar_rocm.so:

  ControlSequence, random_controls, zero_controls   src/Types/ControlSequence.jl:37-158
  QuantumSystem, quantum_system                     src/Types/QuantumSystem.jl:60-66, 318-341
  QuantumTarget, state_target, unitary_target       src/Types/Targets.jl:37-144
  compute_fidelity(sys, ctrl, tgt)                  src/Physics/Objectives.jl:609-639
  compute_grape_gradient(sys, ctrl, tgt)            src/Physics/Gradients.jl:80-127
  GRAPEConfig, OptimizationResult, grape_optimize   src/Optimization/GRAPE.jl:60-211, 469-643
  grape_optimize_ensemble (L-BFGS)                  src/Optimization/GRAPE.jl:1012-1230
  grape_gradient, evaluate_fidelity                 src/Optimization/GRAPE.jl:1462-1490
  grape_step!, adapt_grape_step_size,
  check_grape_convergence                           src/Optimization/GRAPE.jl:261-402
  EnsembleObjective, ensemble_value(_and_grad),
  build_ensemble_from_systems                       src/Optimization/Ensemble/EnsembleObjective.jl:77-307,
                                                    SystemBuilder.jl:49-87
  set_device / get_device / available_devices /
  with_device (with :rocm registered)               src/Backend/Scheduling/DeviceRegistry.jl:51-151

There is no CPU execution path in this package: the CPU device is Pulsar.jl
itself. When the library or a gfx950 GPU is missing, calls raise RocmError.
"""
from __future__ import annotations

import contextlib
import threading
import time
import warnings
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

from . import native
from .problem import (PR_CLOSED, PR_FP64, PR_GATE, PR_GRAD_EXACT, PR_GRAD_REFERENCE, PR_LINDBLAD,
                      PR_PAIRS_REAL, PR_PAIRS_SQUARE, PR_STATE, Problem)

# --------------------------------------------------------------------------- types


class ControlSequence:
    """Piecewise-constant amplitudes [n_steps x n_controls] (ControlSequence.jl:37-59).

    `controls` is the [n_controls x n_steps] view (ControlSequence.jl:99-110).
    """

    def __init__(self, amplitudes, dt: float, n_steps: int):
        a = np.asarray(amplitudes, dtype=np.float64)
        if dt < 0:
            raise ValueError(f"dt must be non-negative, got {dt}")
        if n_steps < 0:
            raise ValueError(f"n_steps must be non-negative, got {n_steps}")
        if a.ndim != 2 or a.shape[0] != n_steps:
            raise ValueError(f"amplitudes has {a.shape[0] if a.ndim == 2 else '?'} rows but n_steps = {n_steps}; "
                             "expected layout [n_steps x n_controls]")
        self.amplitudes = a
        self.dt = float(dt)
        self.n_steps = int(n_steps)

    @classmethod
    def from_controls(cls, controls, dt: float, total_time: float, n_timesteps: int):
        """4-argument form taking [n_controls x n_timesteps] (ControlSequence.jl:77-89)."""
        c = np.asarray(controls, dtype=np.float64)
        if c.shape[1] != n_timesteps:
            raise ValueError(f"controls has {c.shape[1]} columns but n_timesteps = {n_timesteps}")
        return cls(c.T.copy(), dt, n_timesteps)

    @property
    def controls(self) -> np.ndarray:
        return self.amplitudes.T

    @property
    def n_timesteps(self) -> int:
        return self.n_steps

    @property
    def total_time(self) -> float:
        return self.dt * self.n_steps


@dataclass
class QuantumSystem:
    """H_drift, H_controls, dim, n_controls (QuantumSystem.jl:60-66)."""

    H_drift: np.ndarray
    H_controls: List[np.ndarray]
    dim: int = 0
    n_controls: int = 0
    metadata: dict = field(default_factory=dict)

    def __post_init__(self):
        self.H_drift = np.asarray(self.H_drift, dtype=np.complex128)
        self.H_controls = [np.asarray(h, dtype=np.complex128) for h in self.H_controls]
        self.dim = self.H_drift.shape[0]
        self.n_controls = len(self.H_controls)


def _check_hermitian(H, name, tol=1e-10):
    dev = np.abs(H - H.conj().T).max() if H.size else 0.0
    if dev > tol:
        raise ValueError(f"{name} is not Hermitian: max|H - H^dag| = {dev} > tol = {tol}")


def quantum_system(H_drift, H_controls, metadata=None) -> QuantumSystem:
    """quantum_system (QuantumSystem.jl:318-341): validates shapes and Hermiticity."""
    sys_ = QuantumSystem(H_drift, list(H_controls), metadata=dict(metadata or {}))
    d = sys_.dim
    if sys_.H_drift.shape != (d, d):
        raise ValueError("H_drift must be square")
    _check_hermitian(sys_.H_drift, "H_drift")
    for j, h in enumerate(sys_.H_controls):
        if h.shape != (d, d):
            raise ValueError(f"H_controls[{j}] must be {d} x {d}")
        _check_hermitian(h, f"H_controls[{j}]")
    return sys_


def random_controls(system, total_time: float, n_timesteps: int, amplitude: float = 1.0,
                    rng: Optional[np.random.Generator] = None) -> ControlSequence:
    """amplitude * (2 rand - 1), [n_steps x n_controls] (ControlSequence.jl:142-158)."""
    if total_time <= 0:
        raise ValueError(f"total_time must be positive, got {total_time}")
    if n_timesteps <= 0:
        raise ValueError(f"n_timesteps must be positive, got {n_timesteps}")
    rng = rng or np.random.default_rng()
    a = amplitude * (2 * rng.random((n_timesteps, system.n_controls)) - 1)
    return ControlSequence(a, total_time / n_timesteps, n_timesteps)


def zero_controls(system, total_time: float, n_timesteps: int) -> ControlSequence:
    return ControlSequence(np.zeros((n_timesteps, system.n_controls)), total_time / n_timesteps, n_timesteps)


@dataclass
class QuantumTarget:
    """type ('state' | 'unitary'), target_state, target_unitary, initial_state, dim (Targets.jl:37-44)."""

    type: str
    target_state: Optional[np.ndarray]
    target_unitary: Optional[np.ndarray]
    initial_state: Optional[np.ndarray]
    dim: int


def state_target(state, psi_init=None) -> QuantumTarget:
    """state_target (Targets.jl:79-106); like the reference, the state is not normalized."""
    psi = np.asarray(state, dtype=np.complex128).reshape(-1)
    if psi.size == 0:
        raise ValueError("Target state vector must be non-empty")
    if not np.all(np.isfinite(psi)):
        raise ValueError("Target state contains NaN or Inf values")
    if np.linalg.norm(psi) < np.finfo(float).eps:
        raise ValueError("Target state has zero norm")
    psi_i = None
    if psi_init is not None:
        psi_i = np.asarray(psi_init, dtype=np.complex128).reshape(-1)
        if psi_i.size != psi.size:
            raise ValueError(f"psi_init has length {psi_i.size}; expected {psi.size}")
        if not np.all(np.isfinite(psi_i)):
            raise ValueError("psi_init contains NaN or Inf values")
        if np.linalg.norm(psi_i) < np.finfo(float).eps:
            raise ValueError("psi_init has zero norm")
    return QuantumTarget("state", psi, None, psi_i, psi.size)


def unitary_target(U) -> QuantumTarget:
    """unitary_target (Targets.jl:128-144): square and ||U^dag U - I||/dim <= 1e-8."""
    Uc = np.asarray(U, dtype=np.complex128)
    if Uc.ndim != 2 or Uc.shape[0] != Uc.shape[1]:
        raise ValueError(f"Target unitary must be square, got {Uc.shape}")
    m = Uc.shape[0]
    dev = np.linalg.norm(Uc.conj().T @ Uc - np.eye(m)) / m
    if dev > 1e-8:
        raise ValueError(f"Target matrix is not unitary: ||U^dag U - I|| / dim = {dev} > 1e-8")
    return QuantumTarget("unitary", None, Uc, None, m)


# --------------------------------------------------------------------------- devices

_DEVICES = ("cpu", "cuda", "metal", "rocm")
_device_default = "cpu"
_device_local = threading.local()


def rocm_available() -> bool:
    try:
        return native.device_count() > 0
    except native.RocmError:
        return False


def set_device(device: str) -> None:
    """set_device! (DeviceRegistry.jl:51-72) with :rocm registered beside :cpu/:cuda/:metal.
    An unavailable device warns and keeps the current one."""
    global _device_default
    device = device.lstrip(":")
    if device not in _DEVICES:
        raise ValueError(f"Unknown device :{device}. Choose :cpu, :cuda, :metal or :rocm.")
    if device in ("cuda", "metal") or (device == "rocm" and not rocm_available()):
        warnings.warn(f"set_device(:{device}): backend not available here; keeping current device :{get_device()}.")
        return
    _device_local.device = device
    _device_default = device


def get_device() -> str:
    return getattr(_device_local, "device", _device_default)


def available_devices() -> List[str]:
    return ["cpu"] + (["rocm"] if rocm_available() else [])


@contextlib.contextmanager
def with_device(device: str):
    device = device.lstrip(":")
    if device not in _DEVICES:
        raise ValueError(f"Unknown device :{device}. Choose :cpu, :cuda, :metal or :rocm.")
    prev = getattr(_device_local, "device", None)
    _device_local.device = device
    try:
        yield
    finally:
        if prev is None:
            del _device_local.device
        else:
            _device_local.device = prev


# --------------------------------------------------------------------------- problems


def _state_problem(system, n_t, dt, psi0, psit, normalize: bool) -> Problem:
    if normalize:
        psi0 = psi0 / np.linalg.norm(psi0)
        psit = psit / np.linalg.norm(psit)
    return Problem(dim=system.dim, n_ctrl=system.n_controls, n_t=n_t, H0=system.H_drift,
                   Hc=np.stack(system.H_controls), dt=dt, dynamics=PR_CLOSED, objective=PR_STATE,
                   n_members=1, psi_init=psi0[None, :], psi_targ=psit[None, :], weight=np.ones(1))


def _gate_problem(system, n_t, dt, Ut) -> Problem:
    return Problem(dim=system.dim, n_ctrl=system.n_controls, n_t=n_t, H0=system.H_drift,
                   Hc=np.stack(system.H_controls), dt=dt, dynamics=PR_CLOSED, objective=PR_GATE,
                   n_members=1, U_target=Ut)


def _check_target(target: QuantumTarget):
    if target.type == "unitary":
        if target.target_unitary is None:
            raise ValueError('target.type is "unitary" but target.target_unitary is nothing')
    elif target.type == "state":
        if target.target_state is None:
            raise ValueError('target.type is "state" but target.target_state is nothing')
    elif target.type == "subspace":
        raise NotImplementedError("Subspace fidelity is not yet implemented")
    else:
        raise ValueError(f'Unknown target type "{target.type}"; expected "unitary" or "state"')


def _check_controls(system, controls: ControlSequence):
    if controls.amplitudes.shape[1] != system.n_controls:
        raise ValueError(f"controls has {controls.amplitudes.shape[1]} control channels but system has "
                         f"{system.n_controls} control fields")


def fidelity_problem(system, controls: ControlSequence, target: QuantumTarget) -> Problem:
    """compute_fidelity's objective: |Tr(Ut^dag U)/d|^2, or |<psi_t|U psi_i>|^2 with the
    raw (unnormalized) states and psi_i = initial_state or target_state (Objectives.jl:609-630)."""
    _check_target(target)
    _check_controls(system, controls)
    if target.type == "unitary":
        return _gate_problem(system, controls.n_steps, controls.dt, target.target_unitary)
    psi_i = target.target_state if target.initial_state is None else target.initial_state
    return _state_problem(system, controls.n_steps, controls.dt, psi_i, target.target_state, normalize=False)


def gradient_problem(system, controls: ControlSequence, target: QuantumTarget, in_optimizer: bool = False) -> Problem:
    """compute_grape_gradient's objective: states normalized (Gradients.jl:369-374); psi_i is
    initial_state or target_state (Gradients.jl:115-116), and inside grape_optimize always
    target_state (compute_grape_gradient_with!, Gradients.jl:156-157)."""
    _check_target(target)
    _check_controls(system, controls)
    if target.type == "unitary":
        return _gate_problem(system, controls.n_steps, controls.dt, target.target_unitary)
    if in_optimizer or target.initial_state is None:
        psi_i = target.target_state
    else:
        psi_i = target.initial_state
    return _state_problem(system, controls.n_steps, controls.dt, psi_i, target.target_state, normalize=True)


def compute_fidelity(system, controls: ControlSequence, target: QuantumTarget) -> float:
    """compute_fidelity(system, controls, target) on the ROCm device (Objectives.jl:632-639)."""
    p = fidelity_problem(system, controls, target)
    plan = native.Plan([p])
    try:
        F, _ = plan.eval([controls.controls], grad=False)
    finally:
        plan.close()
    return float(F[0])


def compute_grape_gradient(system, controls: ControlSequence, target: QuantumTarget,
                           grad_mode: int = PR_GRAD_REFERENCE) -> np.ndarray:
    """compute_grape_gradient (Gradients.jl:80-127): G[j, k] = dF/du_j[k], shape [n_c x n_t].
    grad_mode=PR_GRAD_EXACT gives the exact derivative instead of the first-order formula."""
    p = gradient_problem(system, controls, target)
    plan = native.Plan([p], grad_mode=grad_mode)
    try:
        _, G = plan.eval([controls.controls], grad=True)
    finally:
        plan.close()
    return G[0]


# --------------------------------------------------------------------------- GRAPE loop


@dataclass
class GRAPEConfig:
    """GRAPEConfig (GRAPE.jl:60-127); `backend` and `use_threads` are stored, as in the reference."""

    max_iter: int = 1000
    convergence_tol: float = 1e-8
    step_size: float = 0.01
    adapt_step_size: bool = True
    min_step_size: float = 1e-10
    max_step_size: float = 10.0
    gradient_norm_tol: float = 1e-6
    verbose: bool = False
    print_interval: int = 100
    callback: Optional[Callable] = None
    use_threads: bool = True
    backend: str = "auto"

    def __post_init__(self):
        if self.max_iter <= 0:
            raise ValueError("max_iter must be positive")
        if self.convergence_tol <= 0:
            raise ValueError("convergence_tol must be positive")
        if self.step_size <= 0:
            raise ValueError("step_size must be positive")
        if self.min_step_size <= 0:
            raise ValueError("min_step_size must be positive")
        if self.max_step_size < self.step_size:
            raise ValueError("max_step_size must be >= step_size")
        if self.min_step_size > self.step_size:
            raise ValueError("min_step_size must be <= step_size")
        if self.gradient_norm_tol <= 0:
            raise ValueError("gradient_norm_tol must be positive")
        if self.print_interval <= 0:
            raise ValueError("print_interval must be positive")
        if self.backend not in ("auto", "cpu", "gpu", "hybrid"):
            raise ValueError("backend must be :auto, :cpu, :gpu, or :hybrid")


@dataclass
class OptimizationResult:
    controls: np.ndarray
    fidelity: float
    fidelity_history: List[float]
    gradient_norm_history: List[float]
    n_iterations: int
    converged: bool
    termination_reason: str
    total_time: float
    n_fidelity_evaluations: int
    n_gradient_evaluations: int
    metadata: dict


def grape_step(controls: np.ndarray, gradient: np.ndarray, step_size: float) -> np.ndarray:
    """grape_step! (GRAPE.jl:261-268): controls += step_size * gradient, in place."""
    controls += step_size * gradient
    return controls


def adapt_grape_step_size(fidelity_history, current_step, min_step=1e-10, max_step=10.0,
                          growth_factor=1.05, shrink_factor=0.5) -> float:
    """adapt_grape_step_size (GRAPE.jl:309-334)."""
    n = len(fidelity_history)
    if n < 2:
        return current_step
    delta = fidelity_history[n - 1] - fidelity_history[n - 2]
    if delta > 0.0:
        new = current_step * growth_factor
    elif delta < -1e-14:
        new = current_step * shrink_factor
    else:
        new = current_step
    return min(max(new, min_step), max_step)


def check_grape_convergence(fidelity_history, gradient_norm, config: GRAPEConfig):
    """check_grape_convergence (GRAPE.jl:371-402)."""
    n = len(fidelity_history)
    if n == 0:
        return False, ""
    cur = fidelity_history[-1]
    if cur >= 1.0 - 1e-12:
        return True, "perfect fidelity (F >= 1 - 1e-12)"
    if gradient_norm < config.gradient_norm_tol:
        return True, f"gradient norm {gradient_norm} < tol {config.gradient_norm_tol}"
    if n >= 3:
        dF = abs(fidelity_history[-1] - fidelity_history[-2])
        if dF < config.convergence_tol:
            return True, f"fidelity change {dF} < tol {config.convergence_tol}"
    return False, ""


def grape_optimize(system, target: QuantumTarget, controls_init: ControlSequence,
                   config: Optional[GRAPEConfig] = None, max_iter: Optional[int] = None,
                   step_size: Optional[float] = None, on_device: bool = True) -> OptimizationResult:
    """grape_optimize (GRAPE.jl:469-643, with the positional-max_iter overload :665-675).

    The loop logic (history, adaptation, convergence, best-controls tracking,
    final re-evaluation) follows the reference. With on_device (the default,
    when no callback or printing is asked for) the whole loop runs in the
    library (pr_grape_optimize): controls, gradient and best controls stay in
    HBM and each iteration returns only F and |G|. Otherwise this host loop
    drives one device evaluation per iteration."""
    if config is None:
        config = GRAPEConfig(**({"max_iter": max_iter} if max_iter else {}),
                             **({"step_size": step_size} if step_size else {}))
    if on_device and config.callback is None and not config.verbose:
        return _grape_optimize_device(system, target, controls_init, config)
    return _grape_optimize_host(system, target, controls_init, config)


def _reason(code: int, f_hist, g_hist, config: GRAPEConfig) -> str:
    if code == 1:
        return "perfect fidelity (F >= 1 - 1e-12)"
    if code == 2:
        return f"gradient norm {g_hist[-1]} < tol {config.gradient_norm_tol}"
    if code == 3:
        # the test ran before the post-convergence evaluation was appended
        dF = abs(f_hist[-2] - f_hist[-3])
        return f"fidelity change {dF} < tol {config.convergence_tol}"
    return "maximum iterations reached"


def _grape_optimize_device(system, target: QuantumTarget, controls_init: ControlSequence,
                           config: GRAPEConfig) -> OptimizationResult:
    t_start = time.time()
    dt, n_t, n_c = controls_init.dt, controls_init.n_timesteps, system.n_controls
    if controls_init.controls.shape != (n_c, n_t):
        raise ValueError(f"controls_init.controls size {controls_init.controls.shape} does not match ({n_c}, {n_t})")
    seq = ControlSequence(np.asarray(controls_init.controls).T, dt, n_t)
    fplan = native.Plan([fidelity_problem(system, seq, target)])
    gplan = None if target.type == "unitary" else native.Plan([gradient_problem(system, seq, target,
                                                                                 in_optimizer=True)])
    cfg = native.GrapeConfig(config.max_iter, config.convergence_tol, config.step_size,
                             1 if config.adapt_step_size else 0, config.min_step_size, config.max_step_size,
                             config.gradient_norm_tol)
    try:
        best, f_hist, g_hist, res = native.grape_optimize_device(fplan, gplan, controls_init.controls, cfg)
    finally:
        fplan.close()
        if gplan is not None:
            gplan.close()
    return OptimizationResult(best, float(res.fidelity), f_hist, g_hist, int(res.n_iterations), bool(res.converged),
                              _reason(res.reason, f_hist, g_hist, config), time.time() - t_start,
                              int(res.n_fidelity_evaluations), int(res.n_gradient_evaluations),
                              {"step_size_final": float(res.step_size_final), "algorithm": "GRAPE",
                               "n_controls": n_c, "n_timesteps": n_t, "dt": dt, "total_time": dt * n_t})


def _grape_optimize_host(system, target: QuantumTarget, controls_init: ControlSequence,
                         config: GRAPEConfig) -> OptimizationResult:
    t_start = time.time()
    dt, n_t, n_c = controls_init.dt, controls_init.n_timesteps, system.n_controls
    if controls_init.controls.shape != (n_c, n_t):
        raise ValueError(f"controls_init.controls size {controls_init.controls.shape} does not match ({n_c}, {n_t})")
    controls = np.array(controls_init.controls, dtype=np.float64)
    best = controls.copy()
    f_hist: List[float] = []
    g_hist: List[float] = []
    n_f = n_g = 0
    step = config.step_size
    converged, reason, n_iter = False, "maximum iterations reached", 0

    seq = ControlSequence(controls.T, dt, n_t)
    fplan = native.Plan([fidelity_problem(system, seq, target)])
    if target.type == "unitary":
        gplan = fplan
    else:
        gplan = native.Plan([gradient_problem(system, seq, target, in_optimizer=True)])

    def fid(c):
        return float(fplan.eval([c], grad=False)[0][0])

    try:
        F = fid(controls)
        n_f += 1
        f_hist.append(F)
        best_F = F
        if config.verbose:
            print(f"[GRAPE] Starting: F0 = {F:.8f}, n_params = {n_c * n_t}")
        for it in range(1, config.max_iter + 1):
            n_iter = it
            if gplan is fplan:
                Fa, Ga = fplan.eval([controls], grad=True)
                F, G = float(Fa[0]), Ga[0]
            else:
                F = fid(controls)
                G = gplan.eval([controls], grad=True)[1][0]
            n_f += 1
            n_g += 1
            gnorm = float(np.linalg.norm(G))
            f_hist.append(F)
            g_hist.append(gnorm)
            if F > best_F:
                best_F = F
                best[:] = controls
            if config.adapt_step_size and len(f_hist) >= 2:
                step = adapt_grape_step_size(f_hist, step, min_step=config.min_step_size,
                                             max_step=config.max_step_size)
            grape_step(controls, G, step)
            if config.verbose and (it % config.print_interval == 0 or it == 1):
                print(f"[GRAPE] iter={it:5d}  F={F:.8f}  |dF|={gnorm:.3e}  a={step:.3e}")
            if config.callback is not None:
                config.callback(it, F, gnorm, controls)
            conv, why = check_grape_convergence(f_hist, gnorm, config)
            if conv:
                converged, reason = True, why
                Fp = fid(controls)
                f_hist.append(Fp)
                n_f += 1
                if Fp > best_F:
                    best_F = Fp
                    best[:] = controls
                break
        F_final = fid(best)
        n_f += 1
    finally:
        fplan.close()
        if gplan is not fplan:
            gplan.close()
    elapsed = time.time() - t_start
    if config.verbose:
        print(f"[GRAPE] Done: F={F_final:.8f}  iters={n_iter}  time={elapsed:.3f} s  converged={converged}")
    return OptimizationResult(best, F_final, f_hist, g_hist, n_iter, converged, reason, elapsed, n_f, n_g,
                              {"step_size_final": step, "algorithm": "GRAPE", "n_controls": n_c,
                               "n_timesteps": n_t, "dt": dt, "total_time": dt * n_t})


def grape_gradient(system, target: QuantumTarget, u, dt: float, parallel: bool = False) -> np.ndarray:
    """grape_gradient (GRAPE.jl:1462-1472): compute_grape_gradient on a raw
    [n_c x n_t] control matrix; `parallel` is accepted and ignored, as there."""
    u = np.asarray(u, dtype=np.float64)
    return compute_grape_gradient(system, ControlSequence(u.T, float(dt), u.shape[1]), target)


def evaluate_fidelity(system, target: QuantumTarget, u, dt: float) -> float:
    """evaluate_fidelity (GRAPE.jl:1474-1490): compute_fidelity of the total
    propagator of a raw [n_c x n_t] control matrix."""
    u = np.asarray(u, dtype=np.float64)
    return compute_fidelity(system, ControlSequence(u.T, float(dt), u.shape[1]), target)


# --------------------------------------------------------------------------- ensembles


class EnsembleObjective:
    """EnsembleObjective (EnsembleObjective.jl:77-163): per-sample closures or a
    batched_eval(theta) -> (F, dF) fast path (aggregator :mean only)."""

    def __init__(self, f_samples=None, grad_samples=None, aggregator="mean", cvar_alpha=0.2,
                 resample=None, batched_eval=None, n_samples=None):
        aggregator = aggregator.lstrip(":")
        if aggregator not in ("mean", "worst_case", "cvar"):
            raise ValueError(f"aggregator must be :mean, :worst_case, or :cvar; got {aggregator}")
        if not 0 < cvar_alpha <= 1:
            raise ValueError(f"cvar_alpha must lie in (0, 1]; got {cvar_alpha}")
        f_samples = list(f_samples or [])
        if batched_eval is not None:
            if aggregator != "mean":
                raise ValueError(f"batched_eval fast path requires aggregator=:mean; got {aggregator}")
            if n_samples is None or n_samples < 1:
                raise ValueError("n_samples must be positive in batched mode")
        else:
            n_samples = len(f_samples) if n_samples is None else n_samples
            if n_samples != len(f_samples):
                raise ValueError(f"n_samples ({n_samples}) does not match length(f_samples) ({len(f_samples)})")
            if grad_samples is not None and len(grad_samples) != n_samples:
                raise ValueError("length(grad_samples) must equal length(f_samples)")
        self.f_samples = f_samples
        self.grad_samples = grad_samples
        self.aggregator = aggregator
        self.cvar_alpha = float(cvar_alpha)
        self.resample = resample
        self.batched_eval = batched_eval
        self.n_samples = n_samples


def _agg_value(fv, aggregator, alpha):
    if aggregator == "mean":
        return float(np.mean(fv))
    if aggregator == "worst_case":
        return float(np.min(fv))
    n_tail = max(1, int(round(alpha * len(fv))))
    return float(np.mean(np.sort(fv)[:n_tail]))


def _agg_grad(fv, gv, aggregator, alpha):
    gv = np.asarray(gv)
    if aggregator == "mean":
        return gv.sum(axis=0) / len(gv)
    if aggregator == "worst_case":
        return gv[int(np.argmin(fv))].copy()
    n_tail = max(1, int(round(alpha * len(fv))))
    order = np.argsort(fv, kind="stable")
    return gv[order[:n_tail]].sum(axis=0) / n_tail


def ensemble_value(obj: EnsembleObjective, theta) -> float:
    if obj.resample is not None:
        obj.resample(obj)
    if obj.batched_eval is not None:
        value = getattr(obj.batched_eval, "value", None)
        return float(value(theta)) if value is not None else float(obj.batched_eval(theta)[0])
    fv = np.array([f(theta) for f in obj.f_samples])
    return _agg_value(fv, obj.aggregator, obj.cvar_alpha)


def ensemble_value_and_grad(obj: EnsembleObjective, theta):
    if obj.resample is not None:
        obj.resample(obj)
    if obj.batched_eval is not None:
        F, G = obj.batched_eval(theta)
        return float(F), np.asarray(G, dtype=np.float64).reshape(-1, order="F")
    if obj.grad_samples is None:
        raise ValueError("ensemble_value_and_grad called without grad_samples; use ensemble_value for "
                         "derivative-free optimizers.")
    fv = np.array([f(theta) for f in obj.f_samples])
    gv = [g(theta) for g in obj.grad_samples]
    return _agg_value(fv, obj.aggregator, obj.cvar_alpha), _agg_grad(fv, gv, obj.aggregator, obj.cvar_alpha)


def _scale_factors(ref: Sequence[np.ndarray], hcs: Sequence[np.ndarray]):
    """Real s_j with hcs[j] == s_j * ref[j] (as _scale_controls produces,
    Platforms/Common.jl:64-69), or None."""
    s = []
    for a, b in zip(ref, hcs):
        na = np.vdot(a, a).real
        if na == 0:
            if np.abs(b).max() != 0:
                return None
            s.append(1.0)
            continue
        c = np.vdot(a, b) / na
        if abs(c.imag) > 1e-14 * abs(c) or np.abs(b - c.real * a).max() > 1e-13 * max(np.abs(b).max(), 1e-300):
            return None
        s.append(c.real)
    return np.array(s)


def ensemble_problems(systems, target: QuantumTarget, n_t: int, dt: float, in_optimizer: bool = True):
    """Problems (one per distinct control set, members = systems sharing it up to
    per-control scale factors) whose F and G sum to the ensemble mean."""
    _check_target(target)
    N = len(systems)
    groups = []  # (ref controls, [(index, scales)])
    for i, s in enumerate(systems):
        for ref, members in groups:
            sc = _scale_factors(ref, s.H_controls)
            if sc is not None:
                members.append((i, sc))
                break
        else:
            groups.append((s.H_controls, [(i, np.ones(len(s.H_controls)))]))
    probs = []
    for ref, members in groups:
        idx = [i for i, _ in members]
        H0 = np.stack([systems[i].H_drift for i in idx])
        sc = np.stack([s for _, s in members])
        w = np.full(len(idx), 1.0 / N)
        common = dict(dim=systems[idx[0]].dim, n_ctrl=len(ref), n_t=n_t, H0=H0, Hc=np.stack(ref), dt=dt,
                      dynamics=PR_CLOSED, n_members=len(idx), ctrl_scale=sc, weight=w)
        if target.type == "unitary":
            probs.append(Problem(objective=PR_GATE, U_target=target.target_unitary, **common))
        else:
            psit = target.target_state / np.linalg.norm(target.target_state)
            psi_i = target.target_state if (in_optimizer or target.initial_state is None) else target.initial_state
            psi_i = psi_i / np.linalg.norm(psi_i)
            probs.append(Problem(objective=PR_STATE, psi_init=psi_i[None], psi_targ=psit[None], **common))
    return probs


class BatchedEnsemble:
    """batched_eval(theta) -> (F, G) on the device for an ensemble of systems."""

    def __init__(self, systems, target, controls: ControlSequence, grad_mode=PR_GRAD_REFERENCE,
                 precision=PR_FP64):
        for s in systems:
            _check_controls(s, controls)
        self.n_c, self.n_t = systems[0].n_controls, controls.n_steps
        self.probs = ensemble_problems(systems, target, controls.n_steps, controls.dt)
        self.plan = native.Plan(self.probs, grad_mode=grad_mode, precision=precision)

    def __call__(self, theta):
        u = np.asarray(theta, dtype=np.float64).reshape(self.n_c, self.n_t, order="F")
        F, G = self.plan.eval([u] * len(self.probs), grad=True)
        return float(F.sum()), np.sum(G, axis=0)

    def value(self, theta) -> float:
        """F only (the gradient walk is skipped), for line searches."""
        u = np.asarray(theta, dtype=np.float64).reshape(self.n_c, self.n_t, order="F")
        F, _ = self.plan.eval([u] * len(self.probs), grad=False)
        return float(F.sum())

    def close(self):
        self.plan.close()


def build_ensemble_from_systems(systems, target: QuantumTarget, controls: ControlSequence,
                                aggregator: str = "mean", cvar_alpha: float = 0.2) -> EnsembleObjective:
    """build_ensemble_from_systems (SystemBuilder.jl:49-87). The :mean aggregator
    takes the batched device path (one launch sequence for the whole ensemble);
    :worst_case and :cvar use one device plan per member, as per-sample closures."""
    if len(systems) == 0:
        raise ValueError("systems must be non-empty")
    aggregator = aggregator.lstrip(":")
    if aggregator == "mean":
        return EnsembleObjective(batched_eval=BatchedEnsemble(systems, target, controls),
                                 n_samples=len(systems))
    n_c, n_t, dt = systems[0].n_controls, controls.n_steps, controls.dt

    def mk(s):
        plan = native.Plan(ensemble_problems([s], target, n_t, dt))

        def f(theta):
            u = np.asarray(theta, np.float64).reshape(n_c, n_t, order="F")
            return float(plan.eval([u], grad=False)[0][0])

        def g(theta):
            u = np.asarray(theta, np.float64).reshape(n_c, n_t, order="F")
            return plan.eval([u], grad=True)[1][0].reshape(-1, order="F")

        return f, g

    fs, gs = zip(*[mk(s) for s in systems])
    return EnsembleObjective(list(fs), grad_samples=list(gs), aggregator=aggregator, cvar_alpha=cvar_alpha,
                             n_samples=len(systems))


# --------------------------------------------------------------------------- ensemble L-BFGS


def _lbfgs_direction(g, s_list, y_list):
    """L-BFGS two-loop recursion (Nocedal & Wright Alg. 7.4) returning an ascent
    direction, with the safeguards of GRAPE.jl:1062-1080."""
    k = len(s_list)
    if k == 0:
        return g.copy()
    q = g.copy()
    alphas = np.zeros(k)
    for i in range(k - 1, -1, -1):
        rho = 1.0 / max(float(np.dot(y_list[i], s_list[i])), 1e-14)
        alphas[i] = rho * float(np.dot(s_list[i], q))
        q -= alphas[i] * y_list[i]
    gamma = float(np.dot(s_list[-1], y_list[-1])) / max(float(np.dot(y_list[-1], y_list[-1])), 1e-14)
    q *= gamma
    for i in range(k):
        rho = 1.0 / max(float(np.dot(y_list[i], s_list[i])), 1e-14)
        beta = rho * float(np.dot(y_list[i], q))
        q += (alphas[i] - beta) * s_list[i]
    return q


def _ensemble_lbfgs(F_fn, G_fn, controls0, config, lbfgs_memory=10, amplitude_limit=np.inf):
    """The loop of grape_optimize_ensemble (GRAPE.jl:1012-1230) over an
    evaluator pair F_fn(u) -> float, G_fn(u) -> [n_c x n_t]: normalised L-BFGS
    ascent direction, backtracking line search accepting F_new >= F0 (step
    halved up to 20 times), curvature pairs kept when y.s > 1e-14 |s|^2,
    amplitude clipping, and the reference's three convergence tests."""
    controls = np.array(controls0, dtype=np.float64)
    f_hist, g_hist = [], []
    n_f = n_g = 0
    s_list, y_list = [], []
    clip = np.isfinite(amplitude_limit)

    def line_search(u, d, F0, a0, max_bt=20):
        a, n_ls = a0, 0
        for _ in range(max_bt):
            un = u + a * d
            if clip:
                np.clip(un, -amplitude_limit, amplitude_limit, out=un)
            Fn = F_fn(un)
            n_ls += 1
            if Fn >= F0:
                return a, un, Fn, n_ls
            a *= 0.5
        un = u + a * d
        if clip:
            np.clip(un, -amplitude_limit, amplitude_limit, out=un)
        return a, un, F_fn(un), max_bt + 1

    F = F_fn(controls)
    n_f += 1
    f_hist.append(F)
    best_F, best = F, controls.copy()
    converged, reason = False, "maximum iterations reached"
    for it in range(1, config.max_iter + 1):
        G = np.asarray(G_fn(controls), dtype=np.float64)
        n_g += 1
        g = G.reshape(-1, order="F")
        gnorm = float(np.linalg.norm(g))
        g_hist.append(gnorm)
        d = _lbfgs_direction(g, s_list, y_list)
        dmax = float(np.max(np.abs(d)))
        if dmax < 1e-14:
            converged, reason = True, "zero gradient direction"
            break
        d /= dmax
        dm = d.reshape(controls.shape, order="F")
        a0 = config.step_size * amplitude_limit if clip else config.step_size
        a, cn, Fn, n_ls = line_search(controls, dm, F, a0)
        n_f += n_ls
        Gn = np.asarray(G_fn(cn), dtype=np.float64)
        n_g += 1
        sv = a * d
        yv = Gn.reshape(-1, order="F") - g
        if float(np.dot(yv, sv)) > 1e-14 * float(np.dot(sv, sv)):
            s_list.append(sv)
            y_list.append(yv)
            if len(s_list) > lbfgs_memory:
                s_list.pop(0)
                y_list.pop(0)
        controls, F = cn, Fn
        f_hist.append(F)
        if F > best_F:
            best_F, best = F, controls.copy()
        if config.verbose and (it % config.print_interval == 0 or it == 1):
            print(f"[Ensemble L-BFGS-GRAPE] iter={it:4d}  F_bb={F:.8f}  a={a:.3e}  |dF|={gnorm:.3e}")
        if config.callback is not None:
            config.callback(it, F, gnorm, controls)
        if gnorm < config.gradient_norm_tol:
            converged, reason = True, f"gradient norm < {config.gradient_norm_tol}"
            break
        if F >= 1.0 - 1e-8:
            converged, reason = True, "near-perfect fidelity"
            break
        if len(f_hist) >= 2 and abs(f_hist[-1] - f_hist[-2]) < config.convergence_tol:
            converged, reason = True, f"fidelity change < {config.convergence_tol}"
            break
    return best, best_F, f_hist, g_hist, converged, reason, n_f, n_g


def grape_optimize_ensemble(systems, target: QuantumTarget, controls_init: ControlSequence,
                            config: Optional[GRAPEConfig] = None, lbfgs_memory: int = 10,
                            amplitude_limit: float = np.inf, on_device: bool = True) -> OptimizationResult:
    """grape_optimize_ensemble (GRAPE.jl:1012-1230): broadband L-BFGS-GRAPE over
    an ensemble of systems, maximising the ensemble-mean fidelity. Every F and
    gradient is one batched device evaluation of the whole ensemble (line-search
    steps evaluate F only). With on_device (and no callback or printing, which
    need the Python loop) the whole loop runs in the library
    (pr_lbfgs_optimize): controls, gradients and curvature pairs stay on the
    device."""
    config = config or GRAPEConfig()
    if len(systems) == 0:
        raise ValueError("systems must be non-empty")
    if lbfgs_memory <= 0:
        raise ValueError("lbfgs_memory must be positive")
    t0 = time.time()
    obj = build_ensemble_from_systems(systems, target, controls_init, aggregator="mean")
    shape = controls_init.controls.shape
    be = obj.batched_eval
    if (on_device and config.callback is None and not config.verbose and len(be.probs) == 1
            and lbfgs_memory <= 64):
        try:
            cfg = native.LbfgsConfig(int(config.max_iter), float(config.convergence_tol), float(config.step_size),
                                     float(config.gradient_norm_tol), int(lbfgs_memory),
                                     float(amplitude_limit) if np.isfinite(amplitude_limit) else 0.0)
            best, f_hist, g_hist, res = native.lbfgs_optimize_device(be.plan, controls_init.controls, cfg)
        finally:
            be.close()
        reason = {native.PR_STOP_CODES["gradient"]: f"gradient norm < {config.gradient_norm_tol}",
                  native.PR_STOP_CODES["perfect"]: "near-perfect fidelity",
                  native.PR_STOP_CODES["fidelity"]: f"fidelity change < {config.convergence_tol}",
                  native.PR_STOP_CODES["zero_direction"]: "zero gradient direction"}.get(
                      res.reason, "maximum iterations reached")
        return OptimizationResult(best, float(res.fidelity), f_hist, g_hist, len(f_hist), bool(res.converged),
                                  reason, time.time() - t0, int(res.n_fidelity_evaluations),
                                  int(res.n_gradient_evaluations),
                                  {"algorithm": "Ensemble L-BFGS-GRAPE", "n_systems": len(systems),
                                   "lbfgs_memory": lbfgs_memory, "on_device": True,
                                   "step_size_final": float(res.step_size_final)})

    def F_fn(u):
        return ensemble_value(obj, u.reshape(-1, order="F"))

    def G_fn(u):
        return ensemble_value_and_grad(obj, u.reshape(-1, order="F"))[1].reshape(shape, order="F")

    try:
        best, best_F, f_hist, g_hist, conv, reason, n_f, n_g = _ensemble_lbfgs(
            F_fn, G_fn, controls_init.controls, config, lbfgs_memory, amplitude_limit)
    finally:
        obj.batched_eval.close()
    return OptimizationResult(best, best_F, f_hist, g_hist, len(f_hist), conv, reason, time.time() - t0, n_f, n_g,
                              {"algorithm": "Ensemble L-BFGS-GRAPE", "n_systems": len(systems),
                               "lbfgs_memory": lbfgs_memory})
