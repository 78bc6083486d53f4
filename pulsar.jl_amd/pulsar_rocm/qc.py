"""The quantum-computing callers of the path: quasi-static robust control and
the open-system LindbladQCControl context.

quasi_static_ensemble / robust_optimcon_qs
(src/Application/QuantumComputing/NoiseModels/QuasiStatic.jl:23-216) build
an ensemble of drift Hamiltonians over the Cartesian product of the noise
samples and run grape_optimize_ensemble over it: here that ensemble is one
device plan (one problem, the members as its ensemble) and the L-BFGS loop
runs in the library (pr_lbfgs_optimize).

LindbladQCControl (src/Application/QuantumComputing/LindbladQCControl.jl:
33-284) precomputes the drift and control Liouvillians and vec(rho) of the
initial and target states; its kernel is grape_lindblad_kernel seen through
an adapter (one drift, power level 1, :square fidelity, uniform dt), which
here is one PR_LINDBLAD problem with one member and one state pair
(qc_kernel). optimcon(ctx) runs its :grape loop (gradient ascent with the
context's fixed step, LindbladQCControl.jl:356-406) over that kernel; the
reference's L-BFGS-B, CG and derivative-free optimizers are host libraries
around the same kernel and are not part of this library.
"""
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np

from . import native
from .api import (ControlSequence, GRAPEConfig, OptimizationResult, QuantumSystem, QuantumTarget,
                  grape_optimize_ensemble, quantum_system)
from .mr import build_control_liouvillian, build_drift_liouvillian, pure_state_to_vec_rho
from .problem import PR_LINDBLAD, PR_PAIRS_SQUARE, Problem


# ------------------------------------------------------------------ quasi-static noise
@dataclass
class QuasiStaticNoise:
    """A quasi-static noise source (QuasiStatic.jl:23-71): the noisy
    parameter, the distribution (:gaussian, :uniform or :custom), sigma
    (standard deviation; half-width for :uniform), the sample count and
    optional explicit samples (which override the distribution)."""

    parameter: str
    sigma: float
    distribution: str = "gaussian"
    n_samples: int = 9
    samples: List[float] = field(default_factory=list)

    def __post_init__(self):
        self.distribution = str(self.distribution).lstrip(":")
        if self.distribution not in ("gaussian", "uniform", "custom") and not self.samples:
            raise ValueError("distribution must be :gaussian, :uniform or :custom")
        if self.n_samples <= 0:
            raise ValueError("n_samples must be positive")


def qs_samples(noise: QuasiStaticNoise, rng: Optional[np.random.Generator] = None) -> np.ndarray:
    """_qs_samples (QuasiStatic.jl:76-98): explicit samples if given; :gaussian
    n evenly spaced z-scores on [-2, 2] times sigma (0 for n = 1); :uniform n
    evenly spaced points on [-sigma, sigma]; :custom sigma times n normal
    draws (from numpy's generator: Julia's stream is not reproduced)."""
    if len(noise.samples):
        return np.asarray(noise.samples, np.float64)
    n = noise.n_samples
    if noise.distribution == "gaussian":
        if n == 1:
            return np.zeros(1)
        return noise.sigma * np.linspace(-2.0, 2.0, n)
    if noise.distribution == "uniform":
        return np.linspace(-noise.sigma, noise.sigma, n)
    rng = rng or np.random.default_rng()
    return noise.sigma * rng.standard_normal(n)


def quasi_static_ensemble(H_drift_fn: Callable, noise_sources, rng=None) -> List[np.ndarray]:
    """quasi_static_ensemble (QuasiStatic.jl:132-155): H_drift_fn(params) over
    the Cartesian product of the sources' samples, the first source varying
    slowest (params[k] is the value drawn for noise_sources[k])."""
    grids = [qs_samples(ns, rng) for ns in noise_sources]
    drifts = []

    def walk(k, params):
        if k == len(grids):
            drifts.append(np.asarray(H_drift_fn(list(params)), np.complex128))
            return
        for s in grids[k]:
            walk(k + 1, params + [float(s)])

    walk(0, [])
    return drifts


def robust_optimcon_qs(H_drift_fn: Callable, H_controls, noise_sources, target: QuantumTarget,
                       ctrl: ControlSequence, config: Optional[GRAPEConfig] = None, pwr_levels=None,
                       weights=None, rng=None, lbfgs_memory: int = 10) -> OptimizationResult:
    """robust_optimcon_qs (QuasiStatic.jl:196-216): the quasi-static ensemble
    of drifts, the controls scaled by pwr_levels (_scale_controls,
    Platforms/Common.jl:64-69), and grape_optimize_ensemble over the systems,
    which evaluates the whole ensemble as one device plan with the L-BFGS
    loop in the library. `weights` is accepted and unused, as there (the
    ensemble mean weights every member 1/M)."""
    del weights
    drifts = quasi_static_ensemble(H_drift_fn, noise_sources, rng)
    H_controls = [np.asarray(h, np.complex128) for h in H_controls]
    pwr = np.ones(len(H_controls)) if pwr_levels is None else np.asarray(pwr_levels, np.float64)
    if len(pwr) != len(H_controls):
        raise ValueError("scale must have the same length as H_controls")
    scaled = [pwr[k] * H_controls[k] for k in range(len(H_controls))]
    systems = [quantum_system(H, scaled) for H in drifts]
    return grape_optimize_ensemble(systems, target, ctrl, config=config, lbfgs_memory=lbfgs_memory)


# ------------------------------------------------------------------ LindbladQCControl
def qc_vectorise(psi_or_rho, N: int) -> np.ndarray:
    """_qc_vectorise (LindbladQCControl.jl:199-226): a pure state of length N
    becomes vec(|psi><psi|); a vec(rho) of length N^2 is kept."""
    v = np.asarray(psi_or_rho, np.complex128).reshape(-1)
    if v.size == N:
        return pure_state_to_vec_rho(v)
    if v.size == N * N:
        return v.copy()
    raise ValueError(f"LindbladQCControl: a state of length {v.size} is neither a pure state of dimension {N} "
                     f"nor a vec(rho) of length {N * N}")


class LindbladQCControl:
    """The open-system QC context (LindbladQCControl.jl:69-191): system,
    state target, initial ControlSequence, jump operators and rates, and the
    precomputed Liouville-space objects (_L_drift, _L_controls, _sigma_init,
    _sigma_targ). rho_init defaults to |0...0>. Penalties are host callables
    subtracted from F and G after the device kernel, as the adapter's
    _apply_penalties! does."""

    def __init__(self, system: QuantumSystem, target: QuantumTarget, ctrl: ControlSequence, jump_ops=(),
                 decay_rates=(), rho_init=None, method: str = "lbfgs", max_iter: int = 500,
                 convergence_tol: float = 1e-8, gradient_norm_tol: float = 1e-6, step_size: float = 0.01,
                 penalty_fns=(), penalty_grad_fns=(), metadata=None):
        if len(jump_ops) != len(decay_rates):
            raise ValueError(f"jump_ops and decay_rates differ in length ({len(jump_ops)} and {len(decay_rates)})")
        if any(r < 0 for r in decay_rates):
            raise ValueError("decay_rates must be non-negative")
        if target.type != "state":
            raise ValueError(f"LindbladQCControl takes state-transfer targets only (target.type = "
                             f"'{target.type}'); use LindbladMRControl for gates in Liouville space")
        self.system, self.target, self.ctrl = system, target, ctrl
        self.jump_ops = [np.asarray(L, np.complex128) for L in jump_ops]
        self.decay_rates = [float(r) for r in decay_rates]
        self.method = str(method).lstrip(":")
        self.grape_config = GRAPEConfig(max_iter=max_iter, convergence_tol=convergence_tol,
                                        gradient_norm_tol=gradient_norm_tol, step_size=step_size)
        self.penalty_fns, self.penalty_grad_fns = list(penalty_fns), list(penalty_grad_fns)
        self.metadata = dict(metadata or {})
        H0 = np.asarray(system.H_drift, np.complex128)
        N = H0.shape[0]
        self._hilbert_dim, self._liouville_dim = N, N * N
        self._L_drift = build_drift_liouvillian(H0, self.jump_ops, self.decay_rates)
        self._L_controls = [build_control_liouvillian(h) for h in system.H_controls]
        if rho_init is None:
            psi0 = np.zeros(N, np.complex128)
            psi0[0] = 1.0
            self._sigma_init = pure_state_to_vec_rho(psi0)
        else:
            self._sigma_init = qc_vectorise(rho_init, N)
        self._sigma_targ = qc_vectorise(target.target_state, N)
        self._plan = None

    def problem(self) -> Problem:
        """The adapter's view for grape_lindblad_kernel (LindbladQCControl.jl:
        237-263) as one PR_LINDBLAD problem: one member (drift Liouvillian,
        power level 1), one pair (sigma_init -> sigma_targ), :square, dt of
        the ControlSequence on every slice."""
        n_t = self.ctrl.controls.shape[1]
        return Problem(dim=self._liouville_dim, n_ctrl=len(self._L_controls), n_t=n_t, H0=self._L_drift,
                       Hc=np.stack(self._L_controls), dt=float(self.ctrl.dt), dynamics=PR_LINDBLAD,
                       objective=PR_PAIRS_SQUARE, n_members=1, psi_init=self._sigma_init[None, :],
                       psi_targ=self._sigma_targ[None, :], weight=np.ones(1))

    def plan(self) -> native.Plan:
        if self._plan is None:
            self._plan = native.Plan([self.problem()])
        return self._plan

    def close(self):
        if self._plan is not None:
            self._plan.close()
            self._plan = None


def qc_kernel(w, ctx: LindbladQCControl):
    """_qc_kernel (LindbladQCControl.jl:276-284): (F, G[n_ctrl, n_t]) of the
    Liouville-space kernel on the device, penalties subtracted."""
    F, G = ctx.plan().eval([np.asarray(w, np.float64)])
    F, G = float(F[0]), G[0]
    for pf, pg in zip(ctx.penalty_fns, ctx.penalty_grad_fns):
        F -= pf(w)
        G = G - pg(w)
    return F, G


def optimcon(ctx: LindbladQCControl) -> OptimizationResult:
    """optimcon(ctx::LindbladQCControl) for :grape (LindbladQCControl.jl:
    356-406): per iteration F, G = qc_kernel(w), stop when |G| <
    gradient_norm_tol or |F - F_prev| < convergence_tol (from the second
    iteration), else w += step_size G; the best w and F are returned."""
    import time
    if ctx.method != "grape":
        raise ValueError(f"method :{ctx.method} runs a host optimizer around qc_kernel (L-BFGS-B, CG or a "
                         "derivative-free method), which is not part of this library; use :grape or call "
                         "qc_kernel from that optimizer")
    cfg = ctx.grape_config
    t0 = time.perf_counter()
    w = np.array(ctx.ctrl.controls, np.float64)
    f_hist, g_hist = [], []
    F_prev, F_best, w_best = -np.inf, -np.inf, w.copy()
    n = 0
    converged, reason = False, "maximum iterations reached"
    try:
        for it in range(1, cfg.max_iter + 1):
            F, G = qc_kernel(w, ctx)
            n += 1
            gn = float(np.linalg.norm(G))
            f_hist.append(F)
            g_hist.append(gn)
            if F > F_best:
                F_best, w_best = F, w.copy()
            if gn < cfg.gradient_norm_tol:
                converged, reason = True, "gradient norm < tol"
                break
            if abs(F - F_prev) < cfg.convergence_tol and it > 1:
                converged, reason = True, "fidelity change < tol"
                break
            F_prev = F
            w = w + cfg.step_size * G
    finally:
        ctx.close()
    return OptimizationResult(w_best, F_best, f_hist, g_hist, len(f_hist), converged, reason,
                              time.perf_counter() - t0, n, n,
                              {"algorithm": "Lindblad QC GRAPE (gradient ascent)", "hilbert_dim": ctx._hilbert_dim,
                               "liouville_dim": ctx._liouville_dim, "n_jump_ops": len(ctx.jump_ops)})
