"""Gradient verification and the AD entry points, executed on the :rocm device.

  finite_difference_gradient, verify_gradient     src/Physics/Gradients.jl:442-554
  grape_gradient_check                            src/Optimization/GRAPE.jl:911-951
  AutoDiffConfig, select_autodiff_backend,
  compute_gradient_autodiff, finite_diff_gradient_ad,
  verify_gradient_autodiff                        src/Physics/AutoDiff.jl:187-565

Central differences need 2 n_c n_t fidelity evaluations at perturbed controls;
they are independent, so they run as problems of one device plan, F only, in
chunks sized to a fixed amount of device memory.

The reference's AD backends differentiate through its eigen-exp with
ForwardDiff or Zygote. On this device the same derivative (exact in dt) is the
Frechet-derivative gradient mode of the library (PR_GRAD_EXACT), which is
what the :forward and :reverse backends return here; :finite_diff takes the
central-difference path above.
"""
from __future__ import annotations

import warnings
from dataclasses import dataclass

import numpy as np

from . import native
from .api import ControlSequence, QuantumTarget, compute_grape_gradient, fidelity_problem
from .problem import PR_GRAD_EXACT

# device memory one chunk of perturbed problems may take for its U images
_CHUNK_BYTES = 1 << 30


def _fidelities(system, target: QuantumTarget, dt: float, controls_list):
    """compute_fidelity of one system at many [n_c x n_t] control matrices."""
    n = len(controls_list)
    n_t = controls_list[0].shape[1]
    p = fidelity_problem(system, ControlSequence(controls_list[0].T, dt, n_t), target)
    nb = (p.dim + 15) // 16
    per = max(1, n_t * nb * nb * 256 * 16)
    chunk = int(max(1, min(n, _CHUNK_BYTES // per)))
    out = np.empty(n)
    plan = native.Plan([p] * chunk)
    try:
        for i0 in range(0, n, chunk):
            part = list(controls_list[i0:i0 + chunk])
            m = len(part)
            part += [part[-1]] * (chunk - m)
            F, _ = plan.eval(part, grad=False)
            out[i0:i0 + m] = F[:m]
    finally:
        plan.close()
    return out


def finite_difference_gradient(system, controls: ControlSequence, target: QuantumTarget,
                               eps: float = 1e-6) -> np.ndarray:
    """Central differences (F(u + eps e_jk) - F(u - eps e_jk)) / (2 eps), [n_c x n_t]
    (Gradients.jl:442-478)."""
    if eps <= 0:
        raise ValueError(f"eps must be positive, got {eps}")
    u = np.array(controls.controls, dtype=np.float64)
    n_c, n_t = u.shape
    pert = []
    for j in range(n_c):
        for k in range(n_t):
            up = u.copy()
            up[j, k] += eps
            um = u.copy()
            um[j, k] -= eps
            pert += [up, um]
    F = _fidelities(system, target, controls.dt, pert)
    return ((F[0::2] - F[1::2]) / (2 * eps)).reshape(n_c, n_t)


def verify_gradient(analytical_grad, numerical_grad, tol: float = 1e-5, relative: bool = True) -> bool:
    """||G_ana - G_num||_inf (divided by max(1, ||G_num||_inf) when relative) <= tol
    (Gradients.jl:536-554); warns and returns False otherwise."""
    a = np.asarray(analytical_grad, dtype=np.float64)
    b = np.asarray(numerical_grad, dtype=np.float64)
    if a.shape != b.shape:
        raise ValueError(f"analytical_grad size {a.shape} != numerical_grad size {b.shape}")
    abs_err = float(np.max(np.abs(a - b))) if a.size else 0.0
    err = abs_err / max(1.0, float(np.max(np.abs(b))) if b.size else 0.0) if relative else abs_err
    if err > tol:
        warnings.warn(f"Gradient verification failed: error = {err} > tol = {tol} "
                      f"(max |G_ana - G_num| = {abs_err})")
        return False
    return True


def grape_gradient_check(system, target: QuantumTarget, controls: ControlSequence, eps: float = 1e-5) -> dict:
    """grape_gradient_check (GRAPE.jl:911-951): the GRAPE gradient against central
    differences; passed when the largest absolute deviation is below 1e-4."""
    G_ana = compute_grape_gradient(system, controls, target)
    G_num = finite_difference_gradient(system, controls, target, eps=eps)
    diff = np.abs(G_ana - G_num)
    max_err = float(diff.max())
    scale = max(1.0, float(np.abs(G_num).max()))
    return {"analytical_gradient": G_ana, "numerical_gradient": G_num, "max_error": max_err,
            "max_relative_error": max_err / scale, "passed": max_err < 1e-4}


@dataclass
class AutoDiffConfig:
    """AutoDiffConfig (AutoDiff.jl): backend :auto, :forward, :reverse or :finite_diff."""

    backend: str = "auto"
    numerical_eps: float = 1e-6
    verify_against_numerical: bool = False
    verbose: bool = False

    def __post_init__(self):
        self.backend = self.backend.lstrip(":")
        if self.backend not in ("auto", "forward", "reverse", "finite_diff"):
            raise ValueError(f"backend must be :auto, :forward, :reverse or :finite_diff; got {self.backend}")


def select_autodiff_backend(system, controls: ControlSequence, config: AutoDiffConfig) -> str:
    """An explicit backend is returned as given; :auto takes :reverse, the
    device's exact adjoint gradient (AutoDiff.jl:473-492 picks the available
    AD package the same way)."""
    return config.backend if config.backend != "auto" else "reverse"


def finite_diff_gradient_ad(system, controls: ControlSequence, target: QuantumTarget,
                            eps: float = 1e-6) -> np.ndarray:
    """finite_diff_gradient_ad (AutoDiff.jl:402-435): the same central differences."""
    return finite_difference_gradient(system, controls, target, eps=eps)


def compute_gradient_autodiff(system, controls: ControlSequence, target: QuantumTarget,
                              config: AutoDiffConfig | None = None) -> np.ndarray:
    """compute_gradient_autodiff (AutoDiff.jl:187-220)."""
    config = config or AutoDiffConfig()
    backend = select_autodiff_backend(system, controls, config)
    if config.verbose:
        print(f"[AutoDiff] Using backend :{backend} for {system.n_controls}x{controls.n_timesteps} problem")
    if backend == "finite_diff":
        G = finite_diff_gradient_ad(system, controls, target, eps=config.numerical_eps)
    else:
        G = compute_grape_gradient(system, controls, target, grad_mode=PR_GRAD_EXACT)
    if config.verify_against_numerical:
        G_fd = finite_diff_gradient_ad(system, controls, target, eps=config.numerical_eps)
        dev = float(np.max(np.abs(G - G_fd)))
        if dev > 1e-4:
            warnings.warn(f"[AutoDiff] Gradient verification FAILED: max|G_ad - G_fd| = {dev}")
        elif config.verbose:
            print(f"[AutoDiff] Gradient verified: max|G_ad - G_fd| = {dev}")
    return G


def verify_gradient_autodiff(system, controls: ControlSequence, target: QuantumTarget, tol: float = 1e-5,
                             verbose: bool = True) -> bool:
    """verify_gradient_autodiff (AutoDiff.jl:535-565): the GRAPE gradient against
    the AD gradient; True when max |G_grape - G_ad| < tol."""
    G_grape = compute_grape_gradient(system, controls, target)
    G_ad = compute_gradient_autodiff(system, controls, target, AutoDiffConfig(backend="auto"))
    diff = np.abs(G_grape - G_ad)
    dev = float(diff.max())
    if verbose:
        j, k = np.unravel_index(int(np.argmax(diff)), diff.shape)
        print(f"[GradVerify] max|G_grape - G_ad| = {dev:.3e} at (j={j + 1}, k={k + 1})")
        print("[GradVerify] " + ("PASSED" if dev < tol else "FAILED") + f" (tol = {tol:.1e})")
    return dev < tol
