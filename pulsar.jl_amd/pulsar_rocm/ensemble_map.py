"""ensemble_map / ensemble_grad! / ensemble_fobj with a :rocm route.

Mirror of src/Computation/EnsembleMap.jl:324-607. The reference maps arbitrary
per-member closures over CPU threads, also under :cuda and :metal
(EnsembleMap.jl:375-415: "GPU acceleration for arbitrary closures is not
supported"), and leaves batched device kernels to extensions that overload the
map (`_ensemble_map_gpu!`, :425-436).

Here the device is :rocm and the batched kernel is the library's plan:

  * a member function that carries a device description of its members (the
    `MemberObjectives` functors below, or any object with the same `rocm_*`
    methods) is evaluated for ALL requested members in one `pr_eval`:
    ensemble_map -> one plan with one problem per member (per-member F, no
    gradient walk); ensemble_grad! -> one plan whose member weights are
    w_i / sum(w), so the library reduces the weighted mean gradient on the device;
  * any other callable is mapped on host threads, as the reference does for
    closures it cannot send to a GPU (an informational message once).

Member objectives are those of build_ensemble_from_systems' sample closures
(SystemBuilder.jl:61-77): F_i = compute_fidelity(systems[i], ctrl(theta), target)
(Objectives.jl:632-639, states as given) and g_i = compute_grape_gradient(
systems[i], ctrl(theta), target) (Gradients.jl:80-127, normalised states),
with ctrl(theta) = _vec_to_ctrl(theta): theta is vec(controls) = [n_c x n_t]
column-major (ControlSequence.jl:99-110).

Signs follow the reference: ensemble_fobj's f returns -mean(F_i) (or
-dot(w, F)/sum(w)) for minimisers, and its grad! the weighted mean of the
member gradients as grad_member! gives them (EnsembleMap.jl:580-606).
"""
from __future__ import annotations

import logging
import threading
from concurrent.futures import ThreadPoolExecutor
from typing import Dict, Optional, Sequence, Tuple

import numpy as np

from . import api, native
from .problem import PR_FP64, PR_GRAD_REFERENCE

_log = logging.getLogger("pulsar_rocm")
_DEVICES = ("cpu", "cuda", "metal", "rocm")
_info_once = set()


def _info(key: str, msg: str) -> None:
    if key not in _info_once:
        _info_once.add(key)
        _log.info(msg)


def _n_threads() -> int:
    import os

    return max(1, min(16, os.cpu_count() or 1))


def _host_map(f, indices) -> np.ndarray:
    """f(idx) for every idx on host threads, in index order (@threadsif, EnsembleMap.jl:368-371)."""
    out = np.empty(len(indices))
    if len(indices) <= 1:
        for k, i in enumerate(indices):
            out[k] = float(f(i))
        return out
    with ThreadPoolExecutor(max_workers=_n_threads()) as ex:
        for k, v in enumerate(ex.map(f, indices)):
            out[k] = float(v)
    return out


def _device(device: Optional[str]) -> str:
    d = (device or api.get_device()).lstrip(":")
    if d not in _DEVICES:
        raise ValueError(f"ensemble_map: unknown device :{d}. Choose :cpu, :cuda, :metal, or :rocm.")
    return d


def ensemble_map(f, iter, device: Optional[str] = None) -> np.ndarray:
    """ensemble_map(f, iter; device=get_device()) (EnsembleMap.jl:365-423): [f(idx) for idx
    in iter] as float64, in index order. Under :rocm a member function with a device
    description (`rocm_map(indices)`) runs as one device evaluation."""
    indices = list(iter)
    d = _device(device)
    if d == "rocm":
        batch = getattr(f, "rocm_map", None)
        if batch is not None:
            return np.asarray(batch(indices), dtype=np.float64)
        _info("map_closure", "ensemble_map: objective function uses CPU threads even with :rocm device. "
                             "Device batching needs a member function with a device description "
                             "(MemberObjectives); arbitrary closures are not supported.")
    elif d in ("cuda", "metal"):
        _info(f"map_{d}", f"ensemble_map: :{d} is not loaded here — using CPU threads.")
    return _host_map(f, indices)


def _weights(weight, n: int, fn: str):
    if weight is None:
        return None
    w = np.asarray(weight, dtype=np.float64).ravel()
    if w.size != n:
        raise ValueError(f"{fn}: weight length ({w.size}) ≠ number of ensemble members ({n}).")
    return w


def ensemble_grad(g_out: np.ndarray, theta, grad_member, iter, weight=None,
                  device: Optional[str] = None) -> np.ndarray:
    """ensemble_grad!(g_out, θ, grad_member!, iter; weight, device) (EnsembleMap.jl:496-533):
    g_out <- sum_i w_i g_i / sum_i w_i (uniform weights when weight is None), in place.
    Under :rocm a grad_member with a device description (`rocm_grad(theta, indices,
    weights)`) reduces the weighted mean on the device in one evaluation."""
    indices = list(iter)
    n = len(indices)
    w = _weights(weight, n, "ensemble_grad!")
    d = _device(device)
    theta = np.asarray(theta, dtype=np.float64)
    if d == "rocm":
        batch = getattr(grad_member, "rocm_grad", None)
        if batch is not None:
            wv = np.ones(n) if w is None else w
            wsum = float(wv.sum())
            scale = wv / wsum if wsum > 0 else np.full(n, 1.0 / n)
            g_out[...] = np.asarray(batch(theta, indices, scale)).reshape(g_out.shape)
            return g_out
        _info("grad_closure", "ensemble_grad!: gradient function uses CPU threads even with :rocm device.")
    p = g_out.size
    acc = np.zeros(p)
    wsum = [0.0]
    lk = threading.Lock()

    def one(k):
        buf = np.zeros(p)
        grad_member(buf, theta, indices[k])
        wk = 1.0 if w is None else w[k]
        with lk:
            acc[:] += wk * buf
            wsum[0] += wk

    if n <= 1:
        for k in range(n):
            one(k)
    else:
        with ThreadPoolExecutor(max_workers=_n_threads()) as ex:
            list(ex.map(one, range(n)))
    scale = 1.0 / wsum[0] if wsum[0] > 0 else 1.0 / n
    g_out[...] = (acc * scale).reshape(g_out.shape)
    return g_out


def ensemble_fobj(f_member, grad_member, iter, weight=None, device: Optional[str] = None):
    """ensemble_fobj(f_member, grad_member!, iter; weight, device) (EnsembleMap.jl:580-607):
    the pair (f, grad!) with f(θ) = -mean_i f_member(θ, i) (-dot(w, F)/sum(w) when
    weighted) and grad!(g, θ) = ensemble_grad!(g, θ, grad_member!, iter). Under :rocm
    f_member with a device description (`rocm_values(theta, indices)`) is one device
    evaluation per call."""
    indices = list(iter)
    w = _weights(weight, len(indices), "ensemble_fobj")

    def f(theta):
        batch = getattr(f_member, "rocm_values", None)
        if _device(device) == "rocm" and batch is not None:
            parts = np.asarray(batch(np.asarray(theta, dtype=np.float64), indices), dtype=np.float64)
        else:
            parts = ensemble_map(lambda i: float(f_member(theta, i)), indices, device=device)
        if w is None:
            return -float(np.mean(parts))
        return -float(np.dot(w, parts) / w.sum())

    def grad(g, theta):
        return ensemble_grad(g, theta, grad_member, indices, weight=w, device=device)

    return f, grad


# --------------------------------------------------------------------------- device descriptions


class _PlanCache:
    """Plans keyed by (kind, member indices, weights); a few kept, oldest closed first."""

    def __init__(self, max_plans: int = 4):
        self.max = max_plans
        self.plans: Dict[Tuple, Tuple[native.Plan, int]] = {}
        self.lock = threading.Lock()

    def get(self, key, build):
        with self.lock:
            hit = self.plans.get(key)
            if hit is not None:
                return hit
            if len(self.plans) >= self.max:
                old = next(iter(self.plans))
                self.plans.pop(old)[0].close()
            hit = build()
            self.plans[key] = hit
            return hit

    def close(self):
        with self.lock:
            for plan, _ in self.plans.values():
                plan.close()
            self.plans.clear()


class MemberObjectives:
    """Device description of the per-member objectives of an ensemble of systems
    sharing a target and a time grid (the sample closures of
    build_ensemble_from_systems, SystemBuilder.jl:61-77).

      mo.fidelity_at(theta)  -> i -> F_i(theta)        for ensemble_map
      mo.f_member            -> (theta, i) -> F_i       for ensemble_fobj
      mo.grad_member         -> (g, theta, i) -> g_i    for ensemble_grad! / ensemble_fobj

    Each is an ordinary callable (one member per call, one device evaluation) that
    also carries the batched form the :rocm route of ensemble_map / ensemble_grad! /
    ensemble_fobj uses. Member indices are positions in `systems` (0-based)."""

    def __init__(self, systems: Sequence[api.QuantumSystem], target: api.QuantumTarget, n_t: int, dt: float,
                 grad_mode: int = PR_GRAD_REFERENCE, precision: int = PR_FP64):
        if len(systems) == 0:
            raise ValueError("systems must be non-empty")
        n_c = systems[0].n_controls
        for s in systems:
            if s.n_controls != n_c:
                raise ValueError("every member must have the same number of control fields")
        api._check_target(target)
        self.systems = list(systems)
        self.target = target
        self.n_t, self.dt, self.n_c = int(n_t), float(dt), n_c
        self.grad_mode, self.precision = int(grad_mode), int(precision)
        self._plans = _PlanCache()
        mo = self

        class _FAt:
            def __init__(self, theta):
                self.theta = np.asarray(theta, dtype=np.float64)

            def __call__(self, i):
                return float(mo.rocm_values(self.theta, [i])[0])

            def rocm_map(self, indices):
                return mo.rocm_values(self.theta, indices)

        class _FMember:
            def __call__(self, theta, i):
                return float(mo.rocm_values(np.asarray(theta, dtype=np.float64), [i])[0])

            def rocm_values(self, theta, indices):
                return mo.rocm_values(theta, indices)

        class _GMember:
            def __call__(self, g, theta, i):
                g[...] = mo.rocm_grad(np.asarray(theta, dtype=np.float64), [i], np.ones(1)).reshape(g.shape)
                return g

            def rocm_grad(self, theta, indices, weights):
                return mo.rocm_grad(theta, indices, weights)

        self._FAt = _FAt
        self.f_member = _FMember()
        self.grad_member = _GMember()

    def fidelity_at(self, theta):
        """i -> F_i at controls theta (the closure of the ensemble_map example, EnsembleMap.jl:354-363)."""
        return self._FAt(theta)

    def _check_indices(self, indices):
        for i in indices:
            if not (isinstance(i, (int, np.integer)) and 0 <= int(i) < len(self.systems)):
                raise IndexError(f"member index {i!r} is not in 0..{len(self.systems) - 1}")

    def _controls(self, theta):
        theta = np.asarray(theta, dtype=np.float64)
        if theta.size != self.n_c * self.n_t:
            raise ValueError(f"theta has {theta.size} entries; expected n_controls*n_timesteps = "
                             f"{self.n_c * self.n_t}")
        return theta.reshape(self.n_c, self.n_t, order="F")

    def _ctrl_seq(self):
        return api.ControlSequence(np.zeros((self.n_t, self.n_c)), self.dt, self.n_t)

    def rocm_values(self, theta, indices) -> np.ndarray:
        """F_i for every i in indices: one plan with one problem per member, F only."""
        self._check_indices(indices)
        u = self._controls(theta)
        key = ("F", tuple(int(i) for i in indices))

        def build():
            cs = self._ctrl_seq()
            probs = [api.fidelity_problem(self.systems[int(i)], cs, self.target) for i in indices]
            return native.Plan(probs, precision=self.precision), len(probs)

        plan, n_probs = self._plans.get(key, build)
        F, _ = plan.eval([u] * n_probs, grad=False)
        return np.asarray(F, dtype=np.float64)

    def rocm_grad(self, theta, indices, weights) -> np.ndarray:
        """sum_i weights_i g_i (vec form, n_c*n_t) over indices: one plan whose members
        carry the weights (members sharing control operators up to per-channel scale
        factors form one problem), reduced on the device."""
        self._check_indices(indices)
        weights = np.asarray(weights, dtype=np.float64).ravel()
        if weights.size != len(indices):
            raise ValueError("one weight per member index is needed")
        u = self._controls(theta)
        key = ("G", tuple(int(i) for i in indices), weights.tobytes())

        def build():
            probs = _weighted_gradient_problems([self.systems[int(i)] for i in indices], weights, self.target,
                                                self.n_t, self.dt)
            return native.Plan(probs, grad_mode=self.grad_mode, precision=self.precision), len(probs)

        plan, n_probs = self._plans.get(key, build)
        _, G = plan.eval([u] * n_probs, grad=True)
        return np.sum(G, axis=0).reshape(-1, order="F")

    def close(self):
        self._plans.close()


def _weighted_gradient_problems(systems, weights, target, n_t, dt):
    """compute_grape_gradient's objective (normalised states, psi_i = initial_state or
    target_state) for every system, grouped into problems of members that share the
    control operators up to per-channel scales, member weights as given."""
    cs = api.ControlSequence(np.zeros((n_t, systems[0].n_controls)), dt, n_t)
    groups = []  # (ref controls, [(position, scales)])
    for pos, s in enumerate(systems):
        for ref, members in groups:
            sc = api._scale_factors(ref, s.H_controls)
            if sc is not None:
                members.append((pos, sc))
                break
        else:
            groups.append((s.H_controls, [(pos, np.ones(len(s.H_controls)))]))
    probs = []
    for ref, members in groups:
        base = api.gradient_problem(systems[members[0][0]], cs, target)
        base.n_members = len(members)
        base.H0 = np.stack([systems[p].H_drift for p, _ in members])
        base.Hc = np.stack(ref)
        base.ctrl_scale = np.stack([sc for _, sc in members])
        base.weight = np.array([weights[p] for p, _ in members])
        probs.append(base)
    return probs
