"""ctypes binding of libpulsar_rocm.so (the C-ABI in include/pulsar_rocm.h).

This is the device path. There is no CPU fallback here: when the library or
the GPU is missing, every call raises `RocmError`; falling back to another
device is the caller's decision (the host API does it only on request).
"""
from __future__ import annotations

import ctypes
import os
from typing import Sequence

import numpy as np

from .problem import COptions, Problem, controls_cm, grad_from_cm, to_c

LIB_NAME = "libpulsar_rocm.so"
LIB_PATH = os.environ.get("PULSAR_ROCM_LIB") or os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

PR_ERR = {
    -1: "PR_ERR_ARG",
    -2: "PR_ERR_DEVICE",
    -3: "PR_ERR_NOMEM",
    -4: "PR_ERR_LAUNCH",
    -5: "PR_ERR_UNSUPPORTED",
    -6: "PR_ERR_INJECTED",
}

EXPORTS = ("pr_abi_version", "pr_device_count", "pr_init", "pr_last_error", "pr_plan_create",
           "pr_eval", "pr_eval_device", "pr_eval_device_static", "pr_plan_n_params", "pr_plan_exp_counts", "pr_plan_shapes", "pr_plan_set_timing",
           "pr_plan_kernel_times", "pr_plan_buffers", "pr_plan_upload_controls",
           "pr_plan_synchronize", "pr_plan_destroy", "pr_grape_optimize", "pr_lbfgs_optimize", "pr_mr_optimcon")
KERNEL_CLASSES = ("segment", "combine", "gradient", "reduce")


class GrapeConfig(ctypes.Structure):
    """pr_grape_config (GRAPEConfig's loop knobs, GRAPE.jl:60-127)."""

    _fields_ = [("max_iter", ctypes.c_int), ("convergence_tol", ctypes.c_double), ("step_size", ctypes.c_double),
                ("adapt_step_size", ctypes.c_int), ("min_step_size", ctypes.c_double),
                ("max_step_size", ctypes.c_double), ("gradient_norm_tol", ctypes.c_double)]


class GrapeResult(ctypes.Structure):
    _fields_ = [("n_iterations", ctypes.c_int), ("converged", ctypes.c_int), ("reason", ctypes.c_int),
                ("n_fidelity_history", ctypes.c_int), ("n_fidelity_evaluations", ctypes.c_int),
                ("n_gradient_evaluations", ctypes.c_int), ("fidelity", ctypes.c_double),
                ("step_size_final", ctypes.c_double)]


class LbfgsConfig(ctypes.Structure):
    """pr_lbfgs_config (the ensemble L-BFGS loop's knobs, GRAPE.jl:1012-1230)."""

    _fields_ = [("max_iter", ctypes.c_int), ("convergence_tol", ctypes.c_double), ("step_size", ctypes.c_double),
                ("gradient_norm_tol", ctypes.c_double), ("lbfgs_memory", ctypes.c_int),
                ("amplitude_limit", ctypes.c_double)]


class MrConfig(ctypes.Structure):
    """pr_mr_config (the MR optimcon loops' knobs, OptControl.jl:380-561)."""

    _fields_ = [("method", ctypes.c_int), ("max_iter", ctypes.c_int), ("grad_tol", ctypes.c_double),
                ("lbfgs_memory", ctypes.c_int), ("l_bound", ctypes.c_double), ("u_bound", ctypes.c_double),
                ("tile", ctypes.c_int)]


PR_MR_LBFGS, PR_MR_GRAPE = 0, 1
PR_STOP = {0: "max_iter", 1: "perfect", 2: "gradient", 3: "fidelity", 4: "zero_direction"}
PR_STOP_CODES = {v: k for k, v in PR_STOP.items()}


class RocmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{PR_ERR.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load(path: str = LIB_PATH):
    """Load the shared library and declare every exported entry point."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RocmError(-2, f"{path} is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(path)
    dp = ctypes.POINTER(ctypes.c_double)
    L.pr_abi_version.restype = ctypes.c_int
    L.pr_device_count.restype = ctypes.c_int
    L.pr_init.argtypes = [ctypes.c_int]
    L.pr_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    L.pr_plan_create.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(COptions),
                                 ctypes.POINTER(ctypes.c_void_p)]
    L.pr_eval.argtypes = [ctypes.c_void_p, ctypes.POINTER(dp), dp, ctypes.POINTER(dp), dp]
    L.pr_eval_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_void_p]
    L.pr_plan_n_params.argtypes = [ctypes.c_void_p]
    L.pr_plan_n_params.restype = ctypes.c_longlong
    L.pr_eval_device_static.argtypes = [ctypes.c_void_p] * 5
    L.pr_plan_exp_counts.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong)]
    L.pr_plan_shapes.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    L.pr_plan_set_timing.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.pr_plan_kernel_times.argtypes = [ctypes.c_void_p, dp, ctypes.POINTER(ctypes.c_longlong)]
    L.pr_plan_buffers.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_void_p)] * 3
    L.pr_plan_upload_controls.argtypes = [ctypes.c_void_p, ctypes.POINTER(dp)]
    L.pr_plan_synchronize.argtypes = [ctypes.c_void_p]
    L.pr_plan_destroy.argtypes = [ctypes.c_void_p]
    L.pr_grape_optimize.argtypes = [ctypes.c_void_p, ctypes.c_void_p, dp, ctypes.POINTER(GrapeConfig), dp, dp, dp,
                                    ctypes.POINTER(GrapeResult)]
    L.pr_lbfgs_optimize.argtypes = [ctypes.c_void_p, dp, ctypes.POINTER(LbfgsConfig), dp, dp, dp,
                                    ctypes.POINTER(GrapeResult)]
    L.pr_mr_optimcon.argtypes = [ctypes.c_void_p, dp, ctypes.POINTER(MrConfig), dp, dp, dp,
                                 ctypes.POINTER(GrapeResult)]
    _lib = L
    return L


def last_error() -> str:
    buf = ctypes.create_string_buffer(512)
    load().pr_last_error(buf, 512)
    return buf.value.decode(errors="replace")


def check(rc: int) -> None:
    if rc != 0:
        raise RocmError(rc, last_error())


def device_count() -> int:
    return int(load().pr_device_count())


def init(device: int = 0) -> None:
    check(load().pr_init(int(device)))


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Plan:
    """A device plan over one or more problems (pr_plan_create). The options are
    pr_options': precision, grad_mode, segment, store_props, and pack (1: small
    problems and the members of small ensembles run as diagonal blocks of
    16 x 16 problems; 0: each problem as given)."""

    def __init__(self, problems: Sequence[Problem], precision: int = 0, grad_mode: int = 0,
                 segment: int = 0, store_props: int = -1, pack: int = 1):
        L = load()
        self.problems = list(problems)
        n = len(self.problems)
        arr = (type(to_c(self.problems[0])) * n)()
        # every struct's arrays stay referenced until the plan has copied them,
        # also when one Problem object appears several times
        self._keeps = []
        for i, p in enumerate(self.problems):
            arr[i] = to_c(p)
            self._keeps.append(p._keep)
        self._structs = arr
        opt = COptions(int(precision), int(grad_mode), int(segment), int(store_props), int(pack))
        h = ctypes.c_void_p()
        check(L.pr_plan_create(ctypes.cast(arr, ctypes.c_void_p), n, ctypes.byref(opt), ctypes.byref(h)))
        self._h = h
        self.n_params = int(L.pr_plan_n_params(h))

    @property
    def handle(self):
        return self._h

    def eval(self, controls: Sequence[np.ndarray], grad: bool = True, diag: bool = False):
        """F (array [n_probs]) and G (list of [n_ctrl, n_t]) for per-problem controls."""
        L = load()
        n = len(self.problems)
        if len(controls) != n:
            raise ValueError(f"expected {n} control arrays, got {len(controls)}")
        bufs = []
        for p, u in zip(self.problems, controls):
            u = np.asarray(u, dtype=np.float64)
            if u.shape != (p.n_ctrl, p.n_t):
                raise ValueError(f"controls shape {u.shape} != ({p.n_ctrl}, {p.n_t})")
            bufs.append(controls_cm(u))
        cptrs = (ctypes.POINTER(ctypes.c_double) * n)(*[_dptr(b) for b in bufs])
        F = np.zeros(n)
        gbufs = [np.zeros(p.n_ctrl * p.n_t) for p in self.problems] if grad else []
        gptrs = (ctypes.POINTER(ctypes.c_double) * n)(*[_dptr(g) for g in gbufs]) if grad else None
        D = np.zeros(4 * n) if diag else None
        check(L.pr_eval(self._h, cptrs, _dptr(F), gptrs, _dptr(D) if diag else None))
        G = [grad_from_cm(g, p.n_ctrl, p.n_t) for g, p in zip(gbufs, self.problems)] if grad else None
        if diag:
            return F, G, D.reshape(n, 4)
        return F, G

    def eval_device(self, d_controls: int, d_F: int, d_G: int | None, stream: int | None = None) -> None:
        """Enqueue an evaluation on device pointers (ints), e.g. from torch tensors."""
        check(load().pr_eval_device(self._h, ctypes.c_void_p(d_controls), ctypes.c_void_p(d_F),
                                    ctypes.c_void_p(d_G) if d_G else None,
                                    ctypes.c_void_p(stream) if stream else None))

    def eval_device_static(self, d_controls: int, d_F: int, d_G: int | None, stream: int | None = None) -> None:
        """pr_eval_device_static: the same evaluation with the library's static
        scratch and no allocation (direct launches)."""
        check(load().pr_eval_device_static(self._h, ctypes.c_void_p(d_controls), ctypes.c_void_p(d_F),
                                           ctypes.c_void_p(d_G) if d_G else None,
                                           ctypes.c_void_p(stream) if stream else None))

    def buffers(self):
        """(d_controls, d_F, d_G) device addresses of the plan-owned buffers."""
        a, b, c = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        check(load().pr_plan_buffers(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def upload_controls(self, controls: Sequence[np.ndarray]) -> None:
        bufs = [controls_cm(np.asarray(u, np.float64)) for u in controls]
        for p, u in zip(self.problems, controls):
            if np.asarray(u).shape != (p.n_ctrl, p.n_t):
                raise ValueError(f"controls shape {np.asarray(u).shape} != ({p.n_ctrl}, {p.n_t})")
        cptrs = (ctypes.POINTER(ctypes.c_double) * len(bufs))(*[_dptr(b) for b in bufs])
        check(load().pr_plan_upload_controls(self._h, cptrs))

    def synchronize(self) -> None:
        check(load().pr_plan_synchronize(self._h))

    def set_timing(self, enable: bool) -> None:
        check(load().pr_plan_set_timing(self._h, 1 if enable else 0))

    def exp_counts(self) -> np.ndarray:
        """[n_probs, 4]: member-slices whose exp took degree 6, degree 9, the
        squarings, and the member-slices a reduced-precision plan redid in FP64,
        of the last evaluation."""
        n = len(self.problems)
        out = (ctypes.c_longlong * (4 * n))()
        check(load().pr_plan_exp_counts(self._h, out))
        return np.array(out[:], dtype=np.int64).reshape(n, 4)

    def shapes(self) -> np.ndarray:
        """[n_probs, 4]: slices per segment, segments per member, trace operand in
        registers (1/0), split assembly passes in registers (pr_plan_shapes)."""
        n = len(self.problems)
        out = (ctypes.c_int * (4 * n))()
        check(load().pr_plan_shapes(self._h, out))
        return np.array(out[:], dtype=np.int64).reshape(n, 4)

    def kernel_times(self) -> dict:
        """{class: (total_ms, launches)} since the previous call (HIP events)."""
        ms = np.zeros(len(KERNEL_CLASSES))
        n = (ctypes.c_longlong * len(KERNEL_CLASSES))()
        check(load().pr_plan_kernel_times(self._h, _dptr(ms), n))
        return {c: (float(ms[i]), int(n[i])) for i, c in enumerate(KERNEL_CLASSES)}

    def close(self) -> None:
        if getattr(self, "_h", None):
            load().pr_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def grape_optimize_device(fid_plan: Plan, grad_plan, controls: np.ndarray, cfg: GrapeConfig):
    """pr_grape_optimize: the gradient-ascent loop with controls resident on the
    device. controls [n_ctrl, n_t]; returns (best controls, fidelity history,
    gradient-norm history, GrapeResult)."""
    u = np.asarray(controls, dtype=np.float64)
    n_c, n_t = u.shape
    u0 = controls_cm(u)
    best = np.zeros_like(u0)
    fh = np.zeros(cfg.max_iter + 2)
    gh = np.zeros(cfg.max_iter)
    res = GrapeResult()
    check(load().pr_grape_optimize(fid_plan._h, grad_plan._h if grad_plan is not None else None, _dptr(u0),
                                   ctypes.byref(cfg), _dptr(best), _dptr(fh), _dptr(gh), ctypes.byref(res)))
    return (grad_from_cm(best, n_c, n_t), fh[:res.n_fidelity_history].tolist(), gh[:res.n_iterations].tolist(), res)


def lbfgs_optimize_device(plan: Plan, controls: np.ndarray, cfg: LbfgsConfig):
    """pr_lbfgs_optimize: the ensemble L-BFGS loop with the controls, gradient
    and curvature pairs resident on the device. controls [n_ctrl, n_t];
    returns (best controls, fidelity history, gradient-norm history,
    GrapeResult)."""
    u = np.asarray(controls, dtype=np.float64)
    n_c, n_t = u.shape
    u0 = controls_cm(u)
    best = np.zeros_like(u0)
    fh = np.zeros(cfg.max_iter + 1)
    gh = np.zeros(cfg.max_iter)
    res = GrapeResult()
    check(load().pr_lbfgs_optimize(plan._h, _dptr(u0), ctypes.byref(cfg), _dptr(best), _dptr(fh), _dptr(gh),
                                   ctypes.byref(res)))
    return (grad_from_cm(best, n_c, n_t), fh[:res.n_fidelity_history].tolist(), gh[:res.n_iterations].tolist(), res)


def mr_optimcon_device(plan: Plan, guess: np.ndarray, cfg: MrConfig):
    """pr_mr_optimcon: the MR optimcon loop (:lbfgs or :grape) with the
    waveform, gradient and curvature pairs resident on the device. guess
    [n_ctrl, n_t] (cfg.tile copies of its rows make the plan's controls);
    returns (best waveform, fidelity history, gradient-norm history, GrapeResult)."""
    u = np.asarray(guess, dtype=np.float64)
    n_c, n_t = u.shape
    u0 = controls_cm(u)
    best = np.zeros_like(u0)
    fh = np.zeros(cfg.max_iter + 1)
    gh = np.zeros(cfg.max_iter + 1)
    res = GrapeResult()
    check(load().pr_mr_optimcon(plan._h, _dptr(u0), ctypes.byref(cfg), _dptr(best), _dptr(fh), _dptr(gh),
                                ctypes.byref(res)))
    n = res.n_fidelity_history
    return grad_from_cm(best, n_c, n_t), fh[:n].tolist(), gh[:n].tolist(), res
