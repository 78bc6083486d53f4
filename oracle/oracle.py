"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the device path. Imported by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg; never by the
product package.

tracking_evaluate is a numpy restatement (density matrices) of the tracking
kernel; tests/test_tracking_oracle.py pins it (squared overlaps from scipy
expm products, a central difference of the left-perturbed propagator, the
exact derivative for commuting generators, Hermitian-part and linearity
invariants). mas_powder_gradient is a numpy restatement of the MAS powder
gradient; tests/test_mas_oracle.py pins it (Wigner d-matrix relations of the
Fourier components, the magic angle, a central difference of the trace of the
right-perturbed propagator, the gate oracle with U_t = I).
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


def tracking_evaluate(p: Problem, u: np.ndarray, grad: bool = True):
    """F and G[n_ctrl, n_t] of the tracking objective (PR_TRACKING), restating
    _grape_tracking_cpu_threaded (src/Application/MR/GRAPETracking.jl:162-263)
    with density matrices, one member (drift x power level) at a time:

      U_n = exp(-i H_n dt_n), H_n = H0_m + sum_j s_mj u_jn A_j        (:200-206)
      rho[1] = psi psi^dag, rho[n+1] = U_n rho[n] U_n^dag            (:214-218)
      F_m += w_c Re Tr(rho_c rho[step_c + 1])                         (:221-225)
      backward n = n_t..1: Lambda += 2 w_c (rho_c - rho[n+1]) at a
      checkpoint step (:231-235); G[j,n] += -dt_n Re(i Tr(Lambda^dag
      [s_mj A_j, rho[n+1]])) (:237-243); Lambda = U_n^dag Lambda U_n (:246)

    F = sum_m w_m F_m and G = sum_m w_m G_m with w_m = 1/(n_members n_pairs)
    unless p.weight is given (the division by N_ens, :252-258). The
    reference's power level pwr is s_mj = pwr for every control. Exps come
    from or_expm_neg_i (eigen-decomposition of the Hermitian upper triangle,
    as _expm_neg_i_into!, src/Computation/Propagators.jl:659-674). Pure
    Python loops over slices: for small cases only.
    """
    d, nc, nt, M = p.dim, p.n_ctrl, p.n_t, p.n_members
    u = np.asarray(u, np.float64)
    H0 = np.asarray(p.H0, np.complex128)
    H0s = H0 if H0.ndim == 3 else np.broadcast_to(H0, (M, d, d))
    Hc = np.asarray(p.Hc, np.complex128)
    sc = np.ones((M, nc)) if p.ctrl_scale is None else np.asarray(p.ctrl_scale, np.float64)
    dts = np.full(nt, float(p.dt)) if p.dt_slices is None else np.asarray(p.dt_slices, np.float64)
    npairs = p.n_pairs
    w = np.full(M, 1.0 / (M * npairs)) if p.weight is None else np.asarray(p.weight, np.float64)
    if p.n_track < 1:
        raise ValueError("tracking objective without checkpoints")
    ck = {}
    for step, wt, tg in zip(np.asarray(p.track_step), np.asarray(p.track_weight), np.asarray(p.track_target)):
        if not 1 <= int(step) <= nt:
            raise ValueError(f"TrackingPoint.step={int(step)} out of range [1, {nt}]")
        ck.setdefault(int(step), []).append((float(wt), np.asarray(tg, np.complex128)))
    steps = sorted(ck)
    F = 0.0
    G = np.zeros((nc, nt))
    for m in range(M):
        Ps = []
        for n in range(nt):
            H = H0s[m].copy()
            for j in range(nc):
                H = H + sc[m, j] * u[j, n] * Hc[j]
            Ps.append(expm_neg_i(H, dts[n]))
        Hk = [sc[m, j] * Hc[j] for j in range(nc)]
        fm = 0.0
        gm = np.zeros((nc, nt))
        for s in range(npairs):
            psi = np.asarray(p.psi_init[s], np.complex128)
            rho = [np.outer(psi, psi.conj())]
            for n in range(nt):
                rho.append(Ps[n] @ rho[n] @ Ps[n].conj().T)
            for step in steps:
                for wt, tg in ck[step]:
                    fm += wt * np.real(np.trace(tg @ rho[step]))
            if not grad:
                continue
            Lam = np.zeros((d, d), np.complex128)
            for n in range(nt - 1, -1, -1):
                for wt, tg in ck.get(n + 1, ()):
                    Lam = Lam + 2.0 * wt * (tg - rho[n + 1])
                r = rho[n + 1]
                for j in range(nc):
                    comm = Hk[j] @ r - r @ Hk[j]
                    gm[j, n] += -dts[n] * np.real(1j * np.trace(Lam.conj().T @ comm))
                Lam = Ps[n].conj().T @ Lam @ Ps[n]
        F += w[m] * fm
        G += w[m] * gm
    return F, (G if grad else None)


_SPIN_I = {"1H": 0.5, "2H": 1.0, "13C": 0.5, "14N": 1.0, "15N": 0.5, "17O": 2.5, "19F": 0.5, "23Na": 1.5,
           "29Si": 0.5, "31P": 0.5, "129Xe": 0.5}


def spin_operators(isotopes):
    """(Ix, Iy, Iz) per spin in the tensor-product space (NMRSpinSystem.jl:65-171):
    Iz = diag(I, I-1, ..., -I), I+ = sqrt(I(I+1) - m(m+1)) above the diagonal."""
    Is = [_SPIN_I[i] for i in isotopes]
    dims = [int(round(2 * I + 1)) for I in Is]

    def single(I):
        n = int(round(2 * I + 1))
        ms = [I - r for r in range(n)]
        Iz = np.diag(ms).astype(complex)
        Ip = np.zeros((n, n), complex)
        for r in range(n - 1):
            m = ms[r + 1]
            Ip[r, r + 1] = np.sqrt(I * (I + 1) - m * (m + 1))
        return (Ip + Ip.conj().T) / 2, (Ip - Ip.conj().T) / 2j, Iz

    def embed(k, op):
        out = np.ones((1, 1), complex)
        for i, di in enumerate(dims):
            out = np.kron(out, op if i == k else np.eye(di))
        return out

    ops = [single(I) for I in Is]
    return ([embed(k, ops[k][a]) for k in range(len(Is))] for a in range(3))


def mas_fourier(isotopes, csa, dipolar, orientation=(0.0, 0.0, 0.0)):
    """{m: H_m} of the MAS drift for the system rotated by `orientation`
    (Euler angles added, MASSpinSystem.jl:190-199; components :108-166: CSA
    tensor i on spin i, dipolar couplings on Iz_i Iz_j, sidebands only when
    sin(beta) > 1e-12). csa: (delta_iso, delta_aniso, eta, (a, b, g));
    dipolar: (i, j, b_ij, (a, b, g)) with 1-based spins."""
    Ix, Iy, Iz = spin_operators(isotopes)
    D = Iz[0].shape[0]
    oa, ob, og = orientation
    Hf = {m: np.zeros((D, D), complex) for m in range(-2, 3)}
    for i, (d_iso, d_an, _eta, (a, b, g)) in enumerate(csa):
        if i >= len(Iz):
            break
        a, b = a + oa, b + ob
        T20 = d_an * np.sqrt(2 / 3)
        Hf[0] += (d_iso + T20 * (3 * np.cos(b) ** 2 - 1) / 2) * Iz[i]
        if np.sin(b) > 1e-12:
            c1 = -np.sqrt(3 / 2) * T20 * np.sin(b) * np.cos(b) * np.exp(-1j * a)
            c2 = np.sqrt(3 / 8) * T20 * np.sin(b) ** 2 * np.exp(-2j * a)
            Hf[1] += c1 * Iz[i]
            Hf[-1] += np.conj(c1) * Iz[i]
            Hf[2] += c2 * Iz[i]
            Hf[-2] += np.conj(c2) * Iz[i]
    for (si, sj, bij, (a, b, g)) in dipolar:
        if not (si <= len(Iz) and sj <= len(Iz)):
            continue
        a, b, g = a + oa, b + ob, g + og
        ZZ = Iz[si - 1] @ Iz[sj - 1]
        Hf[0] += 2 * (bij * (3 * np.cos(b) ** 2 - 1) / 2) * ZZ
        if np.sin(b) > 1e-12:
            c1 = -bij * np.sqrt(6.0) * np.sin(b) * np.cos(b) * np.exp(-1j * (a + g))
            c2 = bij * np.sqrt(6.0) / 4 * np.sin(b) ** 2 * np.exp(-2j * (a + g))
            Hf[1] += c1 * ZZ
            Hf[-1] += np.conj(c1) * ZZ
            Hf[2] += c2 * ZZ
            Hf[-2] += np.conj(c2) * ZZ
    return Hf


def mas_propagators(Hf, Hc, omega_r_hz, u, dt, rotor_phase_0=0.0):
    """U_k = exp(-i dt (sum_m e^{i m w_r t_k} H_m + sum_j u_jk H_j)),
    t_k = (k - 1/2) dt + phi0 / w_r (MASPropagators.jl:14-35)."""
    w_r = 2 * np.pi * omega_r_hz
    u = np.asarray(u, np.float64)
    D = Hf[0].shape[0]
    Us = []
    for k in range(u.shape[1]):
        t = (k + 0.5) * dt + rotor_phase_0 / w_r
        H = np.zeros((D, D), complex)
        for m in range(-2, 3):
            H = H + np.exp(1j * m * w_r * t) * Hf[m]
        for j in range(u.shape[0]):
            H = H + u[j, k] * Hc[j]
        Us.append(expm_neg_i(H, dt))
    return Us


def propagator_gradient(Us, Hc, U_target, dt, psi_init=None, psi_targ=None):
    """_grape_gradient_from_propagators (MRPhysics.jl:312-351):
    P_k = U_{k-1}..U_1, Q_k = (U_n..U_k)^dag (Propagators.jl:277-380);
    unitary branch G[j,k] = 2 Re(conj(Phi) Tr(Q_k^dag (-i dt H_j) P_k)) / d^2;
    with psi_targ given, the state branch (:337-351):
    G[j,k] = 2 Im(conj(ov) <Q_k psi_t| -i dt H_j |P_k psi_i>), ov = <psi_t|U psi_i>,
    psi_i = psi_init or psi_t, states as given."""
    nt, D = len(Us), Us[0].shape[0]
    P = [np.eye(D, dtype=complex)]
    for k in range(nt):
        P.append(Us[k] @ P[k])
    Q = [None] * (nt + 1)
    Q[nt] = np.eye(D, dtype=complex)
    for k in range(nt - 1, -1, -1):
        Q[k] = Us[k].conj().T @ Q[k + 1]
    G = np.zeros((len(Hc), nt))
    if psi_targ is not None:
        pt = np.asarray(psi_targ, complex).reshape(-1)
        pi = pt if psi_init is None else np.asarray(psi_init, complex).reshape(-1)
        ov = np.vdot(pt, P[nt] @ pi)
        for k in range(nt):
            psi_k, lam_k = P[k] @ pi, Q[k] @ pt
            for j, Hj in enumerate(Hc):
                G[j, k] = 2 * np.imag(np.conj(ov) * np.vdot(lam_k, (-1j * dt * Hj) @ psi_k))
        return G
    Phi = np.trace(np.asarray(U_target, complex).conj().T @ P[nt]) / D
    for k in range(nt):
        for j, Hj in enumerate(Hc):
            inner = np.trace(Q[k].conj().T @ (-1j * dt * Hj) @ P[k])
            G[j, k] = 2 * np.real(np.conj(Phi) * inner) / D ** 2
    return G


def mas_powder_gradient(isotopes, csa, dipolar, omega_r_hz, u, dt, U_target, orientations,
                        rotor_phase_0: float = 0.0, psi_init=None, psi_targ=None):
    """compute_grape_gradient_powder (src/Physics/MRPhysics.jl:170-188):
    G = sum_o w_o G_o over orientations (a, b, g, w), each G_o from
    mas_fourier -> mas_propagators -> propagator_gradient with the RF controls
    [Ix_1, Iy_1, Ix_2, Iy_2, ...] (MASSpinSystem.jl:95-100)."""
    Ix, Iy, _ = spin_operators(isotopes)
    Hc = [h for k in range(len(Ix)) for h in (Ix[k], Iy[k])]
    G = 0.0
    for (oa, ob, og, ow) in orientations:
        Hf = mas_fourier(isotopes, csa, dipolar, (oa, ob, og))
        Us = mas_propagators(Hf, Hc, omega_r_hz, u, dt, rotor_phase_0)
        G = G + ow * propagator_gradient(Us, Hc, U_target, dt, psi_init, psi_targ)
    return G


def total_propagator(p: Problem, u: np.ndarray, member: int = 0):
    c = to_c(p)
    ub = controls_cm(u)
    out = np.zeros(2 * p.dim * p.dim)
    rc = lib().or_total_propagator(ctypes.byref(c), int(member), _p(ub), _p(out))
    if rc != 0:
        raise RuntimeError(f"or_total_propagator failed with status {rc}")
    return _from_cm(out, p.dim)


__all__ = ["build", "lib", "expm_neg_i", "expm", "expm_frechet", "eigh", "evaluate",
           "tracking_evaluate", "spin_operators", "mas_fourier", "mas_propagators", "propagator_gradient",
           "mas_powder_gradient", "total_propagator", "PR_GRAD_EXACT", "PR_GRAD_REFERENCE"]
