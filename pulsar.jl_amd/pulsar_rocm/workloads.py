"""Synthetic GRAPE workloads (seeded numpy generators).

Each builder returns (Problem, controls[n_ctrl, n_t]) for one of the five
benchmark configurations:

  single_qubit_state   d=2 state GRAPE, H0=(2pi*10/2) sz, controls 0.5 sx, 0.5 sy,
                       T=pi over 100 slices (the reference README quickstart)
  four_qubit_ensemble  d=16 gate synthesis, 256 members (16 detunings x 16
                       control scales), 8 x/y drives, 2000 slices of 1 us
  three_qubit_lindblad N=8 (64x64 Liouvillian), T1/T2 jump operators, 6 drives,
                       1000 slices of 2 ns, 64 amplitude-scale members,
                       Haar-unitary channel target on the 64 vec(E_ij) pairs
  uneven_batch         independent problems, dims {2..64} with p ~ 1/d
  spin_chain_128       d=128 Heisenberg chain, 14 drives, 500 slices, 64 members
  four_qubit_tracking  d=16 tracking GRAPE (density-matrix checkpoints,
                       GRAPETracking.jl): the four_qubit_ensemble system and
                       ensemble, 4 initial states, 2000 slices, 21 checkpoints

The Haar-random target follows random_unitary_problem
(src/Utilities/ProblemLibrary.jl:742-748): QR of a complex Gaussian matrix,
columns rephased by sign(diag R). The ensemble order is the Cartesian product
order of quasi_static_ensemble (src/Application/QuantumComputing/NoiseModels/
QuasiStatic.jl:132-155): the first noise source is the outer loop.
"""
from __future__ import annotations

import numpy as np

from .problem import (PR_CLOSED, PR_GATE, PR_LINDBLAD, PR_PAIRS_REAL, PR_PAIRS_SQUARE, PR_STATE,
                      PR_TRACKING, Problem)

SX = np.array([[0, 1], [1, 0]], dtype=np.complex128)
SY = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
SZ = np.array([[1, 0], [0, -1]], dtype=np.complex128)
I2 = np.eye(2, dtype=np.complex128)
TWO_PI = 2.0 * np.pi


def embed(op: np.ndarray, n_qubits: int, q: int) -> np.ndarray:
    """Single-qubit operator on qubit q (0-based, qubit 0 = most significant factor)."""
    out = np.array([[1.0 + 0j]])
    for k in range(n_qubits):
        out = np.kron(out, op if k == q else I2)
    return out


def embed2(a: np.ndarray, b: np.ndarray, n_qubits: int, q: int) -> np.ndarray:
    """a on qubit q times b on qubit q+1."""
    out = np.array([[1.0 + 0j]])
    for k in range(n_qubits):
        out = np.kron(out, a if k == q else (b if k == q + 1 else I2))
    return out


def haar_unitary(d: int, rng: np.random.Generator) -> np.ndarray:
    z = (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))) / np.sqrt(2.0)
    q, r = np.linalg.qr(z)
    ph = np.diag(r) / np.abs(np.diag(r))
    return q * ph[None, :]


def single_qubit_state(seed: int = 42, psi_init=None):
    """d=2 GRAPE, the reference README quickstart: fixed-point target |1> (no
    psi_init) or |0> -> |1> transfer when psi_init is given."""
    H0 = (TWO_PI * 10 / 2) * SZ
    Hc = np.stack([0.5 * SX, 0.5 * SY])
    n_t = 100
    dt = np.pi / n_t
    rng = np.random.default_rng(seed)
    u = 0.5 * (2 * rng.random((2, n_t)) - 1)
    tgt = np.array([[0, 1]], dtype=np.complex128)
    ini = tgt if psi_init is None else np.asarray(psi_init, np.complex128).reshape(1, 2)
    p = Problem(dim=2, n_ctrl=2, n_t=n_t, H0=H0, Hc=Hc, dt=dt, dynamics=PR_CLOSED,
                objective=PR_STATE, n_members=1, psi_init=ini, psi_targ=tgt,
                weight=np.ones(1))
    return p, u


def four_qubit_drift(omega, J) -> np.ndarray:
    nq = 4
    H = np.zeros((16, 16), dtype=np.complex128)
    for i in range(nq):
        H += omega[i] * embed(SZ, nq, i) / 2
    for i in range(nq - 1):
        H += J[i] * embed2(SZ, SZ, nq, i)
    return H


def four_qubit_parameters(seed: int = 7):
    """(rng, omega, J, Hc) of the four-qubit system: the qubit frequencies and
    couplings are the first draws of the seeded stream, which then goes on to
    draw the controls."""
    rng = np.random.default_rng(seed)
    nq = 4
    omega = rng.uniform(-TWO_PI * 50, TWO_PI * 50, nq)
    J = rng.uniform(TWO_PI * 5, TWO_PI * 20, nq - 1)
    Hc = np.stack([embed(s, nq, q) / 2 for q in range(nq) for s in (SX, SY)])
    return rng, omega, J, Hc


def four_qubit_nominal(seed: int = 7, n_t: int = 2000):
    """The nominal system of the ensemble (no detuning, amplitude scale 1) as a
    one-member problem with the ensemble's controls; its target is left unset."""
    rng, omega, J, Hc = four_qubit_parameters(seed)
    u = rng.uniform(-TWO_PI * 1000, TWO_PI * 1000, (Hc.shape[0], n_t))
    p = Problem(dim=16, n_ctrl=Hc.shape[0], n_t=n_t, H0=four_qubit_drift(omega, J), Hc=Hc, dt=1e-6,
                dynamics=PR_CLOSED, objective=PR_GATE, n_members=1, U_target=np.eye(16, dtype=np.complex128))
    return p, u


def four_qubit_ensemble(seed: int = 7, n_t: int = 2000, n_det: int = 16, n_amp: int = 16,
                        target_seed: int = 2):
    """d=16 closed-system gate synthesis with a detuning x amplitude ensemble."""
    rng, omega, J, Hc = four_qubit_parameters(seed)
    dets = np.linspace(-0.05, 0.05, n_det)
    amps = np.linspace(0.9, 1.1, n_amp)
    H0s, scales = [], []
    for dlt in dets:
        Hd = four_qubit_drift(omega * (1 + dlt), J)
        for a in amps:
            H0s.append(Hd)
            scales.append(np.full(Hc.shape[0], a))
    M = len(H0s)
    dt = 1e-6
    u = rng.uniform(-TWO_PI * 1000, TWO_PI * 1000, (Hc.shape[0], n_t))
    Ut = haar_unitary(16, np.random.default_rng(target_seed))
    p = Problem(dim=16, n_ctrl=Hc.shape[0], n_t=n_t, H0=np.stack(H0s), Hc=Hc, dt=dt,
                dynamics=PR_CLOSED, objective=PR_GATE, n_members=M,
                ctrl_scale=np.stack(scales), U_target=Ut)
    return p, u


# ---------------------------------------------------------------- Lindblad
def drift_liouvillian(H, jumps, rates):
    """build_drift_liouvillian (src/Physics/Lindblad.jl:126-157), column-stacking vec."""
    N = H.shape[0]
    I = np.eye(N)
    L = -1j * np.kron(I, H) + 1j * np.kron(H.T, I)
    for Lk, g in zip(jumps, rates):
        if g == 0.0 or not np.isfinite(g):
            continue
        A = Lk.conj().T @ Lk
        L = L + g * (np.kron(Lk.conj(), Lk) - 0.5 * np.kron(I, A) - 0.5 * np.kron(A.conj(), I))
    return L


def control_liouvillian(Hk):
    """build_control_liouvillian (src/Physics/Lindblad.jl:189-197)."""
    N = Hk.shape[0]
    I = np.eye(N)
    return -1j * np.kron(I, Hk) + 1j * np.kron(Hk.T, I)


def relaxation_ops(n_qubits, T1s, T2s):
    """mr_relaxation (src/Application/MR/LindbladMR.jl:109-157) for spin-1/2
    qubits: I+ and I- at 1/(2 T1), 2 Iz at 1/T2 - 1/(2 T1) when positive."""
    Ip = np.array([[0, 1], [0, 0]], dtype=np.complex128)
    Im = Ip.T.copy()
    Iz = SZ / 2
    jumps, rates = [], []
    for k in range(n_qubits):
        g1 = 1.0 / (2.0 * T1s[k])
        jumps += [embed(Ip, n_qubits, k), embed(Im, n_qubits, k)]
        rates += [g1, g1]
        gz = 1.0 / T2s[k] - g1
        if gz > 0:
            jumps.append(2.0 * embed(Iz, n_qubits, k))
            rates.append(gz)
    return jumps, rates


def three_qubit_lindblad(seed: int = 11, n_t: int = 1000, n_members: int = 64,
                         objective: int = PR_PAIRS_REAL, target_seed: int = 5,
                         amplitude: float = TWO_PI * 5e6):
    """N=8 open system: 64x64 Liouvillians, channel target on vec(E_ij)."""
    rng = np.random.default_rng(seed)
    nq, N = 3, 8
    omega = rng.uniform(-TWO_PI * 2e6, TWO_PI * 2e6, nq)
    Jzz = rng.uniform(TWO_PI * 0.2e6, TWO_PI * 1e6, nq - 1)
    H = sum(omega[i] * embed(SZ, nq, i) / 2 for i in range(nq))
    H = H + sum(Jzz[i] * embed2(SZ, SZ, nq, i) / 4 for i in range(nq - 1))
    T1 = rng.uniform(50e-6, 200e-6, nq)
    T2 = rng.uniform(20e-6, 100e-6, nq)
    jumps, rates = relaxation_ops(nq, T1, T2)
    L0 = drift_liouvillian(H, jumps, rates)
    Hc = [embed(s, nq, q) / 2 for q in range(nq) for s in (SX, SY)]
    Lc = np.stack([control_liouvillian(h) for h in Hc])
    scales = np.linspace(0.95, 1.05, n_members)
    Ut = haar_unitary(N, np.random.default_rng(target_seed))
    St = np.kron(Ut.conj(), Ut)  # vec(U rho U^dag) = (conj(U) (x) U) vec(rho)
    D = N * N
    psi_i = np.eye(D, dtype=np.complex128)
    psi_t = np.ascontiguousarray(St.T)  # row s = S_t e_s
    dt = 2e-9
    u = rng.uniform(-amplitude, amplitude, (Lc.shape[0], n_t))
    p = Problem(dim=D, n_ctrl=Lc.shape[0], n_t=n_t, H0=L0, Hc=Lc, dt=dt,
                dynamics=PR_LINDBLAD, objective=objective, n_members=n_members,
                ctrl_scale=np.repeat(scales[:, None], Lc.shape[0], axis=1),
                psi_init=psi_i, psi_targ=psi_t)
    return p, u


def random_hermitian(d, rng, spectral_norm):
    a = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    h = (a + a.conj().T) / 2
    return h * (spectral_norm / np.max(np.abs(np.linalg.eigvalsh(h))))


def uneven_batch(n_problems: int = 1024, seed: int = 3, n_min: int = 200, n_max: int = 2000,
                 dims=(2, 4, 8, 16, 32, 64)):
    """Independent gate problems of uneven dimension and length."""
    rng = np.random.default_rng(seed)
    w = np.array([1.0 / d for d in dims])
    w /= w.sum()
    out = []
    for _ in range(n_problems):
        d = int(rng.choice(dims, p=w))
        nq = int(round(np.log2(d)))
        n_t = int(rng.integers(n_min, n_max + 1))
        H0 = random_hermitian(d, rng, TWO_PI * 100)
        Hc = np.stack([TWO_PI * 50 * embed(s, nq, q) / 2 for q in range(nq) for s in (SX, SY)])
        u = rng.uniform(-0.5, 0.5, (Hc.shape[0], n_t))
        Ut = haar_unitary(d, rng)
        p = Problem(dim=d, n_ctrl=Hc.shape[0], n_t=n_t, H0=H0, Hc=Hc, dt=1e-4,
                    dynamics=PR_CLOSED, objective=PR_GATE, n_members=1, U_target=Ut)
        out.append((p, u))
    return out


def spin_chain_128(seed: int = 13, n_t: int = 500, n_members: int = 64, target_seed: int = 17):
    """7-qubit Heisenberg chain, d=128, amplitude-scale ensemble."""
    rng = np.random.default_rng(seed)
    nq, d = 7, 128
    J = rng.uniform(TWO_PI * 10, TWO_PI * 50, nq - 1)
    H0 = np.zeros((d, d), dtype=np.complex128)
    for i in range(nq - 1):
        for s in (SX, SY, SZ):
            H0 += J[i] * embed2(s, s, nq, i) / 4
    Hc = np.stack([embed(s, nq, q) / 2 for q in range(nq) for s in (SX, SY)])
    scales = np.linspace(0.95, 1.05, n_members)
    u = rng.uniform(-TWO_PI * 100, TWO_PI * 100, (Hc.shape[0], n_t))
    Ut = haar_unitary(d, np.random.default_rng(target_seed))
    p = Problem(dim=d, n_ctrl=Hc.shape[0], n_t=n_t, H0=H0, Hc=Hc, dt=1e-4, dynamics=PR_CLOSED,
                objective=PR_GATE, n_members=n_members,
                ctrl_scale=np.repeat(scales[:, None], Hc.shape[0], axis=1), U_target=Ut)
    return p, u


def random_state(d: int, rng: np.random.Generator) -> np.ndarray:
    v = rng.standard_normal(d) + 1j * rng.standard_normal(d)
    return v / np.linalg.norm(v)


def random_density(d: int, rng: np.random.Generator, rank: int = 2) -> np.ndarray:
    """A mixed state: sum of `rank` random pure states with random weights."""
    p = rng.uniform(0.2, 1.0, rank)
    p /= p.sum()
    out = np.zeros((d, d), np.complex128)
    for pi in p:
        v = random_state(d, rng)
        out += pi * np.outer(v, v.conj())
    return out


def tracking_checkpoints(d: int, n_t: int, n_checkpoints: int, rng: np.random.Generator):
    """Checkpoints every n_t / n_checkpoints slices alternating between two
    pure targets (the oscillation example of GRAPETracking.jl:69-80), plus a
    mixed target at the last slice; weights sum to 1."""
    a, b = random_state(d, rng), random_state(d, rng)
    per = max(1, n_t // n_checkpoints)
    steps = [min(n_t, k * per) for k in range(1, n_checkpoints + 1)]
    targets = [np.outer(v, v.conj()) for v in (b if k % 2 else a for k in range(1, n_checkpoints + 1))]
    steps.append(n_t)
    targets.append(random_density(d, rng))
    w = np.full(len(steps), 1.0 / len(steps))
    return np.array(steps, np.int32), w, np.stack(targets)


def four_qubit_tracking(seed: int = 23, n_t: int = 2000, n_det: int = 16, n_amp: int = 16, n_pairs: int = 4,
                        n_checkpoints: int = 20):
    """d=16 tracking GRAPE: the four_qubit_ensemble drifts, drives and
    detuning x amplitude ensemble; n_pairs random pure initial states."""
    p, u = four_qubit_ensemble(n_t=n_t, n_det=n_det, n_amp=n_amp)
    rng = np.random.default_rng(seed)
    steps, w, targets = tracking_checkpoints(16, n_t, n_checkpoints, rng)
    p.objective = PR_TRACKING
    p.U_target = None
    p.psi_init = np.stack([random_state(16, rng) for _ in range(n_pairs)])
    p.track_step, p.track_weight, p.track_target = steps, w, targets
    return p, u


def shrink(p: Problem, u: np.ndarray, n_members: int | None = None, n_t: int | None = None):
    """The same workload with fewer members and/or slices (for bounded checks)."""
    import copy
    q = copy.copy(p)
    uu = u
    if n_members is not None and n_members < p.n_members:
        q.n_members = n_members
        if np.asarray(p.H0).ndim == 3:
            q.H0 = np.asarray(p.H0)[:n_members]
        if p.ctrl_scale is not None:
            q.ctrl_scale = np.asarray(p.ctrl_scale)[:n_members]
        if p.weight is not None:
            q.weight = np.asarray(p.weight)[:n_members]
    if n_t is not None and n_t < p.n_t:
        q.n_t = n_t
        uu = u[:, :n_t]
        if p.dt_slices is not None:
            q.dt_slices = np.asarray(p.dt_slices)[:n_t]
        if p.track_step is not None:
            keep = np.asarray(p.track_step) <= n_t
            q.track_step = np.asarray(p.track_step)[keep]
            q.track_weight = np.asarray(p.track_weight)[keep]
            q.track_target = np.asarray(p.track_target)[keep]
    q._keep = []
    return q, np.ascontiguousarray(uu)


def mr_spin_ensemble(seed: int = 29, n_off: int = 64, n_pwr: int = 16, n_t: int = 500):
    """MR state transfer of one spin-1/2, +z -> -y (the OptControl.jl:296-305
    example): offsets in +-2 kHz, RF power levels 0.8..1.2 x 10 kHz, 2 us
    slices; returns (MRControl, waveform [2 x n_t])."""
    from .mr import MRControl
    Iz, Ix, Iy = SZ / 2, SX / 2, SY / 2
    up = np.array([1, 0], np.complex128)
    mIy = np.array([1, -1j], np.complex128) / np.sqrt(2)
    ctrl = MRControl(drifts=[TWO_PI * o * Iz for o in np.linspace(-2e3, 2e3, n_off)], operators=[Ix, Iy],
                     rho_init=[up], rho_targ=[mIy], pulse_dt=np.full(n_t, 2e-6),
                     pwr_levels=list(TWO_PI * 1e4 * np.linspace(0.8, 1.2, n_pwr)))
    w = 0.05 * np.random.default_rng(seed).standard_normal((2, n_t))
    return ctrl, w
