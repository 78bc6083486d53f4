"""Magic-angle-spinning (MAS) GRAPE gradients on the :rocm device.

  mr_system, MRSpinSystem                 src/Types/NMRSpinSystem.jl:48-174
  CSATensor, DipolarCoupling,
  MASSpinSystem, mas_spin_system,
  rotate_spin_system, build_mas_hamiltonian
                                          src/Types/MASSpinSystem.jl:16-199
  compute_grape_gradient (MAS), compute_grape_gradient_powder
                                          src/Physics/MRPhysics.jl:150-188, 312-351

The MAS drift is H(t) = sum_{m=-2..2} exp(i m w_r t) H_m, evaluated at the
slice midpoints t_k = (k - 1/2) dt + phi0 / w_r (MASPropagators.jl:14-35).
Every H_m is a sum of fixed operators O_b (Iz_i for the i-th CSA tensor,
Iz_i Iz_j for a dipolar coupling) with orientation-dependent coefficients
c_mb, and H_{-m} = H_m^dag, so

    H(t_k) = H_0 + sum_{m=1,2} sum_b [2 Re c_mb cos(m th_k) - 2 Im c_mb sin(m th_k)] O_b,

th_k = w_r t_k. On the device this is one plan problem: the members are the
powder orientations (weights w, drift H_0 of that orientation), the operator
list is the RF controls followed by four rotor channels per O_b (cos th, sin th,
cos 2th, sin 2th) whose waveform rows are fixed and whose per-member scales
are the coefficients above. The objective is PR_GATE_PROPAGATORS (unitary
targets) or PR_STATE_PROPAGATORS (state targets), the two branches of
_grape_gradient_from_propagators (MRPhysics.jl:312-351); the gradient rows of
the rotor channels are discarded.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Sequence, Tuple

import numpy as np

from . import native
from .problem import PR_CLOSED, PR_GATE_PROPAGATORS, PR_STATE_PROPAGATORS, Problem

# nuclear spin quantum numbers (NMRSpinSystem.jl:48-61)
SPIN_QUANTUM_NUMBER = {"1H": 0.5, "2H": 1.0, "13C": 0.5, "14N": 1.0, "15N": 0.5, "17O": 2.5, "19F": 0.5,
                       "23Na": 1.5, "29Si": 0.5, "31P": 0.5, "129Xe": 0.5, "electron": 0.5}


def _single_spin_ops(I: float):
    """Iz = diag(I, I-1, ..., -I), I+ with sqrt(I(I+1) - m(m+1)) above the diagonal."""
    n = int(round(2 * I + 1))
    ms = I - np.arange(n)
    Iz = np.diag(ms).astype(np.complex128)
    Ip = np.zeros((n, n), np.complex128)
    for row in range(n - 1):
        m = ms[row + 1]
        Ip[row, row + 1] = np.sqrt(I * (I + 1) - m * (m + 1))
    Im = Ip.conj().T
    return (Ip + Im) / 2, (Ip - Im) / 2j, Iz, Ip, Im


@dataclass
class MRSpinSystem:
    n_spins: int
    isotopes: List[str]
    spin_I: List[float]
    dim: int
    Ix: List[np.ndarray]
    Iy: List[np.ndarray]
    Iz: List[np.ndarray]
    Ip: List[np.ndarray]
    Im: List[np.ndarray]


def mr_system(isotopes) -> MRSpinSystem:
    """Tensor-product spin operators of nuclear spins (NMRSpinSystem.jl:139-174)."""
    if isinstance(isotopes, str):
        isotopes = [isotopes]
    for iso in isotopes:
        if iso == "electron":
            raise ValueError("mr_system is for nuclear spins only. For electron spins, "
                             "use epr_system(S, I_nuclei, B0_tesla, mw_freq_hz; ...).")
        if iso not in SPIN_QUANTUM_NUMBER:
            raise ValueError(f"Unknown isotope '{iso}'. Supported: "
                             + ", ".join(sorted(k for k in SPIN_QUANTUM_NUMBER if k != "electron")))
    Is = [SPIN_QUANTUM_NUMBER[i] for i in isotopes]
    dims = [int(round(2 * I + 1)) for I in Is]
    singles = [_single_spin_ops(I) for I in Is]

    def embed(k, op):
        out = np.array([[1.0 + 0j]])
        for i, di in enumerate(dims):
            out = np.kron(out, op if i == k else np.eye(di))
        return out

    ops = [[embed(k, singles[k][a]) for k in range(len(Is))] for a in range(5)]
    return MRSpinSystem(len(Is), list(isotopes), Is, int(np.prod(dims)), *ops)


@dataclass
class CSATensor:
    """Chemical-shift anisotropy of one spin (MASSpinSystem.jl:16-24), rad/s."""

    delta_iso: float
    delta_aniso: float
    eta: float = 0.0
    euler_PAS_to_MF: Tuple[float, float, float] = (0.0, 0.0, 0.0)


@dataclass
class DipolarCoupling:
    """Dipolar coupling of spins spin_i, spin_j (1-based), b_ij in rad/s (MASSpinSystem.jl:33-40)."""

    spin_i: int
    spin_j: int
    b_ij: float
    euler_DD_to_MF: Tuple[float, float, float] = (0.0, 0.0, 0.0)


@dataclass
class MASSpinSystem:
    base_system: MRSpinSystem
    csa: List[CSATensor]
    dipolar: List[DipolarCoupling]
    omega_r: float
    rotor_period_s: float
    H_fourier: Dict[int, np.ndarray]
    H_drift: np.ndarray
    H_controls: List[np.ndarray]
    dim: int
    n_controls: int
    # (operator, {m: coefficient}) for m = 0, 1, 2: H_m = sum_b c_mb O_b, H_{-m} = H_m^dag
    terms: List[Tuple[np.ndarray, Dict[int, complex]]] = field(default_factory=list, repr=False)


def _rotor_terms(base: MRSpinSystem, csa, dipolar):
    """The coefficient of every operator in H_0, H_1, H_2: the CSA and dipolar
    Fourier components of MASSpinSystem.jl:108-166 (one term per CSA tensor,
    the i-th on spin i, and per dipolar coupling)."""
    terms = []
    n = base.n_spins
    for i, c in enumerate(csa):
        if i >= n:
            break
        T20 = c.delta_aniso * np.sqrt(2 / 3)
        a, b, _ = c.euler_PAS_to_MF
        cb, sb = np.cos(b), np.sin(b)
        co = {0: complex(c.delta_iso + T20 * (3 * cb ** 2 - 1) / 2), 1: 0j, 2: 0j}
        if sb > 1e-12:
            co[1] = -np.sqrt(3 / 2) * T20 * sb * cb * np.exp(-1j * a)
            co[2] = np.sqrt(3 / 8) * T20 * sb ** 2 * np.exp(-2j * a)
        terms.append((base.Iz[i], co))
    for dd in dipolar:
        i, j = dd.spin_i, dd.spin_j
        if not (i <= n and j <= n):
            continue
        O = base.Iz[i - 1] @ base.Iz[j - 1]
        a, b, g = dd.euler_DD_to_MF
        cb, sb = np.cos(b), np.sin(b)
        co = {0: complex(2 * dd.b_ij * (3 * cb ** 2 - 1) / 2), 1: 0j, 2: 0j}
        if sb > 1e-12:
            co[1] = -dd.b_ij * np.sqrt(6.0) * sb * cb * np.exp(-1j * (a + g))
            co[2] = dd.b_ij * np.sqrt(6.0) / 4 * sb ** 2 * np.exp(-2j * (a + g))
        terms.append((O, co))
    return terms


def mas_spin_system(base_system: MRSpinSystem, omega_r_hz: float, csa: Sequence[CSATensor] = (),
                    dipolar: Sequence[DipolarCoupling] = ()) -> MASSpinSystem:
    """mas_spin_system (MASSpinSystem.jl:80-106): Fourier components of the MAS
    drift and the RF controls [Ix_1, Iy_1, Ix_2, Iy_2, ...]."""
    omega_r = 2 * np.pi * omega_r_hz
    d = base_system.dim
    terms = _rotor_terms(base_system, list(csa), list(dipolar))
    Hf = {m: np.zeros((d, d), np.complex128) for m in range(-2, 3)}
    for O, co in terms:
        Hf[0] = Hf[0] + co[0] * O
        for m in (1, 2):
            Hf[m] = Hf[m] + co[m] * O
            Hf[-m] = Hf[-m] + np.conj(co[m]) * O
    Hc = []
    for k in range(base_system.n_spins):
        Hc.append(base_system.Ix[k])
        Hc.append(base_system.Iy[k])
    return MASSpinSystem(base_system, list(csa), list(dipolar), omega_r, 2 * np.pi / omega_r, Hf,
                         np.zeros((d, d), np.complex128), Hc, d, len(Hc), terms)


def rotate_spin_system(sys: MASSpinSystem, alpha: float, beta: float, gamma: float) -> MASSpinSystem:
    """The system at crystal orientation (alpha, beta, gamma): every tensor's
    Euler angles shifted by it (MASSpinSystem.jl:190-199)."""
    csa = [CSATensor(c.delta_iso, c.delta_aniso, c.eta,
                     (c.euler_PAS_to_MF[0] + alpha, c.euler_PAS_to_MF[1] + beta, c.euler_PAS_to_MF[2] + gamma))
           for c in sys.csa]
    dip = [DipolarCoupling(d.spin_i, d.spin_j, d.b_ij,
                           (d.euler_DD_to_MF[0] + alpha, d.euler_DD_to_MF[1] + beta, d.euler_DD_to_MF[2] + gamma))
           for d in sys.dipolar]
    return mas_spin_system(sys.base_system, sys.omega_r / (2 * np.pi), csa=csa, dipolar=dip)


def build_mas_hamiltonian(sys: MASSpinSystem, t: float) -> np.ndarray:
    """H(t) = sum_m H_m exp(i m w_r t) (MASSpinSystem.jl:174-181)."""
    return sum(np.exp(1j * m * sys.omega_r * t) * sys.H_fourier[m] for m in range(-2, 3))


def _target_fields(target):
    """Objective fields of a QuantumTarget (or a bare unitary matrix) for the
    helper: U_t for a unitary target; psi_t and psi_i = initial_state or
    psi_t, as given (MRPhysics.jl:338-339), for a state target."""
    if isinstance(target, np.ndarray):
        return dict(objective=PR_GATE_PROPAGATORS, U_target=np.asarray(target, np.complex128))
    kind = getattr(target, "type", "unitary")
    if kind == "unitary" and getattr(target, "target_unitary", None) is not None:
        return dict(objective=PR_GATE_PROPAGATORS, U_target=np.asarray(target.target_unitary, np.complex128))
    if kind == "state" and getattr(target, "target_state", None) is not None:
        psit = np.asarray(target.target_state, np.complex128).reshape(-1)
        init = getattr(target, "initial_state", None)
        psii = psit if init is None else np.asarray(init, np.complex128).reshape(-1)
        return dict(objective=PR_STATE_PROPAGATORS, psi_init=psii[None, :], psi_targ=psit[None, :])
    raise ValueError("the MAS gradient needs a unitary or a state target")


def mas_powder_problem(sys: MASSpinSystem, controls, target, orientations, rotor_phase_0: float = 0.0):
    """(Problem, extended controls) for the powder average: one member per
    orientation (alpha, beta, gamma, weight); see the module docstring."""
    ori = [tuple(float(v) for v in o) for o in orientations]
    if not ori:
        raise ValueError("at least one orientation is needed")
    u = np.asarray(controls.controls if hasattr(controls, "controls") else controls, np.float64)
    dt = float(controls.dt)
    nc, nt = u.shape
    if nc != sys.n_controls:
        raise ValueError(f"controls have {nc} channels, the system {sys.n_controls}")
    members = [rotate_spin_system(sys, a, b, g) for a, b, g, _ in ori]
    ops = [np.asarray(h, np.complex128) for h in sys.H_controls]
    nterm = len(members[0].terms)
    for t in range(nterm):
        ops += [members[0].terms[t][0]] * 4
    th = sys.omega_r * ((np.arange(nt) + 0.5) * dt + rotor_phase_0 / sys.omega_r)
    rot = np.stack([np.cos(th), np.sin(th), np.cos(2 * th), np.sin(2 * th)])
    u_ext = np.concatenate([u] + [rot] * nterm) if nterm else u
    scale = np.ones((len(ori), len(ops)))
    H0 = []
    for i, mem in enumerate(members):
        H0.append(sum((co[0] * O for O, co in mem.terms), np.zeros((sys.dim, sys.dim), np.complex128)))
        for t, (_, co) in enumerate(mem.terms):
            scale[i, nc + 4 * t:nc + 4 * t + 4] = [2 * co[1].real, -2 * co[1].imag, 2 * co[2].real, -2 * co[2].imag]
    p = Problem(dim=sys.dim, n_ctrl=len(ops), n_t=nt, H0=np.stack(H0), Hc=np.stack(ops), dt=dt,
                dynamics=PR_CLOSED, n_members=len(ori), ctrl_scale=scale,
                weight=np.array([o[3] for o in ori]), **_target_fields(target))
    return p, np.ascontiguousarray(u_ext)


def compute_grape_gradient_powder(sys: MASSpinSystem, controls, target, orientations) -> np.ndarray:
    """sum_o w_o G(rotate_spin_system(sys, alpha_o, beta_o, gamma_o)) with the
    propagator-helper gradient (MRPhysics.jl:170-188), [n_controls x n_t]."""
    p, u = mas_powder_problem(sys, controls, target, orientations)
    plan = native.Plan([p])
    try:
        _, G = plan.eval([u])
    finally:
        plan.close()
    return G[0][:sys.n_controls]


def compute_grape_gradient_mas(sys: MASSpinSystem, controls, target, rotor_phase_0: float = 0.0) -> np.ndarray:
    """compute_grape_gradient(sys::MASSpinSystem, ...) (MRPhysics.jl:155-162):
    one orientation, the system as given."""
    p, u = mas_powder_problem(sys, controls, target, [(0.0, 0.0, 0.0, 1.0)], rotor_phase_0)
    plan = native.Plan([p])
    try:
        _, G = plan.eval([u])
    finally:
        plan.close()
    return G[0][:sys.n_controls]
