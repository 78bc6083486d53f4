"""Pins the MAS oracle (oracle.mas_fourier / mas_propagators /
propagator_gradient / mas_powder_gradient, CPU only) and the host-side rotor
channel decomposition of pulsar_rocm.mas.

The reference's tests hold no MAS powder case; the checks are what the
mathematics fixes:
  - the Fourier components are rank-2 Wigner rotations: the CSA H_m of a
    tensor is T20 d2_m0(beta) e^{-i m alpha} Iz (plus delta_iso in H_0) and the
    dipolar ones 2 b d2_00, 2 b d2_10 e^{-i(a+g)}, b d2_20 e^{-2i(a+g)} times
    IzIz (the reference's own m = 2 factor), with d2 from the general Wigner
    formula; at the magic angle the secular anisotropic part vanishes;
  - the gradient of the propagator helper is 2 Re(conj(Phi) dTr(U)/de) / d^2
    for U_k -> U_k exp(-i e dt H_j) (central differences), and with U_t = I
    and no sidebands it is the gate oracle's reference gradient divided by d;
  - a powder gradient is the weighted sum of its orientations' gradients.
"""
from math import factorial

import numpy as np
import pytest
import scipy.linalg as sla

import oracle as O
from pulsar_rocm import mas as MAS
from pulsar_rocm.api import ControlSequence
from pulsar_rocm.problem import PR_GATE, Problem

TWO_PI = 2 * np.pi


def wigner_d2(mp, m, beta):
    j = 2
    s = 0.0
    for k in range(0, 2 * j + 1):
        if j + m - k < 0 or j - k - mp < 0 or k - m + mp < 0:
            continue
        num = (-1) ** (k - m + mp) * np.sqrt(factorial(j + m) * factorial(j - m) * factorial(j + mp) * factorial(j - mp))
        den = factorial(j + m - k) * factorial(k) * factorial(j - k - mp) * factorial(k - m + mp)
        s += num / den * np.cos(beta / 2) ** (2 * j - 2 * k + m - mp) * np.sin(beta / 2) ** (2 * k - m + mp)
    return s


CSA = [(TWO_PI * 1e3, TWO_PI * 8e3, 0.0, (0.3, 0.9, 0.1)), (-TWO_PI * 2e3, TWO_PI * 5e3, 0.0, (1.1, 0.4, 0.2))]
DIP = [(1, 2, TWO_PI * 15e3, (0.2, 1.2, 0.5))]


def test_fourier_components_are_wigner_rotations():
    Ix, Iy, Iz = O.spin_operators(["1H", "13C"])
    for ori in [(0.0, 0.0, 0.0), (0.4, 0.7, 1.9), (2.0, 2.5, 0.3), (0.5, 1.6, 0.0)]:
        Hf = O.mas_fourier(["1H", "13C"], CSA[:1], [], ori)
        d_iso, d_an, _, (a, b, g) = CSA[0]
        a, b = a + ori[0], b + ori[1]
        T20 = d_an * np.sqrt(2 / 3)
        assert np.abs(Hf[0] - (d_iso + T20 * wigner_d2(0, 0, b)) * Iz[0]).max() < 1e-9
        # sidebands only where sin(beta) > 1e-12, as the reference tests it
        # (a rotated beta beyond pi has none)
        side = np.sin(b) > 1e-12
        for m in (1, 2):
            want = T20 * wigner_d2(m, 0, b) * np.exp(-1j * m * a) * Iz[0] * side
            assert np.abs(Hf[m] - want).max() < 1e-9 and np.abs(Hf[-m] - want.conj().T).max() < 1e-9
        Hd = O.mas_fourier(["1H", "13C"], [], DIP, ori)
        _, _, bij, (a, b, g) = DIP[0]
        a, b, g = a + ori[0], b + ori[1], g + ori[2]
        ZZ = Iz[0] @ Iz[1]
        assert np.abs(Hd[0] - 2 * bij * wigner_d2(0, 0, b) * ZZ).max() < 1e-9
        side = np.sin(b) > 1e-12
        assert np.abs(Hd[1] - 2 * bij * wigner_d2(1, 0, b) * np.exp(-1j * (a + g)) * ZZ * side).max() < 1e-9
        assert np.abs(Hd[2] - bij * wigner_d2(2, 0, b) * np.exp(-2j * (a + g)) * ZZ * side).max() < 1e-9


def test_magic_angle_and_rotor_period():
    bm = np.arccos(1 / np.sqrt(3))
    Ix, Iy, Iz = O.spin_operators(["1H", "13C"])
    Hf = O.mas_fourier(["1H", "13C"], [(TWO_PI * 1e3, TWO_PI * 8e3, 0.0, (0.3, bm, 0.0))],
                       [(1, 2, TWO_PI * 15e3, (0.2, bm, 0.5))])
    assert np.abs(Hf[0] - TWO_PI * 1e3 * Iz[0]).max() < 1e-9
    # the drift at t and t + T_r agree, and its rotor-period mean is H_0
    s = MAS.mas_spin_system(MAS.mr_system(["1H", "13C"]), 10e3, [MAS.CSATensor(*CSA[0])], [MAS.DipolarCoupling(*DIP[0])])
    T = s.rotor_period_s
    assert np.abs(MAS.build_mas_hamiltonian(s, 0.3 * T) - MAS.build_mas_hamiltonian(s, 1.3 * T)).max() < 1e-6
    ts = (np.arange(64) + 0.5) * T / 64
    mean = sum(MAS.build_mas_hamiltonian(s, t) for t in ts) / 64
    assert np.abs(mean - s.H_fourier[0]).max() < 1e-6


def _case(nt=12, seed=0):
    rng = np.random.default_rng(seed)
    dt = 2e-6
    u = rng.uniform(-TWO_PI * 3e4, TWO_PI * 3e4, (4, nt))
    Ut = np.kron(np.array([[0, 1], [1, 0]]), np.eye(2)).astype(complex)
    return u, dt, Ut


def test_propagator_gradient_is_the_trace_derivative():
    u, dt, Ut = _case()
    Ix, Iy, Iz = O.spin_operators(["1H", "13C"])
    Hc = [Ix[0], Iy[0], Ix[1], Iy[1]]
    Hf = O.mas_fourier(["1H", "13C"], CSA, DIP, (0.3, 0.8, 0.2))
    Us = O.mas_propagators(Hf, Hc, 12e3, u, dt)
    G = O.propagator_gradient(Us, Hc, Ut, dt)
    D = 4

    def total(k, j, e):
        W = np.eye(D, dtype=complex)
        for n, U in enumerate(Us):
            W = (U @ sla.expm(-1j * e * dt * Hc[j]) if n == k else U) @ W
        return W

    UT = total(0, 0, 0.0)
    Phi = np.trace(Ut.conj().T @ UT) / D
    h = 100.0  # e dt ||H_j|| ~ 1e-4: truncation ~1e-8, rounding ~1e-12 relative
    for k, j in [(0, 0), (3, 1), (7, 2), (11, 3), (5, 0), (9, 1)]:
        dtr = (np.trace(total(k, j, h)) - np.trace(total(k, j, -h))) / (2 * h)
        want = 2 * np.real(np.conj(Phi) * dtr) / D ** 2
        assert abs(G[j, k] - want) <= 1e-6 * np.abs(G).max(), (k, j, G[j, k], want)


def test_propagator_gradient_is_the_gate_gradient_over_d_at_identity_target():
    u, dt, _ = _case(seed=1)
    Ix, Iy, Iz = O.spin_operators(["1H", "13C"])
    Hc = [Ix[0], Iy[0], Ix[1], Iy[1]]
    Hf = O.mas_fourier(["1H", "13C"], [(c[0], c[1], c[2], (c[3][0], 0.0, c[3][2])) for c in CSA],
                       [(1, 2, DIP[0][2], (0.2, 0.0, 0.5))])  # beta = 0: no sidebands
    assert all(np.abs(Hf[m]).max() == 0 for m in (-2, -1, 1, 2))
    Us = O.mas_propagators(Hf, Hc, 12e3, u, dt)
    G = O.propagator_gradient(Us, Hc, np.eye(4), dt)
    p = Problem(dim=4, n_ctrl=4, n_t=u.shape[1], H0=Hf[0], Hc=np.stack(Hc), dt=dt, objective=PR_GATE,
                U_target=np.eye(4, dtype=complex))
    _, Gg = O.evaluate(p, u)
    assert np.abs(G - Gg / 4).max() <= 1e-12 * np.abs(G).max()


def test_powder_gradient_is_the_weighted_orientation_sum():
    u, dt, Ut = _case(seed=2)
    ori = [(0.1, 0.5, 0.3, 0.25), (1.0, 1.3, 0.2, 0.75)]
    G = O.mas_powder_gradient(["1H", "13C"], CSA, DIP, 12e3, u, dt, Ut, ori)
    parts = [O.mas_powder_gradient(["1H", "13C"], CSA, DIP, 12e3, u, dt, Ut, [o[:3] + (1.0,)]) for o in ori]
    assert np.abs(G - (0.25 * parts[0] + 0.75 * parts[1])).max() <= 1e-15 * np.abs(G).max() * 10


def test_rotor_channels_reproduce_the_fourier_drift():
    # host logic of the device problem: H0 + sum s u A equals the MAS Hamiltonian
    u, dt, Ut = _case(seed=3)
    s = MAS.mas_spin_system(MAS.mr_system(["1H", "13C"]), 12e3, [MAS.CSATensor(*c) for c in CSA],
                            [MAS.DipolarCoupling(*d) for d in DIP])
    ctrl = ControlSequence.from_controls(u, dt, u.shape[1] * dt, u.shape[1])
    ori = [(0.1, 0.5, 0.3, 0.4), (1.0, 1.3, 0.2, 0.6)]
    p, ue = MAS.mas_powder_problem(s, ctrl, Ut, ori, rotor_phase_0=0.7)
    for i, o in enumerate(ori):
        r = MAS.rotate_spin_system(s, *o[:3])
        for k in range(u.shape[1]):
            t = (k + 0.5) * dt + 0.7 / s.omega_r
            H1 = MAS.build_mas_hamiltonian(r, t) + sum(u[j, k] * s.H_controls[j] for j in range(4))
            H2 = p.H0[i] + sum(p.ctrl_scale[i, j] * ue[j, k] * p.Hc[j] for j in range(p.n_ctrl))
            assert np.abs(H1 - H2).max() <= 1e-12 * np.abs(H1).max()


def test_state_branch_is_the_overlap_derivative_times_minus_i():
    # state branch: G = 2 Im(conj(ov) <Q_k psi_t| -i dt H |P_k psi_i>); the
    # bracket is d<psi_t|U|psi_i>/de for U_k -> U_k exp(-i e dt H_j) (central differences)
    u, dt, _ = _case(seed=4)
    Ix, Iy, Iz = O.spin_operators(["1H", "13C"])
    Hc = [Ix[0], Iy[0], Ix[1], Iy[1]]
    Hf = O.mas_fourier(["1H", "13C"], CSA, DIP, (0.3, 0.8, 0.2))
    Us = O.mas_propagators(Hf, Hc, 12e3, u, dt)
    rng = np.random.default_rng(12)
    pi = rng.standard_normal(4) + 1j * rng.standard_normal(4)  # unnormalised, as given
    pt = rng.standard_normal(4) + 1j * rng.standard_normal(4)
    G = O.propagator_gradient(Us, Hc, None, dt, pi, pt)

    def ov(k, j, e):
        W = np.eye(4, dtype=complex)
        for n, U in enumerate(Us):
            W = (U @ sla.expm(-1j * e * dt * Hc[j]) if n == k else U) @ W
        return np.vdot(pt, W @ pi)

    z = ov(0, 0, 0.0)
    h = 100.0
    for k, j in [(0, 0), (4, 1), (11, 3)]:
        dz = (ov(k, j, h) - ov(k, j, -h)) / (2 * h)
        assert abs(G[j, k] - 2 * np.imag(np.conj(z) * dz)) <= 1e-6 * np.abs(G).max()


def test_per_orientation_rotor_channels_reproduce_the_powder_gradient():
    # the construction of the Julia glue (rocm_hooks.jl): per orientation one
    # PR_GATE_PROPAGATORS problem with channels H_m + H_-m and i (H_m - H_-m)
    # driven by cos(m w_r t_k), sin(m w_r t_k); evaluated by the C oracle
    from pulsar_rocm.problem import PR_GATE_PROPAGATORS
    u, dt, Ut = _case(seed=6)
    w_r = TWO_PI * 12e3
    ori = [(0.1, 0.5, 0.3, 0.4), (1.0, 1.3, 0.2, 0.6)]
    Ix, Iy, Iz = O.spin_operators(["1H", "13C"])
    Hc = [Ix[0], Iy[0], Ix[1], Iy[1]]
    G = 0.0
    for a, b, g, wt in ori:
        Hf = O.mas_fourier(["1H", "13C"], CSA, DIP, (a, b, g))
        chans = Hc + [Hf[1] + Hf[-1], 1j * (Hf[1] - Hf[-1]), Hf[2] + Hf[-2], 1j * (Hf[2] - Hf[-2])]
        th = w_r * (np.arange(u.shape[1]) + 0.5) * dt
        ue = np.vstack([u, np.cos(th), np.sin(th), np.cos(2 * th), np.sin(2 * th)])
        p = Problem(dim=4, n_ctrl=len(chans), n_t=u.shape[1], H0=Hf[0], Hc=np.stack(chans), dt=dt,
                    objective=PR_GATE_PROPAGATORS, U_target=Ut)
        G = G + wt * O.evaluate(p, ue)[1][:4]
    G0 = O.mas_powder_gradient(["1H", "13C"], CSA, DIP, 12e3, u, dt, Ut, ori)
    assert np.abs(G - G0).max() <= 1e-12 * np.abs(G0).max()
