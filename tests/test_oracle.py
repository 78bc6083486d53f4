"""Pins the FP64 CPU oracle (CPU only).

Checks against the golden fixtures (scipy-built), the reference test suite's
closed forms (test/Physics/MathematicalCorrectness.jl, PhysicsValidation.jl)
and mathematical invariants chosen so that a dropped term, a wrong sign or
index, or a transposed operand fails at least one of them.
"""
import os

import numpy as np
import pytest

import oracle as O
from pulsar_rocm import workloads as W
from pulsar_rocm.problem import (PR_CLOSED, PR_GATE, PR_GRAD_EXACT, PR_LINDBLAD, PR_PAIRS_REAL,
                                 PR_PAIRS_SQUARE, Problem)

FIX = np.load(os.path.join(os.path.dirname(__file__), "golden", "fixtures.npz"))
SX = np.array([[0, 1], [1, 0]], complex)
SY = np.array([[0, -1j], [1j, 0]], complex)
SZ = np.diag([1.0 + 0j, -1.0])


@pytest.mark.parametrize("i", range(4))
def test_expm_neg_i_vs_scipy(i):
    U = O.expm_neg_i(FIX[f"herm_{i}_H"], float(FIX[f"herm_{i}_dt"]))
    assert np.abs(U - FIX[f"herm_{i}_U"]).max() < 1e-13


@pytest.mark.parametrize("i", range(6))
def test_expm_general_vs_scipy(i):
    E = O.expm(FIX[f"gen_{i}_A"])
    ref = FIX[f"gen_{i}_E"]
    assert np.abs(E - ref).max() / np.abs(ref).max() < 1e-12


@pytest.mark.parametrize("i", range(2))
def test_frechet_vs_scipy(i):
    L = O.expm_frechet(FIX[f"fre_{i}_X"], FIX[f"fre_{i}_E"])
    ref = FIX[f"fre_{i}_L"]
    assert np.abs(L - ref).max() / np.abs(ref).max() < 1e-12


def test_propagator_closed_forms():
    # MathematicalCorrectness.jl:34-61 — zero H, sigma_z rotation, unitarity
    assert np.abs(O.expm_neg_i(np.zeros((2, 2), complex), 1e-6) - np.eye(2)).max() < 1e-15
    U = O.expm_neg_i(np.pi * SZ, 1.0)
    assert abs(U[0, 0] - np.exp(-1j * np.pi)) < 1e-14 and abs(U[1, 1] - np.exp(1j * np.pi)) < 1e-14
    assert abs(U[0, 1]) < 1e-14
    rng = np.random.default_rng(1)
    for d in (2, 4, 16):
        a = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
        H = a + a.conj().T
        U = O.expm_neg_i(H, 0.3)
        assert np.abs(U.conj().T @ U - np.eye(d)).max() < 1e-13
        # det U = exp(-i dt Tr H)
        assert abs(np.linalg.det(U) - np.exp(-1j * 0.3 * np.trace(H))) < 1e-12
        # group property U(t1 + t2) = U(t2) U(t1)
        assert np.abs(O.expm_neg_i(H, 0.5) - O.expm_neg_i(H, 0.2) @ O.expm_neg_i(H, 0.3)).max() < 1e-13


def test_rabi_pi_pulse():
    # PhysicsValidation.jl:24-48: resonant drive Omega sx/2 for t: P(|1>) = sin^2(Omega t / 2)
    for Om, t in ((1.0, np.pi), (2.0, 0.4), (5.0, 1.1)):
        U = O.expm_neg_i(Om * SX / 2, t)
        assert abs(abs(U[1, 0]) ** 2 - np.sin(Om * t / 2) ** 2) < 1e-14


def test_eigh_reconstructs():
    rng = np.random.default_rng(2)
    for d in (3, 8, 16):
        a = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
        H = (a + a.conj().T) / 2
        w, V = O.eigh(H)
        assert np.abs(V @ np.diag(w) @ V.conj().T - H).max() < 1e-13
        assert np.abs(np.sort(w) - np.linalg.eigvalsh(H)).max() < 1e-13


def gate_fixture_problem():
    return Problem(dim=4, n_ctrl=3, n_t=5, H0=FIX["gate_H0"], Hc=FIX["gate_Hc"], dt=float(FIX["gate_dt"]),
                   objective=PR_GATE, n_members=2, ctrl_scale=FIX["gate_scale"], U_target=FIX["gate_Ut"])


def test_gate_fixture():
    p = gate_fixture_problem()
    F, G = O.evaluate(p, FIX["gate_u"])
    assert abs(F - float(FIX["gate_F"])) < 1e-14
    assert np.abs(G - FIX["gate_G"]).max() / np.abs(FIX["gate_G"]).max() < 1e-12
    # the reference's separate F and G closures give the same values
    F2, G2 = O.evaluate(p, FIX["gate_u"], separate_fg=True)
    assert F2 == F and np.array_equal(G2, G)


def test_pairs_fixture():
    p = Problem(dim=2, n_ctrl=2, n_t=4, H0=FIX["pair_H0"], Hc=FIX["pair_Hc"], dt_slices=FIX["pair_dts"],
                objective=PR_PAIRS_SQUARE, n_members=1, ctrl_scale=np.full((1, 2), float(FIX["pair_pwr"])),
                psi_init=FIX["pair_psi0"], psi_targ=FIX["pair_psit"], weight=np.ones(1))
    F, G = O.evaluate(p, FIX["pair_u"])
    assert abs(F - float(FIX["pair_F"])) < 1e-14
    assert np.abs(G - FIX["pair_G"]).max() < 1e-13


def test_lindblad_fixture():
    p = Problem(dim=4, n_ctrl=2, n_t=4, H0=FIX["lind_L0"], Hc=FIX["lind_Lc"], dt_slices=FIX["lind_dts"],
                dynamics=PR_LINDBLAD, objective=PR_PAIRS_SQUARE, n_members=1,
                ctrl_scale=np.full((1, 2), float(FIX["lind_pwr"])), psi_init=FIX["lind_s0"],
                psi_targ=FIX["lind_st"], weight=np.ones(1))
    F, G = O.evaluate(p, FIX["lind_u"])
    assert abs(F - float(FIX["lind_F"])) < 1e-14
    assert np.abs(G - FIX["lind_G"]).max() < 1e-13


def test_gate_gradient_single_slice_closed_form():
    # n_t = 1: P_1 = I and Q_1 = U_1^dag (Propagators.jl:375-380 builds Q[k] from
    # propagators[k], so Q_k^dag = U_n ... U_k includes slice k), hence
    # G_j = 2 Re(conj(Phi) (-i dt) Tr(Ut^dag U_1 H_j)/d). Leaving U_1 out, as the
    # docstring of compute_backward_propagators reads, fails here.
    rng = np.random.default_rng(5)
    d = 4
    H0 = W.random_hermitian(d, rng, 3.0)
    Hc = np.stack([W.random_hermitian(d, rng, 1.0) for _ in range(2)])
    Ut = W.haar_unitary(d, rng)
    u = rng.uniform(-1, 1, (2, 1))
    dt = 0.3
    p = Problem(dim=d, n_ctrl=2, n_t=1, H0=H0, Hc=Hc, dt=dt, objective=PR_GATE, U_target=Ut)
    F, G = O.evaluate(p, u)
    U1 = O.expm_neg_i(H0 + u[0, 0] * Hc[0] + u[1, 0] * Hc[1], dt)
    Phi = np.trace(Ut.conj().T @ U1) / d
    assert abs(F - abs(Phi) ** 2) < 1e-15
    for j in range(2):
        g = 2 * np.real(np.conj(Phi) * (-1j * dt) * np.trace(Ut.conj().T @ U1 @ Hc[j]) / d)
        assert abs(G[j, 0] - g) < 1e-15


def _fd(p, u, eps=1e-6):
    g = np.zeros_like(u)
    for j in range(u.shape[0]):
        for k in range(u.shape[1]):
            up, um = u.copy(), u.copy()
            up[j, k] += eps
            um[j, k] -= eps
            g[j, k] = (O.evaluate(p, up, grad=False)[0] - O.evaluate(p, um, grad=False)[0]) / (2 * eps)
    return g


def _small_gate(dt, seed=9):
    rng = np.random.default_rng(seed)
    d = 3
    H0 = W.random_hermitian(d, rng, 2.0)
    Hc = np.stack([W.random_hermitian(d, rng, 1.0) for _ in range(2)])
    Ut = W.haar_unitary(d, rng)
    u = rng.uniform(-1, 1, (2, 4))
    return Problem(dim=d, n_ctrl=2, n_t=4, H0=H0, Hc=Hc, dt=dt, objective=PR_GATE, U_target=Ut), u


def test_exact_gradient_matches_finite_differences():
    for dt in (0.05, 0.4):
        p, u = _small_gate(dt)
        _, Gx = O.evaluate(p, u, grad_mode=PR_GRAD_EXACT)
        Gfd = _fd(p, u)
        assert np.abs(Gx - Gfd).max() / np.abs(Gfd).max() < 1e-7


def test_reference_gradient_is_first_order():
    # the reference formula drops U_k: its relative error vs the true
    # derivative is O(dt ||H||) — halving dt halves it (AutoDiffTests.jl:33-41
    # checks the tiny-dt regime; MathematicalCorrectness.jl:148 allows 1e-2)
    errs = []
    for dt in (2e-3, 1e-3, 5e-4):
        p, u = _small_gate(dt)
        _, Gr = O.evaluate(p, u)
        _, Gx = O.evaluate(p, u, grad_mode=PR_GRAD_EXACT)
        errs.append(np.abs(Gr - Gx).max() / np.abs(Gx).max())
    assert errs[0] < 5e-2
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 1.85 < r1 < 2.15 and 1.85 < r2 < 2.15, errs


def test_pairs_and_lindblad_exact_vs_fd():
    rng = np.random.default_rng(4)
    d = 3
    H0 = W.random_hermitian(d, rng, 2.0)
    Hc = np.stack([W.random_hermitian(d, rng, 1.0) for _ in range(2)])
    psi0 = (rng.standard_normal((2, d)) + 1j * rng.standard_normal((2, d)))
    psit = (rng.standard_normal((2, d)) + 1j * rng.standard_normal((2, d)))
    u = rng.uniform(-1, 1, (2, 3))
    for obj in (PR_PAIRS_SQUARE, PR_PAIRS_REAL):
        p = Problem(dim=d, n_ctrl=2, n_t=3, H0=H0, Hc=Hc, dt=0.2, objective=obj, psi_init=psi0, psi_targ=psit)
        _, Gx = O.evaluate(p, u, grad_mode=PR_GRAD_EXACT)
        assert np.abs(Gx - _fd(p, u)).max() / np.abs(Gx).max() < 1e-7
    L0 = FIX["lind_L0"]
    Lc = FIX["lind_Lc"]
    p = Problem(dim=4, n_ctrl=2, n_t=4, H0=L0, Hc=Lc, dt=0.2, dynamics=PR_LINDBLAD, objective=PR_PAIRS_SQUARE,
                psi_init=FIX["lind_s0"], psi_targ=FIX["lind_st"])
    u = FIX["lind_u"]
    _, Gx = O.evaluate(p, u, grad_mode=PR_GRAD_EXACT)
    assert np.abs(Gx - _fd(p, u)).max() / np.abs(Gx).max() < 1e-7


def test_zero_rate_lindblad_equals_closed_system():
    # PhysicsValidation.jl:312-351: with no jump operators exp(L t) vec(rho) = vec(U rho U^dag),
    # so Re<vec(sigma_t)|Phi vec(rho_0)> = |<t|U|i>|^2 and the exact gradients agree
    rng = np.random.default_rng(6)
    N = 3
    H = W.random_hermitian(N, rng, 2.0)
    Hc = np.stack([W.random_hermitian(N, rng, 1.0) for _ in range(2)])
    psi0 = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    psit = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    psi0 /= np.linalg.norm(psi0)
    psit /= np.linalg.norm(psit)
    u = rng.uniform(-1, 1, (2, 5))
    pc = Problem(dim=N, n_ctrl=2, n_t=5, H0=H, Hc=Hc, dt=0.15, objective=PR_PAIRS_SQUARE, psi_init=psi0[None],
                 psi_targ=psit[None], weight=np.ones(1))
    rho0 = np.outer(psi0, psi0.conj()).reshape(-1, order="F")
    sigt = np.outer(psit, psit.conj()).reshape(-1, order="F")
    pl = Problem(dim=N * N, n_ctrl=2, n_t=5, H0=W.drift_liouvillian(H, [], []),
                 Hc=np.stack([W.control_liouvillian(h) for h in Hc]), dt=0.15, dynamics=PR_LINDBLAD,
                 objective=PR_PAIRS_REAL, psi_init=rho0[None], psi_targ=sigt[None], weight=np.ones(1))
    Fc, Gc = O.evaluate(pc, u, grad_mode=PR_GRAD_EXACT)
    Fl, Gl = O.evaluate(pl, u, grad_mode=PR_GRAD_EXACT)
    assert abs(Fc - Fl) < 1e-13
    assert np.abs(Gc - Gl).max() < 1e-12


def test_amplitude_damping_decay():
    # excited population decays as exp(-gamma t) under L = sqrt(gamma) sigma_-
    Lm = np.array([[0, 1], [0, 0]], complex)  # |0><1| with |1> = excited (index 1)
    g = 0.7
    L0 = W.drift_liouvillian(np.zeros((2, 2), complex), [Lm], [g])
    rho = np.array([[0, 0], [0, 1]], complex).reshape(-1, order="F")
    for t in (0.1, 1.0, 3.0):
        r = (O.expm(L0 * t) @ rho).reshape(2, 2, order="F")
        assert abs(r[1, 1].real - np.exp(-g * t)) < 1e-14
        assert abs(np.trace(r) - 1) < 1e-14


def test_liouvillian_trace_preserving():
    p, _ = W.three_qubit_lindblad(n_t=2, n_members=1)
    vecI = np.eye(8).reshape(-1, order="F")
    assert np.abs(vecI @ p.H0).max() < 1e-6 * np.abs(p.H0).max()
    for L in p.Hc:
        assert np.abs(vecI @ L).max() < 1e-12 * max(1, np.abs(L).max())


def test_ensemble_weighting_is_member_mean():
    # EnsembleObjective.jl:169-217: mean of the per-member F and G
    p, u = W.four_qubit_ensemble(n_t=6, n_det=2, n_amp=2)
    F, G, Fm, Gm = O.evaluate(p, u, per_member=True)
    assert abs(F - Fm.mean()) < 1e-16
    assert np.abs(G - Gm.mean(axis=0)).max() < 1e-16 + 1e-14 * np.abs(G).max()
    for m in range(p.n_members):
        q, uu = W.shrink(p, u)
        q.n_members = 1
        q.H0 = p.H0[m]
        q.ctrl_scale = p.ctrl_scale[m:m + 1]
        F1, G1 = O.evaluate(q, uu)
        assert F1 == Fm[m] and np.array_equal(G1, Gm[m])


def test_state_transfer_zero_and_one():
    # orthogonal target gives F = 0 at U = I; the target itself gives F = 1
    p = Problem(dim=2, n_ctrl=1, n_t=1, H0=np.zeros((2, 2), complex), Hc=np.stack([SX]), dt=1.0,
                objective=PR_PAIRS_SQUARE, psi_init=np.array([[1, 0]], complex),
                psi_targ=np.array([[0, 1]], complex), weight=np.ones(1))
    assert O.evaluate(p, np.zeros((1, 1)), grad=False)[0] == 0.0
    F, _ = O.evaluate(p, np.full((1, 1), np.pi), grad=False)  # exp(-i pi sx) = -I
    assert abs(F) < 1e-28
    F, _ = O.evaluate(p, np.full((1, 1), np.pi / 2), grad=False)  # -i sx: |0> -> -i|1>
    assert abs(F - 1) < 1e-15


def test_gate_fidelity_bounds_and_phase():
    rng = np.random.default_rng(8)
    d = 4
    Ut = W.haar_unitary(d, rng)
    H = 1j * np.zeros((d, d))
    p = Problem(dim=d, n_ctrl=1, n_t=1, H0=H, Hc=np.stack([np.eye(d) + 0j]), dt=1.0, objective=PR_GATE,
                U_target=Ut)
    # U = exp(-i u I) = e^{-iu} I: F = |Tr(Ut)/d|^2 for every u (global phase invariance)
    ref = abs(np.trace(Ut) / d) ** 2
    for uu in (0.0, 0.7, 2.5):
        F, _ = O.evaluate(p, np.full((1, 1), uu), grad=False)
        assert abs(F - ref) < 1e-14 and 0 <= F <= 1


def test_gate_gradient_follows_the_backward_propagator_loop():
    # the reference's loop, restated literally on n_t = 3: Q[n+1] = I,
    # Q[k] = U_k^dag Q[k+1]; A_k = P_k Ut^dag Q_k^dag; G = 2 Re(conj(Phi) (-i dt) Tr(A_k H_j)/d)
    rng = np.random.default_rng(8)
    d, nt, dt = 3, 3, 0.2
    H0 = W.random_hermitian(d, rng, 2.0)
    Hc = np.stack([W.random_hermitian(d, rng, 1.0) for _ in range(2)])
    Ut = W.haar_unitary(d, rng)
    u = rng.uniform(-1, 1, (2, nt))
    p = Problem(dim=d, n_ctrl=2, n_t=nt, H0=H0, Hc=Hc, dt=dt, objective=PR_GATE, U_target=Ut)
    _, G = O.evaluate(p, u)
    Us = [O.expm_neg_i(H0 + u[0, k] * Hc[0] + u[1, k] * Hc[1], dt) for k in range(nt)]
    P = [np.eye(d)]
    for k in range(nt):
        P.append(Us[k] @ P[k])
    Q = [None] * (nt + 1)
    Q[nt] = np.eye(d)
    for k in range(nt - 1, -1, -1):
        Q[k] = Us[k].conj().T @ Q[k + 1]
    Phi = np.trace(Ut.conj().T @ P[nt]) / d
    for k in range(nt):
        A = P[k] @ Ut.conj().T @ Q[k].conj().T
        for j in range(2):
            g = 2 * np.real(np.conj(Phi) * (-1j * dt) * np.trace(A @ Hc[j]) / d)
            assert abs(G[j, k] - g) < 1e-14


def test_state_gradient_follows_gradients_jl():
    # PR_STATE: _state_gradient_into! (Gradients.jl:357-399) with lambda_k = Q_k psi_t,
    # slice k included; n_t = 2 restated literally
    from pulsar_rocm.problem import PR_STATE
    rng = np.random.default_rng(9)
    d, nt, dt = 4, 2, 0.3
    H0 = W.random_hermitian(d, rng, 2.0)
    Hc = np.stack([W.random_hermitian(d, rng, 1.0) for _ in range(2)])
    psi_i = W.random_state(d, rng)
    psi_t = W.random_state(d, rng)
    u = rng.uniform(-1, 1, (2, nt))
    p = Problem(dim=d, n_ctrl=2, n_t=nt, H0=H0, Hc=Hc, dt=dt, objective=PR_STATE,
                psi_init=psi_i[None], psi_targ=psi_t[None])
    F, G = O.evaluate(p, u)
    Us = [O.expm_neg_i(H0 + u[0, k] * Hc[0] + u[1, k] * Hc[1], dt) for k in range(nt)]
    Ut = Us[1] @ Us[0]
    chi = psi_t.conj() @ Ut @ psi_i
    assert abs(F - abs(chi) ** 2) < 1e-14
    Pk = [np.eye(d), Us[0]]
    Qk = [Us[0].conj().T @ Us[1].conj().T, Us[1].conj().T]
    for k in range(nt):
        lam, phi = Qk[k] @ psi_t, Pk[k] @ psi_i
        for j in range(2):
            g = 2 * np.real(np.conj(chi) * (-1j * dt) * (lam.conj() @ Hc[j] @ phi))
            assert abs(G[j, k] - g) < 1e-14


# ---------------------------------------------------------------------------
# Independent pins of the state, state-pair and Lindblad gradient conventions.
# Each reference-mode gradient is checked three ways, none of which restates the
# formula it checks:
#   * it converges to the exact gradient (itself pinned by central differences
#     in test_pairs_and_lindblad_exact_vs_fd) at first order in dt: the error
#     halves when dt halves, so a dropped factor 2, a wrong sign or a missing
#     term (whose error would not vanish with dt) fails;
#   * with n_t = 1 the chains are trivial, which fixes where lambda sits:
#     GRAPEState.jl:218-229 pairs lambda[n+1] = psi_t with psi[n] = psi_0, while
#     _state_gradient_into! (Gradients.jl:357-399) uses lambda_1 = U_1^dag psi_t;
#   * the objectives are tied to one another: F_square = Re(z)^2 + Im(z)^2 with
#     Re(z) and Im(z) two :real objectives, checked on F and its exact gradient.


def _pair_problem(obj, dt, nt, seed=21, dynamics=PR_CLOSED, n_pairs=2):
    rng = np.random.default_rng(seed)
    if dynamics == PR_CLOSED:
        d = 3
        H0 = W.random_hermitian(d, rng, 2.0)
        Hc = np.stack([W.random_hermitian(d, rng, 1.0) for _ in range(2)])
        psi0 = np.stack([W.random_state(d, rng) for _ in range(n_pairs)])
        psit = np.stack([W.random_state(d, rng) for _ in range(n_pairs)])
    else:
        N, d = 2, 4
        Hs = W.random_hermitian(N, rng, 2.0)
        Lm = np.array([[0, 1], [0, 0]], complex)
        H0 = W.drift_liouvillian(Hs, [Lm], [0.4])
        Hc = np.stack([W.control_liouvillian(W.random_hermitian(N, rng, 1.0)) for _ in range(2)])
        vs = []
        for _ in range(2 * n_pairs):
            v = W.random_state(N, rng)
            vs.append(np.outer(v, v.conj()).reshape(-1, order="F"))
        psi0, psit = np.stack(vs[:n_pairs]), np.stack(vs[n_pairs:])
    u = rng.uniform(-1, 1, (2, nt))
    p = Problem(dim=d, n_ctrl=2, n_t=nt, H0=H0, Hc=Hc, dt=dt, dynamics=dynamics, objective=obj,
                psi_init=psi0, psi_targ=psit, ctrl_scale=np.array([[0.8, 1.3]]), weight=np.ones(1))
    return p, u


def _objectives():
    from pulsar_rocm.problem import PR_STATE
    return [("state", PR_STATE, PR_CLOSED), ("square", PR_PAIRS_SQUARE, PR_CLOSED),
            ("real", PR_PAIRS_REAL, PR_CLOSED), ("lindblad_square", PR_PAIRS_SQUARE, PR_LINDBLAD),
            ("lindblad_real", PR_PAIRS_REAL, PR_LINDBLAD)]


@pytest.mark.parametrize("name,obj,dyn", _objectives())
def test_pair_reference_gradients_are_first_order(name, obj, dyn):
    # the exact gradients of these objectives match central differences
    p, u = _pair_problem(obj, 0.2, 3, dynamics=dyn)
    _, Gx = O.evaluate(p, u, grad_mode=PR_GRAD_EXACT)
    assert np.abs(Gx - _fd(p, u)).max() / np.abs(Gx).max() < 1e-7
    errs = []
    for dt in (2e-3, 1e-3, 5e-4):
        p, u = _pair_problem(obj, dt, 4, dynamics=dyn)
        _, Gr = O.evaluate(p, u)
        _, Gx = O.evaluate(p, u, grad_mode=PR_GRAD_EXACT)
        errs.append(np.abs(Gr - Gx).max() / np.abs(Gx).max())
    assert errs[0] < 5e-2, errs
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 1.85 < r1 < 2.15 and 1.85 < r2 < 2.15, (name, errs)


@pytest.mark.parametrize("name,obj,dyn", _objectives())
def test_pair_gradients_single_slice_closed_form(name, obj, dyn):
    from pulsar_rocm.problem import PR_STATE
    p, u = _pair_problem(obj, 0.35, 1, dynamics=dyn)
    F, G = O.evaluate(p, u)
    s = p.ctrl_scale[0]
    gen = p.H0 + s[0] * u[0, 0] * p.Hc[0] + s[1] * u[1, 0] * p.Hc[1]
    U1 = O.expm_neg_i(gen, p.dt) if dyn == PR_CLOSED else O.expm(gen * p.dt)
    Fref, Gref = 0.0, np.zeros(2)
    for a, b in zip(p.psi_init, p.psi_targ):
        z = b.conj() @ U1 @ a
        sq = obj in (PR_PAIRS_SQUARE, PR_STATE)
        Fref += abs(z) ** 2 if sq else z.real
        for j in range(2):
            if obj == PR_STATE:
                # lambda_1 = Q_1 psi_t = U_1^dag psi_t, psi_1 = psi_0
                inner = (U1.conj().T @ b).conj() @ (s[j] * p.Hc[j]) @ a
                g = 2 * np.real(np.conj(z) * (-1j * p.dt) * inner)
            else:
                # lambda[2] = psi_t, psi[1] = psi_0
                inner = b.conj() @ (s[j] * p.Hc[j]) @ a
                if dyn == PR_CLOSED:
                    g = 2 * p.dt * np.imag(np.conj(z) * inner) if sq else p.dt * np.imag(inner)
                else:
                    g = 2 * p.dt * np.real(np.conj(z) * inner) if sq else p.dt * np.real(inner)
            Gref[j] += g
    assert abs(F - Fref) < 1e-14
    assert np.abs(G[:, 0] - Gref).max() < 1e-14 * max(1.0, np.abs(Gref).max())


def test_square_and_real_objectives_are_tied():
    # one pair: F_square = |z|^2 and F_real = Re z of the same overlap; the
    # exact gradients satisfy dF_square = 2 Re(conj z dz) where dF_real = Re dz
    # and the derivative of Im z is the :real gradient of the pair (psi_0, i psi_t)
    for dyn in (PR_CLOSED, PR_LINDBLAD):
        ps, u = _pair_problem(PR_PAIRS_SQUARE, 0.2, 3, dynamics=dyn, n_pairs=1)
        pr, _ = _pair_problem(PR_PAIRS_REAL, 0.2, 3, dynamics=dyn, n_pairs=1)
        pi, _ = _pair_problem(PR_PAIRS_REAL, 0.2, 3, dynamics=dyn, n_pairs=1)
        pi.psi_targ = 1j * pi.psi_targ  # Re <i t|U|i> = Im <t|U|i>
        Fs, Gs = O.evaluate(ps, u, grad_mode=PR_GRAD_EXACT)
        Fr, Gr = O.evaluate(pr, u, grad_mode=PR_GRAD_EXACT)
        Fi, Gi = O.evaluate(pi, u, grad_mode=PR_GRAD_EXACT)
        assert abs(Fs - (Fr ** 2 + Fi ** 2)) < 1e-14
        assert np.abs(Gs - 2 * (Fr * Gr + Fi * Gi)).max() < 1e-13 * np.abs(Gs).max()


def test_modulus_objective():
    # :modulus (Objectives.jl:222, 356-361): F = |z|; the reference prefactor
    # dt Im(inner / |z|) is the :real one divided by |z| (per pair), and the
    # exact mode is the derivative of |z| (central differences)
    from pulsar_rocm.problem import PR_PAIRS_MODULUS
    pm, u = _pair_problem(PR_PAIRS_MODULUS, 0.3, 3, n_pairs=1)
    ps, _ = _pair_problem(PR_PAIRS_SQUARE, 0.3, 3, n_pairs=1)
    pr, _ = _pair_problem(PR_PAIRS_REAL, 0.3, 3, n_pairs=1)
    Fm, Gm = O.evaluate(pm, u)
    Fs, _ = O.evaluate(ps, u)
    _, Gr = O.evaluate(pr, u)
    assert abs(Fm - np.sqrt(Fs)) < 1e-15
    assert np.abs(Gm * Fm - Gr).max() < 1e-14 * np.abs(Gr).max()
    _, Gx = O.evaluate(pm, u, grad_mode=PR_GRAD_EXACT)
    assert np.abs(Gx - _fd(pm, u)).max() / np.abs(Gx).max() < 1e-7
    # several pairs: the sum of the single-pair objectives
    p2, u2 = _pair_problem(PR_PAIRS_MODULUS, 0.3, 3, n_pairs=2)
    F2, G2 = O.evaluate(p2, u2)
    Fsum, Gsum = 0.0, 0.0
    for s in range(2):
        q, _ = _pair_problem(PR_PAIRS_MODULUS, 0.3, 3, n_pairs=2)
        q.psi_init, q.psi_targ = q.psi_init[s:s + 1], q.psi_targ[s:s + 1]
        Fq, Gq = O.evaluate(q, u2)
        Fsum, Gsum = Fsum + Fq, Gsum + Gq
    assert abs(F2 - Fsum) < 1e-14 and np.abs(G2 - Gsum).max() < 1e-14 * np.abs(Gsum).max()
    # an overlap below 1e-14 contributes no gradient
    q, _ = _pair_problem(PR_PAIRS_MODULUS, 0.3, 1, n_pairs=1)
    q.H0 = np.zeros_like(q.H0)
    q.psi_init = np.array([[1, 0, 0]], complex)
    q.psi_targ = np.array([[0, 1, 0]], complex)
    Fz, Gz = O.evaluate(q, np.zeros((2, 1)))
    assert Fz == 0.0 and np.all(Gz == 0.0)


def test_optimized_control_fixture_is_near_the_optimum():
    # tests/golden/c2_optimized.npz, the point where the reduced-precision
    # forms are checked near F = 1: the oracle on a spread member sample at
    # full length reaches F > 0.999, and the target is the nominal system's
    # propagator (unit fidelity for the nominal member at the reference controls)
    fx = np.load(os.path.join(os.path.dirname(__file__), "golden", "c2_optimized.npz"))
    nom, u_ref = W.four_qubit_nominal()
    assert np.allclose(O.total_propagator(nom, u_ref), fx["target"], atol=1e-12)
    p, _ = W.four_qubit_ensemble()
    p.U_target = fx["target"]
    pick = np.array([0, 85, 170, 255])
    p.n_members = len(pick)
    p.H0 = np.asarray(p.H0)[pick]
    p.ctrl_scale = np.asarray(p.ctrl_scale)[pick]
    F, _ = O.evaluate(p, fx["controls"])
    assert F > 0.999 and F <= 1.0
