"""The :rocm route of ensemble_map / ensemble_grad! / ensemble_fobj
(EnsembleMap.jl:365-607): member objectives of an ensemble of systems
(SystemBuilder.jl:61-77 closures) evaluated in one device plan, against the
oracle member by member."""
import numpy as np
import pytest

import oracle as O
import pulsar_rocm as P
from pulsar_rocm import native
from pulsar_rocm.api import fidelity_problem, gradient_problem
from pulsar_rocm.ensemble_map import MemberObjectives, ensemble_fobj, ensemble_grad, ensemble_map

pytestmark = pytest.mark.gpu

SX = np.array([[0, 1], [1, 0]], complex) / 2
SY = np.array([[0, -1j], [1j, 0]], complex) / 2
SZ = np.diag([0.5 + 0j, -0.5])
I2 = np.eye(2)


@pytest.fixture(scope="module", autouse=True)
def device():
    native.init(0)


def two_qubit_members(n_members, seed):
    """Two coupled qubits, x/y drives on both; members differ in detuning and in
    the amplitude scale of the drives (and member 0 in its control operators'
    phase, so the members do not all share one problem in the gradient plan)."""
    rng = np.random.default_rng(seed)
    systems = []
    for m in range(n_members):
        w = 2 * np.pi * 3.0 * (1 + 0.05 * rng.uniform(-1, 1, 2))
        H0 = w[0] * np.kron(SZ, I2) + w[1] * np.kron(I2, SZ) + 2 * np.pi * 0.7 * np.kron(SZ, SZ)
        s = 1 + 0.1 * rng.uniform(-1, 1)
        ops = [np.kron(SX, I2), np.kron(SY, I2), np.kron(I2, SX), np.kron(I2, SY)]
        if m == 0:
            ops = [ops[1], ops[0], ops[3], ops[2]]
        systems.append(P.quantum_system(H0, [s * o for o in ops]))
    return systems


@pytest.mark.parametrize("kind", ["gate", "state"])
def test_ensemble_map_and_grad_match_the_oracle_member_by_member(kind):
    M, n_t, dt = 7, 60, 0.02
    systems = two_qubit_members(M, seed=11)
    rng = np.random.default_rng(5)
    if kind == "gate":
        q, _ = np.linalg.qr(rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4)))
        tgt = P.unitary_target(q)
    else:  # unnormalised states: F uses them as given, the gradient normalises (a11 vs a12)
        tgt = P.state_target(np.array([0, 1.3, 0, 0.2j]), psi_init=np.array([0.9, 0, 0.1, 0]))
    u = rng.uniform(-6, 6, size=(4, n_t))
    theta = u.reshape(-1, order="F")
    ctrl = P.ControlSequence(u.T, dt, n_t)
    mo = MemberObjectives(systems, tgt, n_t, dt)
    F = ensemble_map(mo.fidelity_at(theta), range(M), device="rocm")
    F0 = np.array([O.evaluate(fidelity_problem(s, ctrl, tgt), u, grad=False)[0] for s in systems])
    assert np.abs(F - F0).max() <= 1e-13
    # a subset in another order, and the per-member call
    sub = [5, 0, 3]
    assert np.array_equal(ensemble_map(mo.fidelity_at(theta), sub, device="rocm"), F[sub])
    assert abs(mo.fidelity_at(theta)(2) - F[2]) <= 1e-15
    assert abs(mo.f_member(theta, 4) - F[4]) <= 1e-15

    G0 = np.array([O.evaluate(gradient_problem(s, ctrl, tgt), u)[1].reshape(-1, order="F") for s in systems])
    w = rng.uniform(0.2, 1.5, M)
    g = np.zeros(theta.size)
    ensemble_grad(g, theta, mo.grad_member, range(M), weight=w, device="rocm")
    ref = (w[:, None] * G0).sum(0) / w.sum()
    assert np.abs(g - ref).max() <= 1e-11 * np.abs(ref).max()
    # the host-threaded route over the same member functions (one device plan per member)
    g_host = np.zeros(theta.size)
    ensemble_grad(g_host, theta, mo.grad_member, range(M), weight=w, device="cpu")
    assert np.abs(g_host - ref).max() <= 1e-11 * np.abs(ref).max()
    gi = np.zeros(theta.size)
    mo.grad_member(gi, theta, 3)
    assert np.abs(gi - G0[3]).max() <= 1e-11 * np.abs(G0[3]).max()

    f, grad = ensemble_fobj(mo.f_member, mo.grad_member, range(M), device="rocm")
    assert f(theta) == pytest.approx(-F0.mean(), abs=1e-13)
    grad(g, theta)
    assert np.abs(g - G0.mean(0)).max() <= 1e-11 * np.abs(G0).max()
    mo.close()


def test_ensemble_grad_exact_mode_matches_finite_differences():
    M, n_t, dt = 3, 20, 0.03
    systems = two_qubit_members(M, seed=2)
    rng = np.random.default_rng(9)
    q, _ = np.linalg.qr(rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4)))
    tgt = P.unitary_target(q)
    theta = rng.uniform(-4, 4, size=4 * n_t)
    mo = MemberObjectives(systems, tgt, n_t, dt, grad_mode=P.PR_GRAD_EXACT)
    g = np.zeros(theta.size)
    ensemble_grad(g, theta, mo.grad_member, range(M), device="rocm")
    eps = 1e-6
    fd = np.empty(theta.size)
    for j in range(theta.size):
        tp, tm = theta.copy(), theta.copy()
        tp[j] += eps
        tm[j] -= eps
        fd[j] = (ensemble_map(mo.fidelity_at(tp), range(M), device="rocm").mean()
                 - ensemble_map(mo.fidelity_at(tm), range(M), device="rocm").mean()) / (2 * eps)
    assert np.abs(g - fd).max() <= 1e-7 * max(1.0, np.abs(fd).max())
    mo.close()
