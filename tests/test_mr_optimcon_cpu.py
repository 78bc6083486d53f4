"""Host logic of the MR optimcon loops (OptControl.jl:382-611) on the CPU:
the two-loop recursion equals the dense BFGS inverse-Hessian update applied
to the gradient, and the loops driven by oracle evaluations behave as the
reference's (ascent, clipping, stopping rules, evaluation counts)."""
import numpy as np
import pytest

import oracle as O
from pulsar_rocm import mr


def _dense_direction(g, s_list, y_list):
    n = g.size
    gamma = np.dot(s_list[-1], y_list[-1]) / np.dot(y_list[-1], y_list[-1])
    H = gamma * np.eye(n)
    for s, y in zip(s_list, y_list):
        rho = 1.0 / np.dot(y, s)
        V = np.eye(n) - rho * np.outer(y, s)
        H = V.T @ H @ V + rho * np.outer(s, s)
    return H @ g


@pytest.mark.parametrize("k", [1, 2, 5])
def test_two_loop_recursion_is_the_bfgs_inverse_update(k):
    rng = np.random.default_rng(k)
    n = 12
    A = rng.standard_normal((n, n))
    Hess = A @ A.T + n * np.eye(n)  # SPD curvature, so y.s > 0
    s_list = [rng.standard_normal(n) for _ in range(k)]
    y_list = [Hess @ s for s in s_list]
    g = rng.standard_normal(n)
    assert np.allclose(mr._lbfgs_direction(g, s_list, y_list), _dense_direction(g, s_list, y_list), rtol=1e-10)
    assert np.array_equal(mr._lbfgs_direction(g, [], []), g)


def _mr_ctrl(**kw):
    Iz, Ix, Iy = np.diag([0.5, -0.5]).astype(complex), np.array([[0, 0.5], [0.5, 0]], complex), \
        np.array([[0, -0.5j], [0.5j, 0]], complex)
    up, mIy = np.array([1, 0], complex), np.array([1, -1j], complex) / np.sqrt(2)
    return mr.MRControl(drifts=[2 * np.pi * 300 * Iz, -2 * np.pi * 200 * Iz], operators=[Ix, Iy], rho_init=[up],
                        rho_targ=[mIy], pulse_dt=np.full(40, 5e-6), pwr_levels=[2 * np.pi * 8e3, 2 * np.pi * 1e4],
                        **kw)


def _oracle_fg(ctrl):
    p = ctrl.problem()

    def fg(w):
        return O.evaluate(p, w)

    def f(w):
        return O.evaluate(p, w, grad=False)[0]
    return fg, f


def test_lbfgs_loop_with_oracle_evaluations():
    ctrl = _mr_ctrl(max_iter=15, l_bound=-0.8, u_bound=0.9)
    fg, f = _oracle_fg(ctrl)
    guess = 0.3 * np.random.default_rng(0).standard_normal((2, 40))
    guess[0, 0] = 5.0  # clipped to u_bound before the first evaluation
    r = mr.optimcon_lbfgs_loop(fg, f, guess, ctrl.l_bound, ctrl.u_bound, ctrl.max_iter, ctrl.grad_tol,
                               ctrl.lbfgs_memory)
    assert r.controls.max() <= 0.9 and r.controls.min() >= -0.8
    assert r.fidelity == max(r.fidelity_history) and r.fidelity > r.fidelity_history[0]
    assert len(r.fidelity_history) == r.n_iterations + 1 == len(r.gradient_norm_history)
    assert r.n_gradient_evaluations == r.n_iterations + 1
    assert r.n_fidelity_evaluations >= r.n_gradient_evaluations + 2 * r.n_iterations
    F0, G0 = fg(np.clip(guess, -0.8, 0.9))
    assert r.fidelity_history[0] == F0 and r.gradient_norm_history[0] == np.linalg.norm(G0)


def test_grape_loop_with_oracle_evaluations():
    ctrl = _mr_ctrl(max_iter=10, method="grape")
    fg, _ = _oracle_fg(ctrl)
    guess = 0.1 * np.random.default_rng(1).standard_normal((2, 40))
    r = mr.optimcon_grape_loop(fg, guess, ctrl.l_bound, ctrl.u_bound, ctrl.max_iter, ctrl.grad_tol)
    assert r.n_iterations == 10 and r.n_fidelity_evaluations == 11
    # first step: w + 0.05 (u - l) / max|G| * G
    F0, G0 = fg(guess)
    w1 = np.clip(guess + 0.1 / np.abs(G0).max() * G0, -1, 1)
    assert r.fidelity_history[1] == fg(w1)[0]


def test_optimcon_rejects_other_methods_and_bad_shapes():
    ctrl = _mr_ctrl(method="cmaes")
    with pytest.raises(ValueError, match="not part of this library"):
        mr.optimcon(ctrl, np.zeros((2, 40)))
    with pytest.raises(ValueError, match="rows"):
        mr.optimcon(_mr_ctrl(), np.zeros((3, 40)))


@pytest.mark.parametrize("nd,pw,pairs", [(7, [1.0, 1.2], 1), (5, [0.9, 1.0, 1.1], 2), (4, [1.0, 1.1], 2)])
def test_packed_members_equal_the_ensemble(nd, pw, pairs):
    # host logic of PackedMembers: block-diagonal packing of f members per
    # 16 x 16 block gives the ensemble's F and (summed channel rows) G
    Iz, Ix, Iy = np.diag([0.5, -0.5]).astype(complex), np.array([[0, 0.5], [0.5, 0]], complex), \
        np.array([[0, -0.5j], [0.5j, 0]], complex)
    up, dn, mIy = np.array([1, 0], complex), np.array([0, 1], complex), np.array([1, -1j], complex) / np.sqrt(2)
    ctrl = mr.MRControl(drifts=[2 * np.pi * o * Iz for o in np.linspace(-500, 500, nd)], operators=[Ix, Iy],
                        rho_init=[up, dn][:pairs], rho_targ=[mIy, up][:pairs], pulse_dt=np.full(30, 5e-6),
                        pwr_levels=[2 * np.pi * 8e3 * x for x in pw])
    p = ctrl.problem()
    u = np.random.default_rng(nd).uniform(-1, 1, (2, 30))
    assert mr.PackedMembers.applies(p)
    pk = mr.PackedMembers(p)
    assert p.n_members % pk.f == 0 and pk.f * p.dim <= 16 and pk.f * p.n_pairs <= 16 and pk.f >= 2
    F0, G0 = O.evaluate(p, u)
    F1, G1 = O.evaluate(pk.problem, pk.controls(u))
    assert abs(F0 - F1) < 1e-14 and np.abs(G0 - pk.gradient(G1)).max() <= 1e-13 * np.abs(G0).max()
