"""Host-side logic of the Python mirror that needs no device: gradient
verification rules and the ensemble L-BFGS loop on synthetic objectives."""
import numpy as np
import pytest

import pulsar_rocm as P
from pulsar_rocm.api import _ensemble_lbfgs, _lbfgs_direction


def test_verify_gradient_relative_and_absolute():
    a = np.array([[1.0, 2.0], [3.0, 4.0]])
    assert P.verify_gradient(a, a + 1e-6, tol=1e-5)
    # relative: error divided by max(1, |G_num|_inf) = 4
    assert P.verify_gradient(a, a + 3e-5, tol=1e-5)
    with pytest.warns(UserWarning):
        assert not P.verify_gradient(a, a + 3e-5, tol=1e-5, relative=False)
    with pytest.raises(ValueError):
        P.verify_gradient(a, a[:, :1])


def test_finite_difference_rejects_nonpositive_eps():
    sys_ = P.quantum_system(np.diag([1.0, -1.0]).astype(complex), [np.array([[0, 1], [1, 0]], complex)])
    ctrl = P.zero_controls(sys_, 1.0, 4)
    with pytest.raises(ValueError):
        P.finite_difference_gradient(sys_, ctrl, P.state_target([0, 1]), eps=0.0)


def test_autodiff_config_backends():
    assert P.select_autodiff_backend(None, None, P.AutoDiffConfig()) == "reverse"
    assert P.select_autodiff_backend(None, None, P.AutoDiffConfig(backend=":finite_diff")) == "finite_diff"
    with pytest.raises(ValueError):
        P.AutoDiffConfig(backend="zygote")


def test_lbfgs_direction_first_step_and_quadratic():
    g = np.array([1.0, -2.0, 0.5])
    assert np.array_equal(_lbfgs_direction(g, [], []), g)
    # the two-loop recursion applies the BFGS inverse-curvature matrix H, which
    # satisfies the secant equation of the newest pair: H y_k = s_k
    rng = np.random.default_rng(0)
    Q = np.linalg.qr(rng.standard_normal((3, 3)))[0]
    A = Q @ np.diag([1.0, 3.0, 10.0]) @ Q.T
    s_list = [rng.standard_normal(3) for _ in range(3)]
    y_list = [A @ s for s in s_list]
    assert np.allclose(_lbfgs_direction(y_list[-1], s_list, y_list), s_list[-1], rtol=1e-10, atol=1e-12)
    # one pair on a multiple of the identity gives the exact inverse: H = I / c
    s1 = [np.array([1.0, 2.0, -1.0])]
    d = _lbfgs_direction(g, s1, [4.0 * s1[0]])
    assert np.allclose(d, g / 4.0)


def test_ensemble_lbfgs_reaches_optimum_of_concave_objective():
    rng = np.random.default_rng(1)
    ustar = rng.uniform(-1, 1, (2, 6))
    W = rng.uniform(0.5, 2.0, (2, 6))

    def F(u):
        return 1.0 - 0.01 * float(np.sum(W * (u - ustar) ** 2))

    def G(u):
        return -0.02 * W * (u - ustar)

    cfg = P.GRAPEConfig(max_iter=200, step_size=0.05, convergence_tol=1e-14, gradient_norm_tol=1e-12)
    best, best_F, f_hist, g_hist, conv, reason, n_f, n_g = _ensemble_lbfgs(F, G, np.zeros((2, 6)), cfg)
    assert conv
    assert best_F >= 1.0 - 1e-8
    assert np.all(np.diff(f_hist) >= 0)  # the line search only accepts F_new >= F0
    assert n_g == 2 * (len(f_hist) - 1)


def test_ensemble_lbfgs_amplitude_limit_clips():
    def F(u):
        return -float(np.sum((u - 5.0) ** 2))

    def G(u):
        return -2.0 * (u - 5.0)

    cfg = P.GRAPEConfig(max_iter=20, step_size=0.5)
    best, *_ = _ensemble_lbfgs(F, G, np.zeros((1, 4)), cfg, amplitude_limit=1.0)
    assert np.all(np.abs(best) <= 1.0 + 1e-15)
    assert np.allclose(best, 1.0)
