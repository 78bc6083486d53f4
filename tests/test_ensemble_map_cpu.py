"""ensemble_map / ensemble_grad! / ensemble_fobj (EnsembleMap.jl:365-607) without a
device: host-thread mapping of plain closures, weights, signs, errors, and the
:rocm route's dispatch to a member function's batched form (a stand-in object
here; the library's own MemberObjectives run in tests/test_gpu_ensemble_map.py)."""
import logging

import numpy as np
import pytest

from pulsar_rocm import api
from pulsar_rocm.ensemble_map import MemberObjectives, ensemble_fobj, ensemble_grad, ensemble_map


@pytest.fixture(autouse=True)
def reset_device():
    prev = api._device_default
    yield
    api._device_default = prev
    if hasattr(api._device_local, "device"):
        del api._device_local.device


def test_map_keeps_index_order_and_converts_to_float():
    vals = ensemble_map(lambda i: i * i, range(37), device="cpu")
    assert vals.dtype == np.float64
    assert np.array_equal(vals, np.arange(37.0) ** 2)
    assert ensemble_map(lambda i: 1, [], device="cpu").shape == (0,)


def test_map_follows_the_task_device_and_rejects_unknown_devices():
    with api.with_device("metal"):  # not loaded here: CPU threads, as the reference
        assert np.array_equal(ensemble_map(lambda i: -i, [3, 1, 2]), [-3.0, -1.0, -2.0])
    with pytest.raises(ValueError, match="unknown device"):
        ensemble_map(lambda i: i, range(2), device="tpu")
    with pytest.raises(ValueError, match="unknown device"):
        ensemble_grad(np.zeros(2), np.zeros(2), lambda g, t, i: None, range(2), device=":tpu")


def test_grad_is_the_weighted_member_mean():
    rng = np.random.default_rng(3)
    members = rng.normal(size=(9, 5))
    w = rng.uniform(0.1, 2.0, 9)

    def gm(g, theta, i):
        g[:] = members[i] * theta

    theta = rng.normal(size=5)
    g = np.full(5, np.nan)
    out = ensemble_grad(g, theta, gm, range(9), device="cpu")
    assert out is g
    assert np.allclose(g, members.mean(axis=0) * theta, rtol=0, atol=1e-15)
    ensemble_grad(g, theta, gm, range(9), weight=w, device="cpu")
    assert np.allclose(g, (w[:, None] * members).sum(0) / w.sum() * theta, rtol=0, atol=1e-14)
    with pytest.raises(ValueError, match="weight length"):
        ensemble_grad(g, theta, gm, range(9), weight=w[:4], device="cpu")


def test_fobj_negates_the_mean_and_keeps_the_gradient_sign():
    # EnsembleMap.jl:589-606: f = -mean(F_i) (-dot(w, F)/sum(w)), grad! = +weighted mean
    F = np.array([0.2, 0.5, 0.9])
    f, grad = ensemble_fobj(lambda th, i: F[i] * th[0], lambda g, th, i: g.__setitem__(0, F[i]), range(3),
                            device="cpu")
    assert f(np.array([2.0])) == pytest.approx(-2.0 * F.mean(), abs=1e-15)
    g = np.zeros(1)
    grad(g, np.array([2.0]))
    assert g[0] == pytest.approx(F.mean(), abs=1e-15)
    w = np.array([1.0, 0.0, 3.0])
    f, grad = ensemble_fobj(lambda th, i: F[i], lambda g, th, i: g.__setitem__(0, F[i]), range(3), weight=w,
                            device="cpu")
    assert f(np.zeros(1)) == pytest.approx(-(F @ w) / w.sum(), abs=1e-15)
    grad(g, np.zeros(1))
    assert g[0] == pytest.approx((F @ w) / w.sum(), abs=1e-15)


class _Batched:
    """A member function with a device description: the :rocm route must call its
    batched form once for all members instead of the per-member call."""

    def __init__(self):
        self.calls = []

    def __call__(self, i):
        raise AssertionError("the :rocm route must not call members one by one")

    def rocm_map(self, indices):
        self.calls.append(list(indices))
        return [10.0 * i for i in indices]


class _BatchedGrad:
    def __init__(self):
        self.weights = None

    def __call__(self, g, theta, i):
        raise AssertionError("the :rocm route must not call members one by one")

    def rocm_grad(self, theta, indices, weights):
        self.weights = np.array(weights)
        return np.full(theta.size, float(np.dot(weights, indices)))


def test_rocm_route_uses_the_batched_form_once(caplog):
    f = _Batched()
    assert np.array_equal(ensemble_map(f, [4, 2, 7], device="rocm"), [40.0, 20.0, 70.0])
    assert f.calls == [[4, 2, 7]]
    gm = _BatchedGrad()
    g = np.zeros(3)
    ensemble_grad(g, np.ones(3), gm, [1, 2, 3], weight=[1.0, 1.0, 2.0], device="rocm")
    assert np.allclose(gm.weights, [0.25, 0.25, 0.5])  # w / sum(w), reduced by the batched form
    assert np.allclose(g, 0.25 * 1 + 0.25 * 2 + 0.5 * 3)
    # a plain closure under :rocm is mapped on host threads with an informational message
    with caplog.at_level(logging.INFO, logger="pulsar_rocm"):
        assert np.array_equal(ensemble_map(lambda i: i + 0.5, range(3), device="rocm"), [0.5, 1.5, 2.5])


def test_member_objectives_validate_their_inputs():
    sx = np.array([[0, 1], [1, 0]], complex) / 2
    sz = np.diag([0.5 + 0j, -0.5])
    s1 = api.quantum_system(sz, [sx])
    s2 = api.quantum_system(sz, [sx, sx])
    tgt = api.unitary_target(np.eye(2))
    with pytest.raises(ValueError, match="non-empty"):
        MemberObjectives([], tgt, 10, 0.1)
    with pytest.raises(ValueError, match="same number of control fields"):
        MemberObjectives([s1, s2], tgt, 10, 0.1)
    mo = MemberObjectives([s1, s1], tgt, 10, 0.1)
    with pytest.raises(IndexError):
        mo.rocm_values(np.zeros(10), [2])
    with pytest.raises(ValueError, match="theta has 9 entries"):
        mo.rocm_values(np.zeros(9), [0])
