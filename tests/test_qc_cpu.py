"""Host logic of the quantum-computing callers (qc.py) on the CPU: the
quasi-static noise samples and ensemble order (QuasiStatic.jl:76-155), the
LindbladQCControl construction and its validation (LindbladQCControl.jl:
69-226), and the adapter's problem against the oracle's Lindblad evaluation
of the equivalent LindbladMRControl."""
import numpy as np
import pytest

import oracle as O
from pulsar_rocm import api as P, mr, qc, workloads as W
from pulsar_rocm.problem import PR_LINDBLAD, PR_PAIRS_SQUARE


def test_quasi_static_samples():
    g = qc.QuasiStaticNoise("freq", 100e3, n_samples=5)
    assert np.allclose(qc.qs_samples(g), 100e3 * np.array([-2, -1, 0, 1, 2.0]), rtol=0, atol=1e-9)
    assert np.array_equal(qc.qs_samples(qc.QuasiStaticNoise("freq", 3.0, n_samples=1)), [0.0])
    u = qc.qs_samples(qc.QuasiStaticNoise("amp", 0.1, distribution=":uniform", n_samples=3))
    assert np.allclose(u, [-0.1, 0.0, 0.1], rtol=0, atol=1e-15)
    e = qc.qs_samples(qc.QuasiStaticNoise("x", 1.0, samples=[0.5, -0.25]))
    assert np.array_equal(e, [0.5, -0.25])
    c = qc.qs_samples(qc.QuasiStaticNoise("x", 2.0, distribution="custom", n_samples=4), np.random.default_rng(1))
    assert c.shape == (4,) and np.allclose(c, 2.0 * np.random.default_rng(1).standard_normal(4))
    with pytest.raises(ValueError):
        qc.QuasiStaticNoise("x", 1.0, distribution="cauchy")
    with pytest.raises(ValueError):
        qc.QuasiStaticNoise("x", 1.0, n_samples=0)


def test_quasi_static_ensemble_is_the_cartesian_product_first_source_slowest():
    srcs = [qc.QuasiStaticNoise("a", 1.0, samples=[1.0, 2.0]),
            qc.QuasiStaticNoise("b", 1.0, samples=[10.0, 20.0, 30.0])]
    drifts = qc.quasi_static_ensemble(lambda p: np.diag([p[0], p[1]]).astype(complex), srcs)
    got = [(H[0, 0].real, H[1, 1].real) for H in drifts]
    assert got == [(a, b) for a in (1.0, 2.0) for b in (10.0, 20.0, 30.0)]


def test_robust_optimcon_qs_rejects_mismatched_power_levels():
    srcs = [qc.QuasiStaticNoise("f", 1.0, n_samples=3)]
    sysx = P.quantum_system(0.5 * W.SZ, [0.5 * W.SX])
    ctrl = P.random_controls(sysx, 1.0, 10, rng=np.random.default_rng(0))
    with pytest.raises(ValueError, match="same length"):
        qc.robust_optimcon_qs(lambda p: p[0] * W.SZ, [W.SX / 2], srcs, P.unitary_target(W.SX), ctrl,
                              pwr_levels=[1.0, 2.0])


def _ctx(**kw):
    sysx = P.quantum_system(2 * np.pi * 0.3 * W.SZ / 2, [2 * np.pi * W.SX / 2, 2 * np.pi * W.SY / 2])
    tgt = P.state_target(np.array([0, 1], complex))
    ctrl = P.random_controls(sysx, 2.0, 25, amplitude=0.4, rng=np.random.default_rng(3))
    sm = np.array([[0, 1], [0, 0]], complex)
    return qc.LindbladQCControl(sysx, tgt, ctrl, jump_ops=[sm], decay_rates=[0.05], **kw), sysx, tgt, ctrl, sm


def test_lindblad_qc_control_construction_and_validation():
    ctx, sysx, tgt, ctrl, sm = _ctx()
    assert ctx._hilbert_dim == 2 and ctx._liouville_dim == 4
    assert np.allclose(ctx._L_drift, mr.build_drift_liouvillian(sysx.H_drift, [sm], [0.05]))
    assert np.allclose(ctx._sigma_init, mr.pure_state_to_vec_rho([1, 0]))  # |0> by default
    assert np.allclose(ctx._sigma_targ, mr.pure_state_to_vec_rho([0, 1]))
    rho = np.array([[0.5, 0.5], [0.5, 0.5]], complex)
    c2 = qc.LindbladQCControl(sysx, tgt, ctrl, rho_init=mr.vec_rho(rho))
    assert np.allclose(c2._sigma_init, mr.vec_rho(rho))
    with pytest.raises(ValueError, match="length"):
        qc.LindbladQCControl(sysx, tgt, ctrl, jump_ops=[sm], decay_rates=[])
    with pytest.raises(ValueError, match="non-negative"):
        qc.LindbladQCControl(sysx, tgt, ctrl, jump_ops=[sm], decay_rates=[-1.0])
    with pytest.raises(ValueError, match="state-transfer"):
        qc.LindbladQCControl(sysx, P.unitary_target(W.SX), ctrl)
    with pytest.raises(ValueError, match="neither"):
        qc.qc_vectorise(np.ones(3), 2)
    with pytest.raises(ValueError, match="not part of this library"):
        qc.optimcon(qc.LindbladQCControl(sysx, tgt, ctrl, method="lbfgs"))


def test_lindblad_qc_problem_is_the_lindblad_mr_kernel_of_its_adapter():
    # the adapter presents one drift, power level 1, :square and uniform dt to
    # grape_lindblad_kernel: the same F and G as a LindbladMRControl built so
    ctx, sysx, tgt, ctrl, sm = _ctx()
    p = ctx.problem()
    assert p.dynamics == PR_LINDBLAD and p.objective == PR_PAIRS_SQUARE and p.n_members == 1
    mrc = mr.LindbladMRControl(drifts=[sysx.H_drift], operators=sysx.H_controls, rho_init=[ctx._sigma_init],
                               rho_targ=[ctx._sigma_targ], pulse_dt=np.full(25, ctrl.dt), jump_ops=[sm],
                               decay_rates=[0.05], fidelity="square")
    F0, G0 = O.evaluate(p, ctrl.controls)
    F1, G1 = O.evaluate(mrc.problem(), ctrl.controls)
    assert abs(F0 - F1) <= 1e-15 and np.abs(G0 - G1).max() <= 1e-15 * np.abs(G1).max()
    assert 0 < F0 < 1
