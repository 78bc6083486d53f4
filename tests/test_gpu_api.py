"""The Pulsar-shaped host API running on the device, against the oracle."""
import numpy as np
import pytest

import oracle as O
import pulsar_rocm as P
from pulsar_rocm import native, workloads as W
from pulsar_rocm.api import fidelity_problem, gradient_problem

pytestmark = pytest.mark.gpu

SX = np.array([[0, 1], [1, 0]], complex)
SY = np.array([[0, -1j], [1j, 0]], complex)
SZ = np.diag([1.0 + 0j, -1.0])


@pytest.fixture(scope="module", autouse=True)
def device():
    native.init(0)


def quickstart(psi_init=None, seed=42):
    """README quickstart: 1 qubit, H0 = (2pi 10/2) sz, controls 0.5 sx / 0.5 sy, T = pi, 100 slices."""
    sys_ = P.quantum_system((2 * np.pi * 10 / 2) * SZ, [0.5 * SX, 0.5 * SY])
    tgt = P.state_target([0, 1], psi_init=psi_init)
    ctrl = P.random_controls(sys_, np.pi, 100, amplitude=0.5, rng=np.random.default_rng(seed))
    return sys_, tgt, ctrl


def test_compute_fidelity_and_gradient_match_oracle():
    sys_, tgt, ctrl = quickstart(psi_init=[1, 0])
    F = P.compute_fidelity(sys_, ctrl, tgt)
    F0, _ = O.evaluate(fidelity_problem(sys_, ctrl, tgt), ctrl.controls, grad=False)
    assert abs(F - F0) < 1e-13
    G = P.compute_grape_gradient(sys_, ctrl, tgt)
    _, G0 = O.evaluate(gradient_problem(sys_, ctrl, tgt), ctrl.controls)
    assert G.shape == (2, 100)
    assert np.abs(G - G0).max() <= 1e-10 * np.abs(G0).max()


def _oracle_grape(sys_, tgt, ctrl, config):
    """The reference loop (GRAPE.jl:469-643) driven by oracle evaluations."""
    fp = fidelity_problem(sys_, ctrl, tgt)
    gp = gradient_problem(sys_, ctrl, tgt, in_optimizer=True)
    c = np.array(ctrl.controls)
    hist = [O.evaluate(fp, c, grad=False)[0]]
    step = config.step_size
    for it in range(config.max_iter):
        F = O.evaluate(fp, c, grad=False)[0]
        G = O.evaluate(gp, c)[1]
        hist.append(F)
        if len(hist) >= 2:
            step = P.adapt_grape_step_size(hist, step, config.min_step_size, config.max_step_size)
        c += step * G
        conv, _ = P.check_grape_convergence(hist, float(np.linalg.norm(G)), config)
        if conv:
            break
    return hist


def test_grape_optimize_trajectory_matches_oracle_loop():
    sys_, tgt, ctrl = quickstart()
    cfg = P.GRAPEConfig(max_iter=60, step_size=0.05)
    res = P.grape_optimize(sys_, tgt, ctrl, config=cfg)
    ref = _oracle_grape(sys_, tgt, ctrl, cfg)
    n = min(len(ref), len(res.fidelity_history))
    assert n >= 30
    assert np.abs(np.array(res.fidelity_history[:n]) - np.array(ref[:n])).max() < 1e-10
    # history[1] repeats the initial evaluation (GRAPE.jl:391-393)
    assert res.fidelity_history[0] == res.fidelity_history[1]


def test_grape_optimize_converges_quickstart():
    # test/Algorithms/GRAPETests.jl acceptance: the fixed-point problem reaches high fidelity
    sys_, tgt, ctrl = quickstart()
    res = P.grape_optimize(sys_, tgt, ctrl, config=P.GRAPEConfig(max_iter=500, step_size=0.05))
    assert res.fidelity > 0.99
    assert res.n_gradient_evaluations == res.n_iterations
    assert res.controls.shape == (2, 100)


def test_grape_optimize_gate_target_is_reproducible():
    rng = np.random.default_rng(3)
    sys_ = P.quantum_system(0.5 * SZ, [0.5 * SX, 0.5 * SY])
    tgt = P.unitary_target(SX)
    ctrl = P.random_controls(sys_, np.pi, 50, amplitude=1.0, rng=rng)
    r1 = P.grape_optimize(sys_, tgt, ctrl, config=P.GRAPEConfig(max_iter=200, step_size=0.05))
    r2 = P.grape_optimize(sys_, tgt, ctrl, config=P.GRAPEConfig(max_iter=200, step_size=0.05))
    assert r1.fidelity == r2.fidelity  # GRAPETests.jl:141-161 bit-reproducibility
    assert r1.fidelity > r1.fidelity_history[0]


def test_build_ensemble_from_systems_batched_mean():
    p, u = W.four_qubit_ensemble(n_t=40, n_det=3, n_amp=2)
    Hc = list(p.Hc)
    systems = [P.quantum_system(p.H0[m], [p.ctrl_scale[m, j] * Hc[j] for j in range(8)]) for m in range(6)]
    tgt = P.unitary_target(p.U_target)
    ctrl = P.ControlSequence(u.T, p.dt, p.n_t)
    obj = P.build_ensemble_from_systems(systems, tgt, ctrl)
    theta = u.reshape(-1, order="F")
    F, G = P.ensemble_value_and_grad(obj, theta)
    F0, G0 = O.evaluate(p, u)
    assert abs(F - F0) < 1e-13
    assert np.abs(G - G0.reshape(-1, order="F")).max() <= 1e-10 * np.abs(G0).max()
    assert abs(P.ensemble_value(obj, theta) - F) < 1e-15
    # worst-case aggregator: per-member closures (EnsembleObjective.jl:187-217)
    objw = P.build_ensemble_from_systems(systems, tgt, ctrl, aggregator="worst_case")
    Fw, Gw = P.ensemble_value_and_grad(objw, theta)
    _, _, Fm, Gm = O.evaluate(p, u, per_member=True)
    w = int(np.argmin(Fm))
    assert abs(Fw - Fm[w]) < 1e-13
    assert np.abs(Gw - Gm[w].reshape(-1, order="F")).max() <= 1e-10 * np.abs(Gm[w]).max()


def test_grape_state_kernel_and_lindblad_kernel():
    rng = np.random.default_rng(12)
    N = 2
    ctrl = P.MRControl(drifts=[0.3 * SZ, -0.2 * SZ], operators=[0.5 * SX, 0.5 * SY],
                       rho_init=[np.array([1, 0], complex)], rho_targ=[np.array([0, 1], complex)],
                       pulse_dt=np.full(30, 0.05), pwr_levels=[2 * np.pi * 1.0, 2 * np.pi * 1.1])
    w = rng.uniform(-1, 1, (2, 30))
    F, G = P.grape_state_kernel(w, ctrl)
    F0, G0 = O.evaluate(ctrl.problem(), w)
    assert abs(F - F0) < 1e-13 and np.abs(G - G0).max() <= 1e-10 * np.abs(G0).max()
    lctrl = P.LindbladMRControl(drifts=[0.3 * SZ], operators=[0.5 * SX, 0.5 * SY],
                                rho_init=[np.array([[1, 0], [0, 0]], complex)],
                                rho_targ=[np.array([[0, 0], [0, 1]], complex)], pulse_dt=np.full(30, 0.05),
                                jump_ops=[np.array([[0, 1], [0, 0]], complex)], decay_rates=[0.2],
                                pwr_levels=[2 * np.pi * 1.0])
    F, G = P.grape_lindblad_kernel(w, lctrl)
    F0, G0 = O.evaluate(lctrl.problem(), w)
    assert abs(F - F0) < 1e-13 and np.abs(G - G0).max() <= 1e-10 * np.abs(G0).max()


# ----------------------------------------------------------------- verification


def _oracle_fd(sys_, tgt, ctrl, eps):
    u = np.array(ctrl.controls)
    G = np.zeros_like(u)
    p = fidelity_problem(sys_, ctrl, tgt)
    for j in range(u.shape[0]):
        for k in range(u.shape[1]):
            up, um = u.copy(), u.copy()
            up[j, k] += eps
            um[j, k] -= eps
            G[j, k] = (O.evaluate(p, up, grad=False)[0] - O.evaluate(p, um, grad=False)[0]) / (2 * eps)
    return G


def test_finite_difference_gradient_matches_oracle_differences():
    sys_, tgt, ctrl = quickstart(psi_init=[1, 0])
    ctrl = P.ControlSequence(ctrl.amplitudes[:12], ctrl.dt, 12)
    G = P.finite_difference_gradient(sys_, ctrl, tgt, eps=1e-6)
    G0 = _oracle_fd(sys_, tgt, ctrl, 1e-6)
    assert G.shape == (2, 12)
    assert np.abs(G - G0).max() <= 1e-7 * np.abs(G0).max()


def _autodiff_qubit(dt=5e-9, n_ts=20):
    # test/AdvancedFeatures/AutoDiffTests.jl:11-21
    sys_ = P.quantum_system(2 * np.pi * 1e3 * SZ, [SX, SY])
    tgt = P.state_target([0.0, 1.0])
    u = np.zeros((2, n_ts)) + 0.05
    return sys_, tgt, P.ControlSequence(u.T, dt, n_ts)


def test_autodiff_backends_match_grape_gradient():
    sys_, tgt, ctrl = _autodiff_qubit()
    g_fd = P.compute_gradient_autodiff(sys_, ctrl, tgt, P.AutoDiffConfig(backend="finite_diff"))
    g_grape = P.compute_grape_gradient(sys_, ctrl, tgt)
    assert g_fd.shape == (2, 20) and np.all(np.isfinite(g_fd))
    assert np.abs(g_fd - g_grape).max() < 1e-5  # AutoDiffTests.jl:33-40
    g_ad = P.compute_gradient_autodiff(sys_, ctrl, tgt)  # exact adjoint
    assert np.abs(g_ad - g_fd).max() <= 1e-6 * max(1.0, np.abs(g_fd).max())
    assert isinstance(P.verify_gradient_autodiff(sys_, ctrl, tgt, tol=1e-4, verbose=False), bool)


def test_grape_gradient_check_and_wrappers():
    sys_, tgt, ctrl = quickstart(psi_init=[1, 0])
    ctrl = P.ControlSequence(ctrl.amplitudes[:20], ctrl.dt, 20)
    chk = P.grape_gradient_check(sys_, tgt, ctrl, eps=1e-5)
    G = P.compute_grape_gradient(sys_, ctrl, tgt)
    assert np.array_equal(chk["analytical_gradient"], G)
    assert chk["max_error"] == pytest.approx(np.abs(G - chk["numerical_gradient"]).max())
    assert chk["passed"] == (chk["max_error"] < 1e-4)
    assert np.array_equal(P.grape_gradient(sys_, tgt, ctrl.controls, ctrl.dt), G)
    assert P.evaluate_fidelity(sys_, tgt, ctrl.controls, ctrl.dt) == P.compute_fidelity(sys_, ctrl, tgt)


def test_grape_optimize_ensemble_matches_oracle_driven_loop():
    # broadband Rx(pi) over a detuning ensemble (GRAPE.jl:1000-1011 docstring example, shortened)
    from pulsar_rocm.api import _ensemble_lbfgs, ensemble_problems
    Iz, Ix = 0.5 * SZ, 0.5 * SX
    systems = [P.quantum_system(2 * np.pi * df * Iz, [2 * np.pi * Ix]) for df in np.linspace(-5000, 5000, 7)]
    tgt = P.unitary_target(np.array([[0, -1j], [-1j, 0]]))
    ctrl = P.random_controls(systems[3], 500e-6, 50, amplitude=1000.0, rng=np.random.default_rng(4))
    cfg = P.GRAPEConfig(max_iter=6, step_size=0.05)
    res = P.grape_optimize_ensemble(systems, tgt, ctrl, config=cfg, amplitude_limit=10_000.0)
    probs = ensemble_problems(systems, tgt, ctrl.n_steps, ctrl.dt)

    def F(u):
        return sum(O.evaluate(q, u, grad=False)[0] for q in probs)

    def G(u):
        return sum(O.evaluate(q, u)[1] for q in probs)

    ref = _ensemble_lbfgs(F, G, ctrl.controls, cfg, 10, 10_000.0)
    assert len(res.fidelity_history) == len(ref[2])
    assert np.abs(np.array(res.fidelity_history) - np.array(ref[2])).max() < 1e-10
    assert res.fidelity >= res.fidelity_history[0]
    assert res.metadata["algorithm"] == "Ensemble L-BFGS-GRAPE"


@pytest.mark.parametrize("memory,limit", [(3, 10_000.0), (10, np.inf)])
def test_device_ensemble_lbfgs_matches_host_loop(memory, limit):
    # pr_lbfgs_optimize against the Python loop over the same device
    # evaluations: a short history (the oldest pairs dropped) with clipping, and
    # the default history without it
    Iz, Ix, Iy = 0.5 * SZ, 0.5 * SX, 0.5 * SY
    systems = [P.quantum_system(2 * np.pi * df * Iz, [2 * np.pi * Ix, 2 * np.pi * Iy])
               for df in np.linspace(-4000, 4000, 5)]
    tgt = P.unitary_target(np.array([[0, -1j], [-1j, 0]]))
    ctrl = P.random_controls(systems[2], 400e-6, 40, amplitude=800.0, rng=np.random.default_rng(12))
    cfg = P.GRAPEConfig(max_iter=14, step_size=0.05, convergence_tol=1e-12)
    print(f"case memory={memory} limit={limit}", flush=True)
    dev = P.grape_optimize_ensemble(systems, tgt, ctrl, config=cfg, lbfgs_memory=memory, amplitude_limit=limit)
    host = P.grape_optimize_ensemble(systems, tgt, ctrl, config=cfg, lbfgs_memory=memory, amplitude_limit=limit,
                                     on_device=False)
    assert dev.metadata.get("on_device") is True
    assert dev.converged == host.converged and dev.termination_reason == host.termination_reason
    assert len(dev.fidelity_history) == len(host.fidelity_history)
    assert np.abs(np.array(dev.fidelity_history) - np.array(host.fidelity_history)).max() <= 1e-10
    assert np.allclose(dev.gradient_norm_history, host.gradient_norm_history, rtol=1e-8, atol=0)
    assert dev.n_fidelity_evaluations == host.n_fidelity_evaluations
    assert dev.n_gradient_evaluations == host.n_gradient_evaluations
    assert abs(dev.fidelity - host.fidelity) <= 1e-10
    assert np.abs(dev.controls - host.controls).max() <= 1e-6 * np.abs(host.controls).max()
    assert dev.fidelity > dev.fidelity_history[0]


# ----------------------------------------------------------------- device GRAPE loop


@pytest.mark.parametrize("case", ["state_fixed_point", "state_transfer", "gate"])
def test_device_grape_loop_matches_host_loop(case):
    if case == "gate":
        rng = np.random.default_rng(3)
        sys_ = P.quantum_system(0.5 * SZ, [0.5 * SX, 0.5 * SY])
        tgt = P.unitary_target(SX)
        ctrl = P.random_controls(sys_, np.pi, 50, amplitude=1.0, rng=rng)
        cfg = P.GRAPEConfig(max_iter=150, step_size=0.05)
    else:
        sys_, tgt, ctrl = quickstart(psi_init=[1, 0] if case == "state_transfer" else None)
        cfg = P.GRAPEConfig(max_iter=300, step_size=0.05)
    dev = P.grape_optimize(sys_, tgt, ctrl, config=cfg)  # pr_grape_optimize
    host = P.grape_optimize(sys_, tgt, ctrl, config=cfg, on_device=False)
    assert dev.n_iterations == host.n_iterations
    assert dev.converged == host.converged and dev.termination_reason.split()[0] == host.termination_reason.split()[0]
    assert len(dev.fidelity_history) == len(host.fidelity_history)
    assert np.abs(np.array(dev.fidelity_history) - np.array(host.fidelity_history)).max() <= 1e-13
    # |G| is summed on the device in one loop and by numpy in the other; near
    # the optimum it falls to rounding level, so the tolerance has an absolute floor
    gn = np.asarray(host.gradient_norm_history)
    assert np.allclose(dev.gradient_norm_history, gn, rtol=1e-12, atol=1e-12 * gn.max())
    assert np.abs(dev.controls - host.controls).max() <= 1e-12 * max(1.0, np.abs(host.controls).max())
    assert abs(dev.fidelity - host.fidelity) <= 1e-13
    assert dev.n_fidelity_evaluations == host.n_fidelity_evaluations
    assert dev.n_gradient_evaluations == host.n_gradient_evaluations
    assert dev.metadata["step_size_final"] == pytest.approx(host.metadata["step_size_final"], rel=1e-12)


@pytest.mark.parametrize("ftype", ["real", "square"])
def test_fidelity_forward_single_and_population(ftype):
    # fidelity_forward (GRAPEState.jl:406-566): F only, members drift x pwr,
    # averaged over members and pairs; a population plan gives the same F per
    # waveform as single evaluations
    rng = np.random.default_rng(31)
    Iz, Ix, Iy = SZ / 2, SX / 2, SY / 2
    H2 = np.kron(Iz, np.eye(2)) * 2 * np.pi * 40 + np.kron(np.eye(2), Iz) * 2 * np.pi * -25
    H2 = H2 + 2 * np.pi * 5 * np.kron(Iz, Iz)
    drifts = [H2, 1.02 * H2]
    ops = [np.kron(Ix, np.eye(2)) + np.kron(np.eye(2), Ix), np.kron(Iy, np.eye(2)) + np.kron(np.eye(2), Iy)]
    rho_i = [W.random_state(4, rng) for _ in range(3)]
    rho_t = [W.random_state(4, rng) for _ in range(3)]
    dts = rng.uniform(1e-6, 3e-6, 80)
    pwr = [2 * np.pi * 4e4, 2 * np.pi * 5e4]
    wfs = [rng.uniform(-1, 1, (2, 80)) for _ in range(5)]
    ctrl = P.MRControl(drifts=drifts, operators=ops, rho_init=rho_i, rho_targ=rho_t, pulse_dt=dts,
                       pwr_levels=pwr, fidelity=ftype)
    F1 = P.fidelity_forward(wfs[0], drifts, pwr, ops, rho_i, rho_t, dts, fidelity_type=ftype)
    F0, _ = O.evaluate(ctrl.problem(), wfs[0], grad=False)
    assert abs(F1 - F0) <= 1e-12, (F1, F0)
    ff = P.FidelityForward(drifts, pwr, ops, rho_i, rho_t, dts, fidelity_type=ftype, population=5)
    try:
        Fb = ff.batch(wfs)
    finally:
        ff.close()
    assert Fb[0] == F1
    for w, f in zip(wfs, Fb):
        Fo, _ = O.evaluate(ctrl.problem(), w, grad=False)
        assert abs(f - Fo) <= 1e-12
    with pytest.raises(ValueError):
        P.fidelity_forward(wfs[0], drifts, pwr, ops, rho_i, rho_t, dts, backend="cuda")


# ----------------------------------------------------------------- quantum-computing callers


def test_robust_optimcon_qs_is_one_device_ensemble():
    # robust_optimcon_qs (QuasiStatic.jl:196-216): frequency and amplitude
    # noise (5 x 3 members) as one plan with the L-BFGS loop in the library;
    # the same as grape_optimize_ensemble over the systems built by hand and
    # as the Python loop over the same evaluations
    from pulsar_rocm import qc
    Iz, Ix, Iy = 0.5 * SZ, 0.5 * SX, 0.5 * SY
    srcs = [qc.QuasiStaticNoise("freq", 2e3, n_samples=5), qc.QuasiStaticNoise("amp", 0.05, distribution="uniform",
                                                                              n_samples=3)]
    H_fn = lambda p: 2 * np.pi * p[0] * Iz + p[1] * 2 * np.pi * 1e3 * Ix  # noqa: E731
    tgt = P.unitary_target(np.array([[0, -1j], [-1j, 0]]))
    sys0 = P.quantum_system(H_fn([0.0, 0.0]), [Ix, Iy])
    ctrl = P.random_controls(sys0, 300e-6, 30, amplitude=2 * np.pi * 800, rng=np.random.default_rng(21))
    cfg = P.GRAPEConfig(max_iter=10, step_size=0.05, convergence_tol=1e-12)
    pwr = [2 * np.pi, 2 * np.pi]
    res = qc.robust_optimcon_qs(H_fn, [Ix, Iy], srcs, tgt, ctrl, config=cfg, pwr_levels=pwr)
    assert res.metadata["on_device"] is True and res.metadata["n_systems"] == 15
    drifts = qc.quasi_static_ensemble(H_fn, srcs)
    systems = [P.quantum_system(H, [pwr[0] * Ix, pwr[1] * Iy]) for H in drifts]
    ref = P.grape_optimize_ensemble(systems, tgt, ctrl, config=cfg)
    host = P.grape_optimize_ensemble(systems, tgt, ctrl, config=cfg, on_device=False)
    assert res.fidelity_history == ref.fidelity_history and np.array_equal(res.controls, ref.controls)
    assert np.abs(np.array(res.fidelity_history) - np.array(host.fidelity_history)).max() <= 1e-10
    assert res.fidelity > res.fidelity_history[0]


def test_lindblad_qc_control_kernel_and_grape_loop():
    # LindbladQCControl (LindbladQCControl.jl:69-406): its kernel is one
    # PR_LINDBLAD problem on the device, equal to the oracle; the :grape loop
    # over it equals the loop driven by the oracle
    from pulsar_rocm import qc
    sysx = P.quantum_system(2 * np.pi * 0.3 * SZ / 2, [2 * np.pi * SX / 2, 2 * np.pi * SY / 2])
    tgt = P.state_target(np.array([0, 1], complex))
    ctrl = P.random_controls(sysx, 2.0, 25, amplitude=0.4, rng=np.random.default_rng(3))
    sm = np.array([[0, 1], [0, 0]], complex)
    pen = (lambda w: 1e-3 * float(np.sum(w ** 2)), lambda w: 2e-3 * w)
    mk = lambda: qc.LindbladQCControl(sysx, tgt, ctrl, jump_ops=[sm], decay_rates=[0.05], method="grape",  # noqa
                                      max_iter=30, step_size=0.5, penalty_fns=[pen[0]], penalty_grad_fns=[pen[1]])
    ctx = mk()
    F, G = qc.qc_kernel(ctrl.controls, ctx)
    F0, G0 = O.evaluate(ctx.problem(), ctrl.controls)
    F0 -= pen[0](ctrl.controls)
    G0 = G0 - pen[1](ctrl.controls)
    ctx.close()
    assert abs(F - F0) <= 1e-12 and np.abs(G - G0).max() <= 1e-10 * np.abs(G0).max()
    res = qc.optimcon(mk())
    # the same loop with oracle evaluations
    ref_ctx = mk()
    w = np.array(ctrl.controls)
    fh, Fp = [], -np.inf
    for it in range(1, 31):
        Fo, Go = O.evaluate(ref_ctx.problem(), w)
        Fo -= pen[0](w)
        Go = Go - pen[1](w)
        fh.append(Fo)
        if np.linalg.norm(Go) < ref_ctx.grape_config.gradient_norm_tol or (abs(Fo - Fp) < 1e-8 and it > 1):
            break
        Fp = Fo
        w = w + 0.5 * Go
    assert len(res.fidelity_history) == len(fh)
    assert np.abs(np.array(res.fidelity_history) - np.array(fh)).max() <= 1e-10
    assert res.fidelity >= res.fidelity_history[0] and res.metadata["liouville_dim"] == 4
