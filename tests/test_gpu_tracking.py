"""Tracking objective on the device (k_segment images + k_track walk) against
the tracking oracle (GRAPETracking.jl:162-263 restated with density matrices).

Both sides are FP64; the device propagates state columns with Taylor exps, the
oracle density matrices with eigen-exps: |dF| <= 1e-12 and
max|dG| / max|G| <= 1e-10, as for the other objectives.
"""
import numpy as np
import pytest

import oracle as O
from pulsar_rocm import mr, native, workloads as W
from pulsar_rocm.problem import PR_GATE, PR_MIXED, PR_TRACKING, Problem

pytestmark = pytest.mark.gpu

TOL_F = 1e-12
TOL_G = 1e-10


@pytest.fixture(scope="module", autouse=True)
def device():
    native.init(0)


def tracking_case(d, M, nc, nt, npairs, n_ck, seed, per_slice_dt=False, norm=3.0):
    rng = np.random.default_rng(seed)
    H0 = np.stack([W.random_hermitian(d, rng, norm) for _ in range(M)])
    Hc = np.stack([W.random_hermitian(d, rng, 1.0) for _ in range(nc)])
    scale = rng.uniform(0.8, 1.2, (M, 1)) * np.ones((1, nc))
    psi = np.stack([W.random_state(d, rng) for _ in range(npairs)])
    steps, w, targets = W.tracking_checkpoints(d, nt, n_ck, rng)
    # a checkpoint at the first slice and one sharing a step with another
    steps = np.concatenate([steps, [1, steps[0]]]).astype(np.int32)
    w = np.concatenate([w, [0.05, 0.07]])
    targets = np.concatenate([targets, [W.random_density(d, rng), W.random_density(d, rng, rank=1)]])
    dts = rng.uniform(0.05, 0.15, nt) if per_slice_dt else None
    p = Problem(dim=d, n_ctrl=nc, n_t=nt, H0=H0, Hc=Hc, dt=0.1, dt_slices=dts, objective=PR_TRACKING,
                n_members=M, ctrl_scale=scale, psi_init=psi, track_step=steps, track_weight=w,
                track_target=targets)
    u = rng.uniform(-2, 2, (nc, nt))
    return p, u


def run(p, u, **kw):
    plan = native.Plan([p], **kw)
    try:
        F, G = plan.eval([u])
    finally:
        plan.close()
    return F[0], G[0]


def check(p, u, **kw):
    print(f"tracking case d={p.dim} M={p.n_members} nc={p.n_ctrl} n_t={p.n_t} pairs={p.n_pairs} "
          f"checkpoints={p.n_track} {kw}", flush=True)
    F0, G0 = O.tracking_evaluate(p, u)
    F, G = run(p, u, **kw)
    assert abs(F - F0) <= TOL_F, (F, F0)
    assert np.abs(G - G0).max() <= TOL_G * np.abs(G0).max(), np.abs(G - G0).max() / np.abs(G0).max()
    return F, G


@pytest.mark.parametrize("d,M,nc,nt,npairs", [
    (2, 4, 2, 40, 1), (4, 3, 3, 30, 2), (16, 3, 8, 40, 4), (7, 2, 3, 25, 3),
    (24, 2, 4, 20, 5), (40, 2, 3, 12, 6), (64, 1, 2, 8, 3)])
def test_tracking_matches_oracle(d, M, nc, nt, npairs):
    p, u = tracking_case(d, M, nc, nt, npairs, n_ck=5, seed=d + nt)
    check(p, u)


def test_tracking_per_slice_dt_and_full_pair_count():
    p, u = tracking_case(16, 2, 4, 30, 16, n_ck=6, seed=3, per_slice_dt=True)
    check(p, u)


def test_tracking_every_slice_a_checkpoint():
    p, u = tracking_case(8, 2, 3, 15, 2, n_ck=15, seed=4)
    check(p, u)


def test_tracking_ignores_precision_and_follows_store_props_options():
    # tracking runs FP64 whatever the precision option says; store_props = 0
    # rebuilds U_k in the walk, other segments only reorder the chain sums
    p, u = tracking_case(16, 2, 4, 24, 3, n_ck=4, seed=5)
    F, G = run(p, u)
    for kw in ({"precision": PR_MIXED}, {"store_props": 0}, {"segment": 5}):
        F2, G2 = run(p, u, **kw)
        assert abs(F2 - F) <= 1e-14 and np.abs(G2 - G).max() <= 1e-13 * np.abs(G).max(), kw


@pytest.mark.parametrize("d,M,nc,nt,npairs", [(4, 3, 3, 30, 2), (16, 3, 8, 40, 4), (16, 2, 4, 30, 12),
                                               (24, 2, 4, 20, 5), (64, 1, 2, 8, 3)])
def test_tracking_recompute_matches_stored_images(d, M, nc, nt, npairs):
    # store_props = 0: no U_k images; the tracking walk rebuilds each U_k with
    # the segment kernel's own calls (packed and unpacked single-wavefront
    # walks, 2 x 2 and 4 x 4 blocks), so F and G equal the stored-image plan's
    # bit for bit at the same segments, and the oracle's
    p, u = tracking_case(d, M, nc, nt, npairs, n_ck=5, seed=d + nt + 1)
    print(f"tracking recompute d={d} M={M} n_t={nt} pairs={npairs}", flush=True)
    out = []
    for sp in (1, 0):
        plan = native.Plan([p], store_props=sp, segment=7)
        try:
            out.append(plan.eval([u]))
        finally:
            plan.close()
    (F1, G1), (F0, G0) = out
    assert F0[0] == F1[0]
    assert np.array_equal(G0[0], G1[0])
    check(p, u, store_props=0)


def test_tracking_beside_a_gate_problem_in_one_plan():
    pt, ut = tracking_case(16, 2, 4, 24, 3, n_ck=4, seed=6)
    pg, ug = W.shrink(*W.four_qubit_ensemble(n_t=24, n_det=2, n_amp=1))
    plan = native.Plan([pg, pt])
    try:
        F, G = plan.eval([ug, ut])
    finally:
        plan.close()
    Fg, Gg = O.evaluate(pg, ug)
    Ft, Gt = O.tracking_evaluate(pt, ut)
    assert abs(F[0] - Fg) <= TOL_F and np.abs(G[0] - Gg).max() <= TOL_G * np.abs(Gg).max()
    assert abs(F[1] - Ft) <= TOL_F and np.abs(G[1] - Gt).max() <= TOL_G * np.abs(Gt).max()


def test_tracking_fidelity_only_evaluation_and_replays():
    p, u = tracking_case(4, 2, 2, 20, 2, n_ck=4, seed=7)
    plan = native.Plan([p])
    try:
        F, G = plan.eval([u])
        F1, _ = plan.eval([u], grad=False)
        F2, G2 = plan.eval([u])
    finally:
        plan.close()
    assert F1[0] == F[0] and F2[0] == F[0] and np.array_equal(G2[0], G[0])


def test_mr_tracking_kernel_single_spin_oscillation():
    # the GRAPETracking.jl:69-80 example: one spin driven between +z and -z,
    # checkpoints every 10 slices, two power levels and two offsets
    Ix = W.SX / 2
    Iy = W.SY / 2
    Iz = W.SZ / 2
    up, dn = np.array([1, 0], complex), np.array([0, 1], complex)
    cps = [mr.TrackingPoint(10 * k, np.outer(dn, dn) if k % 2 else np.outer(up, up), 0.1) for k in range(1, 11)]
    ctrl = mr.MRControl(drifts=[2 * np.pi * 50 * Iz, -2 * np.pi * 80 * Iz], operators=[Ix, Iy],
                        rho_init=[up], rho_targ=[dn], pulse_dt=np.full(100, 2e-6), pwr_levels=[0.9, 1.1],
                        tracking=cps)
    w = np.random.default_rng(8).uniform(-2 * np.pi * 5e4, 2 * np.pi * 5e4, (2, 100))
    F, G = mr.grape_tracking_kernel(w, ctrl)
    F0, G0 = O.tracking_evaluate(ctrl.tracking_problem(), w)
    assert abs(F - F0) <= TOL_F and np.abs(G - G0).max() <= TOL_G * np.abs(G0).max()
    with pytest.raises(ValueError, match="tracking is empty"):
        mr.grape_tracking_kernel(w, mr.MRControl(drifts=ctrl.drifts, operators=ctrl.operators, rho_init=[up],
                                                 rho_targ=[dn], pulse_dt=ctrl.pulse_dt))


def test_four_qubit_tracking_workload_members_sample():
    p, u = W.shrink(*W.four_qubit_tracking(n_t=200), n_members=3)
    check(p, u)


@pytest.mark.parametrize("kind", ["state_lbfgs", "state_grape", "tracking_lbfgs", "lindblad_lbfgs"])
def test_mr_optimcon_on_device_matches_oracle_driven_loop(kind):
    # optimcon (OptControl.jl:310-611, 1040-1236): the same loop with device
    # evaluations and with oracle evaluations gives the same trajectory
    Ix, Iy, Iz = W.SX / 2, W.SY / 2, W.SZ / 2
    up, dn = np.array([1, 0], complex), np.array([0, 1], complex)
    mIy = np.array([1, -1j], complex) / np.sqrt(2)
    common = dict(drifts=[2 * np.pi * 300 * Iz, -2 * np.pi * 200 * Iz], operators=[Ix, Iy], rho_init=[up],
                  rho_targ=[mIy], pulse_dt=np.full(40, 5e-6), pwr_levels=[2 * np.pi * 8e3, 2 * np.pi * 1e4],
                  max_iter=12)
    if kind == "lindblad_lbfgs":
        ctrl = mr.LindbladMRControl(**common, jump_ops=[np.array([[0, 1], [0, 0]], complex)], decay_rates=[300.0])
        p = ctrl.problem()
        fg = lambda w: O.evaluate(p, w)  # noqa: E731
        f = lambda w: O.evaluate(p, w, grad=False)[0]  # noqa: E731
    elif kind == "tracking_lbfgs":
        cps = [mr.TrackingPoint(10 * k, np.outer(dn, dn) if k % 2 else np.outer(up, up), 0.25) for k in range(1, 5)]
        ctrl = mr.MRControl(**common, tracking=cps)
        p = ctrl.tracking_problem()
        fg = lambda w: O.tracking_evaluate(p, w)  # noqa: E731
        f = lambda w: O.tracking_evaluate(p, w, grad=False)[0]  # noqa: E731
    else:
        ctrl = mr.MRControl(**common, method="grape" if kind == "state_grape" else "lbfgs")
        p = ctrl.problem()
        fg = lambda w: O.evaluate(p, w)  # noqa: E731
        f = lambda w: O.evaluate(p, w, grad=False)[0]  # noqa: E731
    guess = 0.2 * np.random.default_rng(4).standard_normal((2, 40))
    dev = mr.optimcon(ctrl, guess)  # the loop in the library (pr_mr_optimcon)
    assert dev.metadata["on_device"] is True
    if kind == "state_grape":
        ref = mr.optimcon_grape_loop(fg, guess, ctrl.l_bound, ctrl.u_bound, ctrl.max_iter, ctrl.grad_tol)
    else:
        ref = mr.optimcon_lbfgs_loop(fg, f, guess, ctrl.l_bound, ctrl.u_bound, ctrl.max_iter, ctrl.grad_tol,
                                     ctrl.lbfgs_memory)
    assert dev.n_iterations == ref.n_iterations and dev.termination_reason == ref.termination_reason
    assert dev.n_fidelity_evaluations == ref.n_fidelity_evaluations
    assert np.abs(np.array(dev.fidelity_history) - np.array(ref.fidelity_history)).max() <= 1e-10
    assert np.abs(dev.controls - ref.controls).max() <= 1e-7


@pytest.mark.parametrize("method,pack,bound", [("lbfgs", True, 1.0), ("lbfgs", False, 0.35), ("grape", True, 0.35),
                                               ("grape", False, 1.0)])
def test_mr_optimcon_library_loop_matches_host_loop(method, pack, bound):
    # pr_mr_optimcon against the host loop over the same device plan: packed
    # (4 members per block, the waveform tiled and the gradient summed on the
    # device) and unpacked, with the bounds clipping (0.35) or not
    Ix, Iy, Iz = W.SX / 2, W.SY / 2, W.SZ / 2
    up = np.array([1, 0], complex)
    mIy = np.array([1, -1j], complex) / np.sqrt(2)
    ctrl = mr.MRControl(drifts=[2 * np.pi * 300 * Iz, -2 * np.pi * 200 * Iz], operators=[Ix, Iy], rho_init=[up],
                        rho_targ=[mIy], pulse_dt=np.full(40, 5e-6), pwr_levels=[2 * np.pi * 8e3, 2 * np.pi * 1e4],
                        max_iter=25, method=method, l_bound=-bound, u_bound=bound, lbfgs_memory=4)
    guess = 0.3 * np.random.default_rng(6).standard_normal((2, 40))
    print(f"case method={method} pack={pack} bound={bound}", flush=True)
    kp = mr.KernelPlan(ctrl)
    assert (kp.packed is not None) and kp.packed.f == 4
    kp.close()
    if pack:
        dev = mr.optimcon(ctrl, guess)
        host = mr.optimcon(ctrl, guess, on_device=False)
    else:
        # the unpacked plan: one member per block, tile 1
        kpu = mr.KernelPlan(ctrl, pack=False)
        try:
            cfg = native.MrConfig(native.PR_MR_LBFGS if method == "lbfgs" else native.PR_MR_GRAPE, ctrl.max_iter,
                                  ctrl.grad_tol, ctrl.lbfgs_memory, ctrl.l_bound, ctrl.u_bound, 1)
            best, fh, gh, res = native.mr_optimcon_device(kpu.plan, guess, cfg)
            fg = lambda w: kpu.eval(w)  # noqa: E731
            f = lambda w: kpu.eval(w, grad=False)[0]  # noqa: E731
            if method == "lbfgs":
                host = mr.optimcon_lbfgs_loop(fg, f, guess, ctrl.l_bound, ctrl.u_bound, ctrl.max_iter,
                                              ctrl.grad_tol, ctrl.lbfgs_memory)
            else:
                host = mr.optimcon_grape_loop(fg, guess, ctrl.l_bound, ctrl.u_bound, ctrl.max_iter, ctrl.grad_tol)
        finally:
            kpu.close()
        dev = type(host)(best, res.fidelity, fh, gh, res.n_iterations, bool(res.converged),
                         host.termination_reason, 0.0, res.n_fidelity_evaluations, res.n_gradient_evaluations, {})
        assert native.PR_STOP[res.reason] in ("max_iter", "perfect", "gradient", "fidelity", "zero_direction")
    assert dev.n_iterations == host.n_iterations and dev.termination_reason == host.termination_reason
    assert dev.n_fidelity_evaluations == host.n_fidelity_evaluations
    assert dev.n_gradient_evaluations == host.n_gradient_evaluations
    assert np.abs(np.array(dev.fidelity_history) - np.array(host.fidelity_history)).max() <= 1e-12
    assert np.allclose(dev.gradient_norm_history, host.gradient_norm_history, rtol=1e-9, atol=0)
    assert np.abs(dev.controls - host.controls).max() <= 1e-9
    assert abs(dev.fidelity - host.fidelity) <= 1e-12
    assert np.all(dev.controls <= bound) and np.all(dev.controls >= -bound)
    assert dev.fidelity > dev.fidelity_history[0]


@pytest.mark.parametrize("nd,pw,pairs", [(16, [0.9, 1.0, 1.1, 1.2], 1), (5, [0.9, 1.1], 2)])
def test_packed_mr_ensemble_on_device(nd, pw, pairs):
    # d = 2 members packed f per block (PackedMembers) against the unpacked oracle
    Ix, Iy, Iz = W.SX / 2, W.SY / 2, W.SZ / 2
    up, dn = np.array([1, 0], complex), np.array([0, 1], complex)
    mIy = np.array([1, -1j], complex) / np.sqrt(2)
    ctrl = mr.MRControl(drifts=[2 * np.pi * o * Iz for o in np.linspace(-800, 800, nd)], operators=[Ix, Iy],
                        rho_init=[up, dn][:pairs], rho_targ=[mIy, up][:pairs], pulse_dt=np.full(60, 4e-6),
                        pwr_levels=[2 * np.pi * 1e4 * x for x in pw])
    w = np.random.default_rng(nd).uniform(-1, 1, (2, 60))
    kp = mr.KernelPlan(ctrl)
    assert kp.packed is not None and kp.packed.f >= 2
    try:
        F, G = kp(w)
    finally:
        kp.close()
    F0, G0 = O.evaluate(ctrl.problem(), w)
    assert abs(F - F0) <= TOL_F and np.abs(G - G0).max() <= TOL_G * np.abs(G0).max()
    Fs, Gs = mr.grape_state_kernel(w, ctrl)
    assert Fs == F and np.array_equal(Gs, G)


def test_packed_lindblad_mr_ensemble_on_device():
    # N = 2 Lindblad members (4 x 4 Liouvillians) packed 4 per block
    Ix, Iy, Iz = W.SX / 2, W.SY / 2, W.SZ / 2
    up = np.array([1, 0], complex)
    mIy = np.array([1, -1j], complex) / np.sqrt(2)
    ctrl = mr.LindbladMRControl(drifts=[2 * np.pi * o * Iz for o in np.linspace(-300, 300, 6)], operators=[Ix, Iy],
                                rho_init=[up], rho_targ=[mIy], pulse_dt=np.full(40, 5e-6),
                                pwr_levels=[2 * np.pi * 8e3, 2 * np.pi * 1e4],
                                jump_ops=[np.array([[0, 1], [0, 0]], complex)], decay_rates=[300.0])
    w = np.random.default_rng(3).uniform(-1, 1, (2, 40))
    kp = mr.KernelPlan(ctrl)
    assert kp.packed is not None and kp.packed.f == 4
    try:
        F, G = kp(w)
    finally:
        kp.close()
    F0, G0 = O.evaluate(ctrl.problem(), w)
    assert abs(F - F0) <= TOL_F and np.abs(G - G0).max() <= TOL_G * np.abs(G0).max()


def test_packed_tracking_mr_ensemble_on_device():
    # one spin, 12 members packed 6 per block, checkpoints blockdiag(C, .., C)
    Ix, Iy, Iz = W.SX / 2, W.SY / 2, W.SZ / 2
    up, dn = np.array([1, 0], complex), np.array([0, 1], complex)
    cps = [mr.TrackingPoint(10 * k, np.outer(dn, dn) if k % 2 else np.outer(up, up), 0.25) for k in range(1, 5)]
    ctrl = mr.MRControl(drifts=[2 * np.pi * o * Iz for o in np.linspace(-300, 300, 6)], operators=[Ix, Iy],
                        rho_init=[up], rho_targ=[dn], pulse_dt=np.full(40, 5e-6),
                        pwr_levels=[2 * np.pi * 8e3, 2 * np.pi * 1e4], tracking=cps)
    w = np.random.default_rng(6).uniform(-1, 1, (2, 40))
    kp = mr.KernelPlan(ctrl, tracking=True)
    assert kp.packed is not None and kp.packed.f == 6
    try:
        F, G = kp(w)
    finally:
        kp.close()
    F0, G0 = O.tracking_evaluate(ctrl.tracking_problem(), w)
    assert abs(F - F0) <= TOL_F and np.abs(G - G0).max() <= TOL_G * np.abs(G0).max()
