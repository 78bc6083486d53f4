"""Pins the tracking oracle (oracle.tracking_evaluate, CPU only) and the C-ABI
validation of tracking problems.

The reference's test suite holds no tracking case, so the oracle is pinned by
what the mathematics fixes:
  - F: a pure-target checkpoint is the squared overlap |<phi|P_j psi>|^2 with
    P_j the product of the first j propagators (scipy.linalg.expm, independent
    of the oracle's eigen-exp);
  - G: the reference's formula is -dt Re(i Tr(Lambda^dag [s A, rho])) with
    Lambda built from 2 w (rho_t - rho); it equals twice the derivative of F
    when slice n's propagator is replaced by exp(-i eps dt s A) U_n, which a
    central difference of density-matrix propagations computes to ~1e-10;
  - only the Hermitian part of a target matters, and checkpoints at one step
    add linearly (what the device's combined checkpoint matrices rely on).
A dropped factor (the 2, dt or s), a sign, a checkpoint read one slice early
or late, or a transposed commutator fails one of these.
"""
import numpy as np
import pytest
import scipy.linalg as sla

import oracle as O
from pulsar_rocm import native, workloads as W
from pulsar_rocm.problem import PR_CLOSED, PR_GRAD_EXACT, PR_LINDBLAD, PR_TRACKING, Problem


def small_problem(d=4, M=2, nc=3, nt=12, npairs=2, seed=0, per_slice_dt=False):
    rng = np.random.default_rng(seed)
    H0 = np.stack([W.random_hermitian(d, rng, 3.0) for _ in range(M)])
    Hc = np.stack([W.random_hermitian(d, rng, 1.0) for _ in range(nc)])
    scale = rng.uniform(0.8, 1.2, (M, 1)) * np.ones((1, nc))
    psi = np.stack([W.random_state(d, rng) for _ in range(npairs)])
    steps = np.array([3, 3, 7, nt, 1], np.int32)
    targets = np.stack([W.random_density(d, rng, rank=1), W.random_density(d, rng), W.random_density(d, rng),
                        W.random_density(d, rng, rank=1), W.random_density(d, rng)])
    weights = np.array([0.3, 0.2, 0.25, 0.15, 0.1])
    dts = rng.uniform(0.05, 0.15, nt) if per_slice_dt else None
    p = Problem(dim=d, n_ctrl=nc, n_t=nt, H0=H0, Hc=Hc, dt=0.1, dt_slices=dts, dynamics=PR_CLOSED,
                objective=PR_TRACKING, n_members=M, ctrl_scale=scale, psi_init=psi,
                track_step=steps, track_weight=weights, track_target=targets)
    u = rng.uniform(-2, 2, (nc, nt))
    return p, u


def props(p, u, m, eps=None):
    """U_n = exp(-i dt_n H_n) by scipy; eps = (n, j, e) left-multiplies slice n
    by exp(-i e dt_n s_mj A_j)."""
    dts = np.full(p.n_t, p.dt) if p.dt_slices is None else np.asarray(p.dt_slices)
    H0 = np.asarray(p.H0)[m]
    s = np.asarray(p.ctrl_scale)[m]
    out = []
    for n in range(p.n_t):
        H = H0 + sum(s[j] * u[j, n] * p.Hc[j] for j in range(p.n_ctrl))
        U = sla.expm(-1j * dts[n] * H)
        if eps is not None and eps[0] == n:
            U = sla.expm(-1j * eps[2] * dts[n] * s[eps[1]] * p.Hc[eps[1]]) @ U
        out.append(U)
    return out


def direct_F(p, u, eps_member=None, eps=None):
    """sum_m w_m sum_s sum_c w_c Re Tr(rho_c rho_s(step_c)) by density-matrix propagation."""
    M, P = p.n_members, p.n_pairs
    F = 0.0
    for m in range(M):
        Us = props(p, u, m, eps if m == eps_member else None)
        for s in range(P):
            rho = np.outer(p.psi_init[s], p.psi_init[s].conj())
            traj = [rho]
            for U in Us:
                rho = U @ rho @ U.conj().T
                traj.append(rho)
            for st, w, tg in zip(p.track_step, p.track_weight, p.track_target):
                F += w * np.real(np.trace(tg @ traj[st])) / (M * P)
    return F


def test_pure_checkpoint_is_squared_overlap():
    p, u = small_problem(d=3, M=1, nt=6, npairs=2, seed=1)
    phi = W.random_state(3, np.random.default_rng(9))
    for j in (1, 4, 6):
        p.track_step = np.array([j], np.int32)
        p.track_weight = np.array([1.0])
        p.track_target = np.outer(phi, phi.conj())[None]
        Us = props(p, u, 0)
        Pj = np.eye(3)
        for U in Us[:j]:
            Pj = U @ Pj
        want = np.mean([abs(phi.conj() @ Pj @ p.psi_init[s]) ** 2 for s in range(2)])
        F, _ = O.tracking_evaluate(p, u)
        assert abs(F - want) < 1e-13, (j, F, want)


@pytest.mark.parametrize("per_slice_dt", [False, True])
def test_gradient_is_twice_the_left_perturbation_derivative(per_slice_dt):
    p, u = small_problem(per_slice_dt=per_slice_dt, seed=2)
    F, G = O.tracking_evaluate(p, u)
    assert abs(F - direct_F(p, u)) < 1e-13
    # G sums the members: the FD perturbs slice n of one member at a time
    h = 1e-5
    rng = np.random.default_rng(4)
    for n, j in [(0, 0), (2, 1), (3, 2), (6, 0), (7, 1), (p.n_t - 1, 2)] + [
            (int(rng.integers(p.n_t)), int(rng.integers(p.n_ctrl))) for _ in range(4)]:
        fd = 0.0
        for m in range(p.n_members):
            fd += (direct_F(p, u, m, (n, j, h)) - direct_F(p, u, m, (n, j, -h))) / (2 * h)
        assert abs(G[j, n] - 2 * fd) <= 1e-8 * np.abs(G).max(), (n, j, G[j, n], 2 * fd)


def test_gradient_is_twice_the_control_derivative_for_commuting_generators():
    # with H_n = s u_n A (no drift, one control) the left perturbation equals
    # the change of u itself, so G = 2 dF/du exactly
    p, u = small_problem(d=3, M=2, nc=1, nt=8, seed=3)
    p.H0 = np.zeros_like(np.asarray(p.H0))
    _, G = O.tracking_evaluate(p, u)
    h = 1e-5
    for n in range(p.n_t):
        up, um = u.copy(), u.copy()
        up[0, n] += h
        um[0, n] -= h
        fd = (direct_F(p, up) - direct_F(p, um)) / (2 * h)
        assert abs(G[0, n] - 2 * fd) <= 1e-8 * np.abs(G).max(), (n, G[0, n], 2 * fd)


def test_only_the_hermitian_part_of_a_target_counts():
    p, u = small_problem(seed=5)
    F0, G0 = O.tracking_evaluate(p, u)
    K = W.random_hermitian(p.dim, np.random.default_rng(6), 1.0)
    t = np.array(p.track_target)
    t[2] = t[2] + 1j * K  # anti-Hermitian addition
    p.track_target = t
    F1, G1 = O.tracking_evaluate(p, u)
    assert abs(F1 - F0) < 1e-14 and np.abs(G1 - G0).max() < 1e-13 * np.abs(G0).max()


def test_checkpoints_at_one_step_add_linearly():
    p, u = small_problem(seed=7)
    F0, G0 = O.tracking_evaluate(p, u)
    # the two checkpoints at step 3 merged into one of weight 1
    t = np.array(p.track_target)
    w = np.array(p.track_weight)
    merged = w[0] * t[0] + w[1] * t[1]
    p.track_step = np.array([3, 7, p.n_t, 1], np.int32)
    p.track_weight = np.array([1.0, w[2], w[3], w[4]])
    p.track_target = np.stack([merged, t[2], t[3], t[4]])
    F1, G1 = O.tracking_evaluate(p, u)
    assert abs(F1 - F0) < 1e-14 and np.abs(G1 - G0).max() < 1e-13 * np.abs(G0).max()


def test_oracle_rejects_out_of_range_steps():
    p, u = small_problem()
    p.track_step = np.array([0, 3, 7, 12, 1], np.int32)
    with pytest.raises(ValueError, match="out of range"):
        O.tracking_evaluate(p, u)


@pytest.mark.parametrize("change,code", [
    ("step0", -1), ("step_past_end", -1), ("no_points", -1), ("lindblad", -5), ("exact", -5)])
def test_c_abi_validates_tracking_problems(change, code):
    # validation precedes any device use, so this runs without a GPU
    p, u = small_problem()
    kw = {}
    if change == "step0":
        p.track_step = np.array([0, 3, 7, 12, 1], np.int32)
    elif change == "step_past_end":
        p.track_step = np.array([13, 3, 7, 12, 1], np.int32)
    elif change == "no_points":
        p.track_step = np.zeros(0, np.int32)
        p.track_weight = np.zeros(0)
        p.track_target = np.zeros((0, p.dim, p.dim), complex)
    elif change == "lindblad":
        p.dynamics = PR_LINDBLAD
    elif change == "exact":
        kw["grad_mode"] = PR_GRAD_EXACT
    with pytest.raises(native.RocmError) as e:
        native.Plan([p], **kw)
    assert e.value.code == code, str(e.value)
