"""Packing of small problems (pr_options.pack): closed-dynamics problems with
one member and dim <= 8, of the gate or a state-pair objective, run as
diagonal blocks of one problem of dim <= 16.

Each problem keeps its own F and G (compute_fidelity / compute_grape_gradient
per problem, Objectives.jl:609-639, Gradients.jl:80-163, and the MR state
kernel's pairs, GRAPEState.jl:218-229): packed plans are
checked against the FP64 oracle at the parity tolerances, against the same
plan without packing at rounding level, and their device paths (host buffers,
device buffers, static scratch, graph replays) against each other bit for bit.
"""
import numpy as np
import pytest

import oracle as O
from pulsar_rocm import native, workloads as W
from pulsar_rocm.problem import (PR_GATE, PR_GRAD_EXACT, PR_MIXED, PR_PAIRS_MODULUS, PR_PAIRS_REAL,
                                 PR_PAIRS_SQUARE, PR_SEGMENT_AUTOTUNE, PR_STATE, Problem, controls_cm)

pytestmark = pytest.mark.gpu

# (imported before the library initialises the device: imported after it,
# torch found no GPU in the process)
torch = pytest.importorskip("torch")

TOL_F = 1e-12
TOL_G = 1e-10


@pytest.fixture(scope="module", autouse=True)
def device():
    native.init(0)


def small_gates(n, seed, dims=(1, 2, 3, 4, 5, 6, 7, 8), n_t=(30, 90), dts=(0.02, 0.03)):
    """Independent gate problems: every dimension up to 8, lengths in n_t, two
    slice lengths, 1-3 controls, some with control scales or an explicit weight 1."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        d = int(rng.choice(dims))
        nc = int(rng.integers(1, 4))
        nt = int(rng.integers(n_t[0], n_t[1] + 1))
        kw = {}
        if i % 3 == 1:
            kw["ctrl_scale"] = rng.uniform(0.8, 1.2, (1, nc))
        if i % 5 == 2:
            kw["weight"] = np.ones(1)
        p = Problem(dim=d, n_ctrl=nc, n_t=nt, dt=float(dts[i % len(dts)]),
                    H0=W.random_hermitian(d, rng, 3.0) if d > 1 else np.array([[1.3 + 0j]]),
                    Hc=np.stack([W.random_hermitian(d, rng, 2.0) if d > 1 else np.array([[0.7 + 0j]])
                                 for _ in range(nc)]),
                    objective=PR_GATE, U_target=W.haar_unitary(d, rng), **kw)
        out.append((p, rng.uniform(-3, 3, (nc, nt))))
    return out


def evaluate(probs, us, **kw):
    plan = native.Plan(probs, **kw)
    try:
        F, G, D = plan.eval(us, diag=True)
        counts = plan.exp_counts()
    finally:
        plan.close()
    return F, G, D, counts


def packed_rows(counts):
    """Problems whose exps were reported by their pack's first problem."""
    return int(np.sum(np.all(counts == 0, axis=1)))


def g_close(g, g0, rel):
    """max|g - g0| <= rel max|g0|; a d = 1 problem's gradient is zero up to
    rounding (|F| = 1 whatever the controls), hence the floor of 1e-4."""
    return np.abs(g - g0).max() <= rel * max(np.abs(g0).max(), 1e-4)


def check_oracle(probs, us, F, G, grad_mode=0, tol_f=TOL_F, tol_g=TOL_G):
    for i, (p, u) in enumerate(zip(probs, us)):
        F0, G0 = O.evaluate(p, u, grad_mode=grad_mode)
        assert abs(F[i] - F0) <= tol_f, (i, p.dim, p.n_t, F[i], F0)
        assert g_close(G[i], G0, tol_g), (i, p.dim, p.n_t)


def test_packed_plan_matches_unpacked_plan_and_oracle():
    batch = small_gates(48, seed=1)
    probs, us = [b[0] for b in batch], [b[1] for b in batch]
    print("packing", [(p.dim, p.n_t, p.dt) for p in probs], flush=True)
    F1, G1, D1, c1 = evaluate(probs, us, pack=1)
    F0, G0, D0, c0 = evaluate(probs, us, pack=0)
    assert packed_rows(c0) == 0
    assert packed_rows(c1) >= 10, c1  # packs were formed
    # exps actually run: the packs' slices instead of every problem's
    assert c1[:, :2].sum() < c0[:, :2].sum()
    assert np.abs(F1 - F0).max() <= 1e-13
    for g1, g0 in zip(G1, G0):
        assert g_close(g1, g0, 1e-11)
    assert D1[:, 0].max() < 1e-12  # unitarity defect of every pack's U_total
    check_oracle(probs, us, F1, G1)


def test_packing_leaves_other_problems_to_their_own_kernels():
    # packable problems beside an ensemble, a pairs objective, a Lindblad
    # problem, per-slice dt, a weight other than 1 and a 2 x 2-block problem
    rng = np.random.default_rng(7)
    batch = small_gates(20, seed=4, n_t=(40, 44), dts=(0.025,))
    e, ue = W.four_qubit_ensemble(n_t=40)
    e, ue = W.shrink(e, ue, n_members=3)
    d = 5
    pairs = Problem(dim=d, n_ctrl=2, n_t=30, dt=0.05, H0=W.random_hermitian(d, rng, 2.0),
                    Hc=np.stack([W.random_hermitian(d, rng, 1.0) for _ in range(2)]),
                    objective=PR_PAIRS_SQUARE,
                    psi_init=np.stack([W.random_state(d, rng) for _ in range(2)]),
                    psi_targ=np.stack([W.random_state(d, rng) for _ in range(2)]))
    lind, ul = W.three_qubit_lindblad(n_t=8, n_members=1, objective=PR_PAIRS_SQUARE)
    d = 3
    base = dict(dim=d, n_ctrl=2, H0=W.random_hermitian(d, rng, 2.0),
                Hc=np.stack([W.random_hermitian(d, rng, 1.0) for _ in range(2)]), objective=PR_GATE,
                U_target=W.haar_unitary(d, rng))
    per_slice = Problem(n_t=42, dt=0.025, dt_slices=rng.uniform(0.01, 0.04, 42), **base)
    weighted = Problem(n_t=42, dt=0.025, weight=np.array([0.5]), **base)
    d = 24
    big = Problem(dim=d, n_ctrl=2, n_t=20, dt=0.01, H0=W.random_hermitian(d, rng, 5.0),
                  Hc=np.stack([W.random_hermitian(d, rng, 2.0) for _ in range(2)]), objective=PR_GATE,
                  U_target=W.haar_unitary(d, rng))
    others = [(e, ue), (pairs, rng.uniform(-1, 1, (2, 30))), (lind, ul),
              (per_slice, rng.uniform(-1, 1, (2, 42))), (weighted, rng.uniform(-1, 1, (2, 42))),
              (big, rng.uniform(-1, 1, (2, 20)))]
    # the other problems sit between the packable ones
    mixed = batch[:7] + others[:3] + batch[7:14] + others[3:] + batch[14:]
    probs, us = [b[0] for b in mixed], [b[1] for b in mixed]
    F, G, D, c = evaluate(probs, us)
    assert packed_rows(c) >= 8
    check_oracle(probs, us, F, G)


def test_packed_exact_gradient_and_recompute():
    batch = small_gates(24, seed=9, n_t=(50, 53))
    probs, us = [b[0] for b in batch], [b[1] for b in batch]
    Fe, Ge, _, c = evaluate(probs, us, grad_mode=PR_GRAD_EXACT)
    Fu, Gu, _, _ = evaluate(probs, us, grad_mode=PR_GRAD_EXACT, pack=0)
    assert packed_rows(c) >= 10
    assert np.abs(Fe - Fu).max() <= 1e-13
    for a, b in zip(Ge, Gu):
        assert g_close(a, b, 1e-11)
    # without stored images the walk rebuilds the pack's propagators bit for bit
    for mode in (0, PR_GRAD_EXACT):
        Fs, Gs, _, _ = evaluate(probs, us, grad_mode=mode, store_props=1)
        Fr, Gr, _, _ = evaluate(probs, us, grad_mode=mode, store_props=0)
        assert np.array_equal(Fs, Fr)
        assert all(np.array_equal(a, b) for a, b in zip(Gs, Gr))
    check_oracle(probs, us, Fe, Ge, grad_mode=PR_GRAD_EXACT)


def test_packed_device_paths_agree_bit_for_bit():
    batch = small_gates(30, seed=12, n_t=(60, 62), dts=(0.02,))
    probs, us = [b[0] for b in batch], [b[1] for b in batch]
    plan = native.Plan(probs)
    try:
        F, G = plan.eval(us)
        F2, G2 = plan.eval(us)  # a replay of the captured graph
        assert np.array_equal(F, F2) and all(np.array_equal(a, b) for a, b in zip(G, G2))
        assert packed_rows(plan.exp_counts()) >= 15
        dc = torch.tensor(np.concatenate([controls_cm(u) for u in us]), dtype=torch.float64, device="cuda")
        gref = np.concatenate([controls_cm(g) for g in G])
        st = torch.cuda.current_stream().cuda_stream
        for f in (plan.eval_device, plan.eval_device_static, plan.eval_device):
            dF = torch.full((len(us),), np.nan, dtype=torch.float64, device="cuda")
            dG = torch.full((plan.n_params,), np.nan, dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            f(dc.data_ptr(), dF.data_ptr(), dG.data_ptr(), st)
            torch.cuda.synchronize()
            assert np.array_equal(dF.cpu().numpy(), F), f.__name__
            assert np.array_equal(dG.cpu().numpy(), gref), f.__name__
    finally:
        plan.close()


def test_packed_fp32_form_and_autotuned_plan():
    batch = small_gates(32, seed=21, n_t=(70, 74), dts=(0.02,))
    probs, us = [b[0] for b in batch], [b[1] for b in batch]
    Fm, Gm, _, c = evaluate(probs, us, precision=PR_MIXED)
    assert packed_rows(c) >= 10
    F0, G0, _, _ = evaluate(probs, us, pack=0)
    # the FP32 form's gate against FP64 (|dF| <= 1e-9, relative gradient 1e-6)
    assert np.abs(Fm - F0).max() <= 1e-9
    for a, b in zip(Gm, G0):
        assert g_close(a, b, 1e-6)
    Fa, Ga, _, c = evaluate(probs, us, segment=PR_SEGMENT_AUTOTUNE)
    assert packed_rows(c) >= 10
    check_oracle(probs, us, Fa, Ga)


def test_whole_uneven_batch_packed_against_unpacked_and_oracle():
    # the benchmarked 1024-problem batch (d in {2..64}, 200-2000 slices):
    # every problem's F and G packed against unpacked, and a sample of the
    # problems that run behind a longer one of their pack against the oracle
    batch = W.uneven_batch()
    probs, us = [b[0] for b in batch], [b[1] for b in batch]
    F1, G1, _, c1 = evaluate(probs, us)
    F0, G0, _, _ = evaluate(probs, us, pack=0)
    behind = np.flatnonzero(np.all(c1 == 0, axis=1))
    print("packed behind a pack's first problem:", behind.size, flush=True)
    assert behind.size > 500
    # full-length chains: the packed exps differ from the unpacked ones in
    # rounding only (shared scaling choices, block offsets inside the tile)
    assert np.abs(F1 - F0).max() <= TOL_F
    for g1, g0 in zip(G1, G0):
        assert g_close(g1, g0, TOL_G)
    sample = [int(i) for i in behind[:12]]
    print("oracle sample", [(probs[i].dim, probs[i].n_t) for i in sample], flush=True)
    check_oracle([probs[i] for i in sample], [us[i] for i in sample], F1[sample], [G1[i] for i in sample])


def small_states(n, seed, objective, dims=(1, 2, 3, 4, 5, 6, 7, 8), n_t=(40, 43), dt=0.02):
    """Independent state-pair problems: 1-3 pairs, 1-3 controls, some with
    control scales, some with an explicit weight."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        d = int(rng.choice(dims))
        nc = int(rng.integers(1, 4))
        nt = int(rng.integers(n_t[0], n_t[1] + 1))
        npairs = int(rng.integers(1, min(d, 3) + 1))
        kw = {}
        if i % 3 == 1:
            kw["ctrl_scale"] = rng.uniform(0.8, 1.2, (1, nc))
        if i % 4 == 2:
            kw["weight"] = np.array([0.7])
        herm = (lambda s: W.random_hermitian(d, rng, s)) if d > 1 else (lambda s: np.array([[0.9 * s + 0j]]))
        p = Problem(dim=d, n_ctrl=nc, n_t=nt, dt=dt, H0=herm(3.0), Hc=np.stack([herm(2.0) for _ in range(nc)]),
                    objective=objective, psi_init=np.stack([W.random_state(d, rng) for _ in range(npairs)]),
                    psi_targ=np.stack([W.random_state(d, rng) for _ in range(npairs)]), **kw)
        out.append((p, rng.uniform(-3, 3, (nc, nt))))
    return out


@pytest.mark.parametrize("objective,grad_mode", [(PR_PAIRS_SQUARE, 0), (PR_PAIRS_REAL, 0), (PR_PAIRS_MODULUS, 0),
                                                 (PR_STATE, 0), (PR_PAIRS_SQUARE, PR_GRAD_EXACT)])
def test_packed_state_pair_objectives(objective, grad_mode):
    batch = small_states(36, seed=40 + objective, objective=objective)
    probs, us = [b[0] for b in batch], [b[1] for b in batch]
    F1, G1, _, c1 = evaluate(probs, us, grad_mode=grad_mode)
    F0, G0, _, _ = evaluate(probs, us, grad_mode=grad_mode, pack=0)
    assert packed_rows(c1) >= 12, c1
    assert np.abs(F1 - F0).max() <= 1e-13
    for g1, g0 in zip(G1, G0):
        assert g_close(g1, g0, 1e-11)
    check_oracle(probs, us, F1, G1, grad_mode=grad_mode)


def test_packs_keep_objectives_apart():
    # gate, pairs and state problems of one plan, interleaved: each objective
    # packs with its own kind only
    mixed = []
    for a, b, c in zip(small_gates(12, seed=51, n_t=(40, 43), dts=(0.02,)),
                       small_states(12, seed=52, objective=PR_PAIRS_SQUARE),
                       small_states(12, seed=53, objective=PR_STATE)):
        mixed += [a, b, c]
    probs, us = [b[0] for b in mixed], [b[1] for b in mixed]
    F, G, _, c = evaluate(probs, us)
    assert packed_rows(c) >= 15
    check_oracle(probs, us, F, G)


def small_ensemble(M, d, seed, objective=PR_GATE, npairs=1, per_member_drift=True, weights=False, nc=2, n_t=45,
                   per_slice_dt=False, dt=0.02):
    rng = np.random.default_rng(seed)
    herm = (lambda s: W.random_hermitian(d, rng, s)) if d > 1 else (lambda s: np.array([[0.9 * s + 0j]]))
    kw = dict(ctrl_scale=rng.uniform(0.8, 1.2, (M, nc)))
    if weights:
        kw["weight"] = rng.uniform(0.1, 1.0, M)
    if objective == PR_GATE:
        kw["U_target"] = W.haar_unitary(d, rng)
    else:
        kw["psi_init"] = np.stack([W.random_state(d, rng) for _ in range(npairs)])
        kw["psi_targ"] = np.stack([W.random_state(d, rng) for _ in range(npairs)])
    if per_slice_dt:
        kw["dt_slices"] = rng.uniform(0.01, 0.03, n_t)
    H0 = np.stack([herm(3.0) for _ in range(M)]) if per_member_drift else herm(3.0)
    p = Problem(dim=d, n_ctrl=nc, n_t=n_t, dt=dt, H0=H0, Hc=np.stack([herm(2.0) for _ in range(nc)]),
                objective=objective, n_members=M, **kw)
    return p, rng.uniform(-3, 3, (nc, n_t))


@pytest.mark.parametrize("case", ["gate-d2-M37", "gate-d5-M7-weights", "gate-d1-M20", "square-d3-M10-2pairs",
                                  "real-d4-M9", "state-d2-M16-shared-drift", "modulus-d2-M11", "exact-gate-d3-M13",
                                  "exact-square-d2-M9", "real-d2-M12-dtslices"])
def test_packed_ensemble_members(case):
    # ensemble members as diagonal blocks of a wider ensemble (the last
    # packed member filled with weight-0 copies), against the unpacked plan
    # and the oracle
    obj = {"gate": PR_GATE, "square": PR_PAIRS_SQUARE, "real": PR_PAIRS_REAL, "state": PR_STATE,
           "modulus": PR_PAIRS_MODULUS}
    parts = case.split("-")
    mode = PR_GRAD_EXACT if parts[0] == "exact" else 0
    kind = parts[1] if mode else parts[0]
    d = int([x for x in parts if x.startswith("d")][0][1:])
    M = int([x for x in parts if x.startswith("M")][0][1:])
    p, u = small_ensemble(M, d, seed=M * 10 + d, objective=obj[kind], npairs=2 if "2pairs" in parts else 1,
                          per_member_drift="shared" not in parts, weights="weights" in parts,
                          per_slice_dt="dtslices" in parts)
    F1, G1, D1, c1 = evaluate([p], [u], grad_mode=mode)
    F0, G0, D0, c0 = evaluate([p], [u], grad_mode=mode, pack=0)
    # exps run: the packed members' slices, not every member's
    assert c1[0, :2].sum() < c0[0, :2].sum(), (c1, c0)
    assert abs(F1[0] - F0[0]) <= 1e-13
    assert g_close(G1[0], G0[0], 1e-11)
    check_oracle([p], [u], F1, G1, grad_mode=mode)


def lindblad_spin(offsets, pwr, fidelity, n_t=40, rate=3.0):
    """A relaxing spin 1/2 (4 x 4 Liouvillians, LindbladMRControl) over the
    given offsets and power levels: one member per (offset, power)."""
    from pulsar_rocm import mr
    sx = np.array([[0, 1], [1, 0]], complex) / 2
    sy = np.array([[0, -1j], [1j, 0]], complex) / 2
    sz = np.diag([0.5 + 0j, -0.5])
    return mr.LindbladMRControl(drifts=[2 * np.pi * o * sz for o in offsets], operators=[sx, sy],
                                rho_init=[np.array([[1, 0], [0, 0]], complex)],
                                rho_targ=[np.array([[0, 0], [0, 1]], complex)], pulse_dt=np.full(n_t, 0.01),
                                jump_ops=[np.array([[0, 1], [0, 0]], complex)], decay_rates=[rate],
                                pwr_levels=list(pwr), fidelity=fidelity).problem()


@pytest.mark.parametrize("fidelity", ["square", "real"])
def test_packed_lindblad_ensemble_and_problems(fidelity):
    # a 15-member ensemble (5 offsets x 3 power levels, per-slice dt) packs 4
    # members to a block problem; 10 one-member problems (uniform dt) pack 4
    # to a problem
    import dataclasses
    rng = np.random.default_rng(70)
    ens = lindblad_spin(np.linspace(-20, 20, 5), 2 * np.pi * np.array([8.0, 10.0, 12.0]), fidelity)
    singles = []
    for o in rng.uniform(-30, 30, 10):
        q = lindblad_spin([o], [2 * np.pi * 10.0], fidelity, n_t=int(rng.integers(40, 43)),
                          rate=float(rng.uniform(1, 5)))
        singles.append(dataclasses.replace(q, dt_slices=None, dt=0.01))
    probs = [ens] + singles
    us = [rng.uniform(-1, 1, (2, q.n_t)) for q in probs]
    F1, G1, _, c1 = evaluate(probs, us)
    F0, G0, _, c0 = evaluate(probs, us, pack=0)
    assert c1[0, :2].sum() < c0[0, :2].sum()  # the ensemble's members were packed
    assert packed_rows(c1) >= 4  # and so were the one-member problems
    assert np.abs(F1 - F0).max() <= 1e-13
    for g1, g0 in zip(G1, G0):
        assert g_close(g1, g0, 1e-11)
    check_oracle(probs, us, F1, G1)


@pytest.mark.parametrize("d,M,npairs,weights,per_slice_dt", [(2, 12, 1, False, False), (3, 10, 2, True, False),
                                                            (4, 6, 3, False, True), (8, 4, 1, False, False)])
def test_packed_tracking_ensembles(d, M, npairs, weights, per_slice_dt):
    # tracking members packed f to a block member (f the largest divisor of M
    # that fits) with checkpoints blockdiag(rho_c, ..., rho_c), against the
    # unpacked plan and the tracking oracle (GRAPETracking.jl:162-263)
    from pulsar_rocm.problem import PR_TRACKING
    rng = np.random.default_rng(80 + d)
    nc, nt = 2, 36
    steps, w, targets = W.tracking_checkpoints(d, nt, 4, rng)
    kw = dict(weight=np.full(M, 0.3)) if weights else {}
    if per_slice_dt:
        kw["dt_slices"] = rng.uniform(0.05, 0.15, nt)
    p = Problem(dim=d, n_ctrl=nc, n_t=nt, dt=0.1, objective=PR_TRACKING, n_members=M,
                H0=np.stack([W.random_hermitian(d, rng, 3.0) for _ in range(M)]),
                Hc=np.stack([W.random_hermitian(d, rng, 1.0) for _ in range(nc)]),
                ctrl_scale=rng.uniform(0.8, 1.2, (M, 1)) * np.ones((1, nc)),
                psi_init=np.stack([W.random_state(d, rng) for _ in range(npairs)]),
                track_step=np.asarray(steps, np.int32), track_weight=w, track_target=targets, **kw)
    u = rng.uniform(-2, 2, (nc, nt))
    F1, G1, _, c1 = evaluate([p], [u])
    F0, G0, _, _ = evaluate([p], [u], pack=0)
    assert np.abs(F1 - F0).max() <= 1e-13
    assert g_close(G1[0], G0[0], 1e-11)
    Fo, Go = O.tracking_evaluate(p, u)
    assert abs(F1[0] - Fo) <= TOL_F
    assert g_close(G1[0], Go, TOL_G)


def test_packed_one_member_tracking_problems():
    # independent tracking problems (d = 1..5, 1-2 initial states, their own
    # checkpoints and weights) packed as blocks: each keeps its checkpoints
    # and its F splits by its initial states; against the unpacked plan and
    # the tracking oracle
    from pulsar_rocm.problem import PR_TRACKING
    rng = np.random.default_rng(91)
    probs, us = [], []
    for i in range(14):
        d = int(rng.integers(1, 6))
        nc, nt, npairs = int(rng.integers(1, 3)), int(rng.integers(36, 39)), int(rng.integers(1, 3))
        steps, w, targets = W.tracking_checkpoints(d, nt, int(rng.integers(1, 4)), rng)
        herm = (lambda s: W.random_hermitian(d, rng, s)) if d > 1 else (lambda s: np.array([[0.9 * s + 0j]]))
        kw = dict(weight=np.array([0.6])) if i % 4 == 1 else {}
        probs.append(Problem(dim=d, n_ctrl=nc, n_t=nt, dt=0.1, objective=PR_TRACKING, H0=herm(3.0),
                             Hc=np.stack([herm(1.0) for _ in range(nc)]),
                             psi_init=np.stack([W.random_state(d, rng) for _ in range(npairs)]),
                             track_step=np.asarray(steps, np.int32), track_weight=w, track_target=targets,
                             **kw))
        us.append(rng.uniform(-2, 2, (nc, nt)))
    F1, G1, _, c1 = evaluate(probs, us)
    F0, G0, _, _ = evaluate(probs, us, pack=0)
    assert np.abs(F1 - F0).max() <= 1e-13
    for g1, g0 in zip(G1, G0):
        assert g_close(g1, g0, 1e-11)
    for i, (p, u) in enumerate(zip(probs, us)):
        Fo, Go = O.tracking_evaluate(p, u)
        assert abs(F1[i] - Fo) <= TOL_F, (i, p.dim)
        assert g_close(G1[i], Go, TOL_G), (i, p.dim)


def test_packed_split16_form():
    # the FP16x2 split form on packed problems and a packed ensemble: each
    # pack bounds its generator by its largest block's, so no slice of these
    # (|X|_1 <= 0.018 per block, 0.03 the split bound) takes the FP64 redo,
    # which the sum over all blocks' channels would have sent there; against
    # FP64 unpacked at the split form's gate
    from pulsar_rocm.problem import PR_SPLIT16
    batch = small_gates(32, seed=31, n_t=(80, 84), dts=(0.0005,))
    probs, us = [b[0] for b in batch], [b[1] for b in batch]
    e, ue = small_ensemble(37, 2, seed=32, dt=0.0005, n_t=80)
    probs.append(e)
    us.append(ue)
    Fs, Gs, _, cs = evaluate(probs, us, precision=PR_SPLIT16)
    assert packed_rows(cs) >= 10
    assert cs[:, 3].sum() == 0, cs[:, 3]
    F0, G0, _, c0 = evaluate(probs, us, pack=0)
    assert cs[-1, :2].sum() < c0[-1, :2].sum()  # the ensemble's members were packed
    assert np.abs(Fs - F0).max() <= 1e-9
    for a, b in zip(Gs, G0):
        assert g_close(a, b, 1e-6)
