"""Device (libpulsar_rocm.so, gfx950) vs the FP64 oracle on identical inputs.

FP64 device path tolerances: |dF| <= 1e-12 and max|dG| / max|G| <= 1e-10
(the device uses a degree-9 Taylor exp and a unitarity-based adjoint chain,
the oracle LAPACK-style eigen-exps and stored P/Q chains; both are exact to
rounding, ~1e-13 relative after 2000 slices).
"""
import os

import numpy as np
import pytest

import oracle as O
from pulsar_rocm import native, workloads as W
from pulsar_rocm.problem import (PR_GATE, PR_LINDBLAD, PR_PAIRS_REAL, PR_PAIRS_SQUARE, Problem)

pytestmark = pytest.mark.gpu

TOL_F = 1e-12
TOL_G = 1e-10
FIX = np.load(os.path.join(os.path.dirname(__file__), "golden", "fixtures.npz"))


@pytest.fixture(scope="module", autouse=True)
def device():
    native.init(0)


def run(p, u, **kw):
    plan = native.Plan([p], **kw)
    try:
        F, G, D = plan.eval([u], diag=True)
    finally:
        plan.close()
    return F[0], G[0], D[0]


def check(p, u, tol_f=TOL_F, tol_g=TOL_G, **kw):
    print(f"case d={p.dim} M={p.n_members} n_t={p.n_t} obj={p.objective} dyn={p.dynamics} {kw}", flush=True)
    F0, G0 = O.evaluate(p, u)
    F, G, D = run(p, u, **kw)
    dF = abs(F - F0)
    rel = np.abs(G - G0).max() / max(np.abs(G0).max(), 1e-300)
    assert dF <= tol_f, (F, F0)
    assert rel <= tol_g, rel
    return F, G, D


def test_four_qubit_ensemble_small():
    p, u = W.four_qubit_ensemble(n_t=64)
    p, u = W.shrink(p, u, n_members=4)
    F, G, D = check(p, u)
    assert D[0] < 1e-12  # unitarity defect of U_total


@pytest.mark.parametrize("seg", [1, 7, 64])
def test_segment_length_does_not_change_results(seg):
    p, u = W.four_qubit_ensemble(n_t=64)
    p, u = W.shrink(p, u, n_members=3)
    check(p, u, segment=seg)


def test_four_qubit_ensemble_full_length_member_sample():
    # all 2000 slices, 16 of the 256 members
    p, u = W.four_qubit_ensemble()
    p, u = W.shrink(p, u, n_members=16)
    check(p, u)


@pytest.mark.parametrize("d", [2, 3, 4, 8, 16, 20, 32, 48, 64])
def test_random_gate_problems(d):
    rng = np.random.default_rng(100 + d)
    H0 = W.random_hermitian(d, rng, 3.0)
    Hc = np.stack([W.random_hermitian(d, rng, 1.0) for _ in range(3)])
    u = rng.uniform(-1, 1, (3, 37))
    p = Problem(dim=d, n_ctrl=3, n_t=37, H0=H0, Hc=Hc, dt=0.05, objective=PR_GATE, n_members=2,
                ctrl_scale=rng.uniform(0.9, 1.1, (2, 3)), U_target=W.haar_unitary(d, rng))
    check(p, u)


def test_large_norm_uses_scaling_and_squaring():
    rng = np.random.default_rng(7)
    d = 8
    p = Problem(dim=d, n_ctrl=2, n_t=12, H0=W.random_hermitian(d, rng, 40.0),
                Hc=np.stack([W.random_hermitian(d, rng, 10.0) for _ in range(2)]), dt=0.1,
                objective=PR_GATE, U_target=W.haar_unitary(d, rng))
    u = rng.uniform(-3, 3, (2, 12))
    F, G, D = check(p, u)
    assert D[1] >= 3  # several squarings were needed
    assert D[2] == 9  # the largest degree used


def test_exp_counts_report_degree_and_squarings():
    # pr_plan_exp_counts: per problem, member-slices at degree 6, at degree 9, squarings.
    # The small-norm C2 slices all take degree 6; the large-norm problem takes
    # degree 9 with ceil(log2(||X||_1 / 0.11)) squarings per slice, which the
    # 1-norm of the host-built generator bounds from below
    rng = np.random.default_rng(7)
    d = 8
    big = Problem(dim=d, n_ctrl=2, n_t=12, H0=W.random_hermitian(d, rng, 40.0),
                  Hc=np.stack([W.random_hermitian(d, rng, 10.0) for _ in range(2)]), dt=0.1,
                  objective=PR_GATE, U_target=W.haar_unitary(d, rng))
    ub = rng.uniform(-3, 3, (2, 12))
    small, us = W.four_qubit_ensemble(n_t=40)
    small, us = W.shrink(small, us, n_members=3)
    plan = native.Plan([small, big])
    try:
        plan.eval([us, ub])
        c = plan.exp_counts()
    finally:
        plan.close()
    assert list(c[0]) == [3 * 40, 0, 0, 0]
    assert c[1][0] + c[1][1] == 12 and c[1][1] == 12
    lower = 0
    for k in range(12):
        X = 0.1 * (big.H0 + ub[0, k] * big.Hc[0] + ub[1, k] * big.Hc[1])
        n1 = np.abs(X).sum(axis=0).max()
        lower += max(0, int(np.ceil(np.log2(n1 / 0.11))))
    assert lower <= c[1][2] <= lower + 12


def test_gate_fixture_on_device():
    p = Problem(dim=4, n_ctrl=3, n_t=5, H0=FIX["gate_H0"], Hc=FIX["gate_Hc"], dt=float(FIX["gate_dt"]),
                objective=PR_GATE, n_members=2, ctrl_scale=FIX["gate_scale"], U_target=FIX["gate_Ut"])
    F, G, _ = run(p, FIX["gate_u"])
    assert abs(F - float(FIX["gate_F"])) < 1e-13
    assert np.abs(G - FIX["gate_G"]).max() / np.abs(FIX["gate_G"]).max() < 1e-11


def test_pairs_fixture_on_device():
    p = Problem(dim=2, n_ctrl=2, n_t=4, H0=FIX["pair_H0"], Hc=FIX["pair_Hc"], dt_slices=FIX["pair_dts"],
                objective=PR_PAIRS_SQUARE, n_members=1, ctrl_scale=np.full((1, 2), float(FIX["pair_pwr"])),
                psi_init=FIX["pair_psi0"], psi_targ=FIX["pair_psit"], weight=np.ones(1))
    F, G, _ = run(p, FIX["pair_u"])
    assert abs(F - float(FIX["pair_F"])) < 1e-13
    assert np.abs(G - FIX["pair_G"]).max() < 1e-12


def test_lindblad_fixture_on_device():
    p = Problem(dim=4, n_ctrl=2, n_t=4, H0=FIX["lind_L0"], Hc=FIX["lind_Lc"], dt_slices=FIX["lind_dts"],
                dynamics=PR_LINDBLAD, objective=PR_PAIRS_SQUARE, n_members=1,
                ctrl_scale=np.full((1, 2), float(FIX["lind_pwr"])), psi_init=FIX["lind_s0"],
                psi_targ=FIX["lind_st"], weight=np.ones(1))
    F, G, _ = run(p, FIX["lind_u"])
    assert abs(F - float(FIX["lind_F"])) < 1e-13
    assert np.abs(G - FIX["lind_G"]).max() < 1e-12


@pytest.mark.parametrize("psi_init", [None, [1, 0]])
def test_single_qubit_state(psi_init):
    p, u = W.single_qubit_state(psi_init=psi_init)
    check(p, u)


@pytest.mark.parametrize("obj", [PR_PAIRS_SQUARE, PR_PAIRS_REAL])
def test_state_pairs_per_slice_dt(obj):
    rng = np.random.default_rng(11)
    d, npairs = 6, 3
    p = Problem(dim=d, n_ctrl=2, n_t=25, H0=np.stack([W.random_hermitian(d, rng, 2.0) for _ in range(2)]),
                Hc=np.stack([W.random_hermitian(d, rng, 1.0) for _ in range(2)]), dt_slices=rng.uniform(0.01, 0.1, 25),
                objective=obj, n_members=2, ctrl_scale=np.array([[1.0, 1.0], [1.1, 1.1]]),
                psi_init=rng.standard_normal((npairs, d)) + 1j * rng.standard_normal((npairs, d)),
                psi_targ=rng.standard_normal((npairs, d)) + 1j * rng.standard_normal((npairs, d)))
    check(p, rng.uniform(-1, 1, (2, 25)))


@pytest.mark.parametrize("obj", [PR_PAIRS_REAL, PR_PAIRS_SQUARE])
def test_three_qubit_lindblad_small(obj):
    p, u = W.three_qubit_lindblad(n_t=40, n_members=2, objective=obj)
    check(p, u)


def test_three_qubit_lindblad_full_length_one_member():
    p, u = W.three_qubit_lindblad(n_members=1)
    check(p, u)


def test_three_qubit_lindblad_four_members_full_length():
    # the three-qubit Lindblad ensemble (64 x 64 superoperators) at 4 of its
    # 64 members and all 1000 slices
    p, u = W.three_qubit_lindblad(n_members=4)
    check(p, u)


def test_three_qubit_lindblad_whole_ensemble():
    # the benchmarked configuration itself: all 64 members, all 1000 slices
    # (the oracle takes about a minute on 16 host threads)
    p, u = W.three_qubit_lindblad()
    assert p.n_members == 64 and p.n_t == 1000
    check(p, u)


def test_uneven_batch_full_length_subset_matches_oracle():
    # 32 problems of the benchmarked 1024-problem uneven batch at full length
    # (the first 6 of d = 2 and 4, the first 5 of d = 8, 16, 32 and 64; 233 to
    # 1945 slices), reference gradients, all in one plan
    from concurrent.futures import ThreadPoolExecutor
    batch = W.uneven_batch()
    pick = []
    for d, n in ((2, 6), (4, 6), (8, 5), (16, 5), (32, 5), (64, 5)):
        pick += [i for i, b in enumerate(batch) if b[0].dim == d][:n]
    probs = [batch[i][0] for i in pick]
    us = [batch[i][1] for i in pick]
    print("uneven subset", [(q.dim, q.n_t) for q in probs], flush=True)
    plan = native.Plan(probs)
    try:
        F, G = plan.eval(us)
    finally:
        plan.close()
    small = [i for i, q in enumerate(probs) if q.dim < 32]
    with ThreadPoolExecutor(16) as ex:
        ref = dict(zip(small, ex.map(lambda i: O.evaluate(probs[i], us[i], n_threads=1), small)))
    for i, q in enumerate(probs):
        if q.dim >= 32:
            ref[i] = O.evaluate(q, us[i])  # one member: its slices in parallel
    for i, q in enumerate(probs):
        F0, G0 = ref[i]
        assert abs(F[i] - F0) <= TOL_F, (i, q.dim, q.n_t, F[i], F0)
        assert np.abs(G[i] - G0).max() / np.abs(G0).max() <= TOL_G, (i, q.dim, q.n_t)


def test_uneven_batch_in_one_plan():
    batch = W.uneven_batch(n_problems=24, seed=21, n_min=20, n_max=60)
    probs = [b[0] for b in batch]
    us = [b[1] for b in batch]
    plan = native.Plan(probs)
    try:
        F, G = plan.eval(us)
    finally:
        plan.close()
    for i, (p, u) in enumerate(batch):
        F0, G0 = O.evaluate(p, u)
        assert abs(F[i] - F0) <= TOL_F, (i, p.dim)
        assert np.abs(G[i] - G0).max() / np.abs(G0).max() <= TOL_G, (i, p.dim)


def test_edge_cases_single_slice_zero_dt_zero_controls():
    rng = np.random.default_rng(3)
    d = 5
    base = dict(dim=d, n_ctrl=2, H0=W.random_hermitian(d, rng, 1.0),
                Hc=np.stack([W.random_hermitian(d, rng, 1.0) for _ in range(2)]), objective=PR_GATE,
                U_target=W.haar_unitary(d, rng))
    check(Problem(n_t=1, dt=0.3, **base), rng.uniform(-1, 1, (2, 1)))
    check(Problem(n_t=9, dt=0.2, **base), np.zeros((2, 9)))
    # dt = 0: every propagator is the identity
    F, G, _ = run(Problem(n_t=3, dt=0.0, **base), rng.uniform(-1, 1, (2, 3)))
    assert abs(F - abs(np.trace(base["U_target"]) / d) ** 2) < 1e-14
    assert np.all(G == 0)


def test_device_resident_path_matches_host_path():
    p, u = W.four_qubit_ensemble(n_t=50)
    p, u = W.shrink(p, u, n_members=5)
    plan = native.Plan([p])
    try:
        F, G = plan.eval([u])
        plan.upload_controls([u])
        dc, dF, dG = plan.buffers()
        plan.eval_device(dc, dF, dG)
        plan.synchronize()
        F2, G2 = plan.eval([u])
        assert F2[0] == F[0] and np.array_equal(G2[0], G[0])  # deterministic
    finally:
        plan.close()


def test_host_buffer_ranges_match_the_device_resident_evaluation():
    # 256 problems of the uneven batch: pr_eval packs, uploads, downloads and
    # unpacks their controls and gradients in several ranges on host threads;
    # its F and G equal the device-resident evaluation's device buffers
    import ctypes
    from pulsar_rocm.problem import controls_cm
    batch = W.uneven_batch(n_problems=256, seed=8)
    probs, us = [b[0] for b in batch], [b[1] for b in batch]
    plan = native.Plan(probs)
    try:
        assert plan.n_params > 1 << 19  # more than one range of 2 MiB
        F, G = plan.eval(us)
        plan.upload_controls(us)
        dc, dF, dG = plan.buffers()
        plan.eval_device(dc, dF, dG)
        plan.synchronize()
        hip = ctypes.CDLL("libamdhip64.so")
        Fd = np.zeros(len(probs))
        Gd = np.zeros(plan.n_params)
        assert hip.hipMemcpy(Fd.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(dF), ctypes.c_size_t(Fd.nbytes), 2) == 0
        assert hip.hipMemcpy(Gd.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(dG), ctypes.c_size_t(Gd.nbytes), 2) == 0
        assert np.array_equal(F, Fd)
        off = 0
        for p, g in zip(probs, G):
            n = p.n_ctrl * p.n_t
            assert np.array_equal(controls_cm(g), Gd[off:off + n])
            off += n
        assert off == plan.n_params
    finally:
        plan.close()


def test_segment_autotuner_plan_matches_oracle():
    # pr_options.segment = PR_SEGMENT_AUTOTUNE: candidate segment lengths are
    # timed per dimension group at plan creation; the plan it keeps evaluates
    # like any other (the oracle's values, deterministic replays)
    from pulsar_rocm.problem import PR_SEGMENT_AUTOTUNE
    batch = W.uneven_batch(n_problems=24, seed=31, n_min=40, n_max=160)
    probs, us = [b[0] for b in batch], [b[1] for b in batch]
    plan = native.Plan(probs, segment=PR_SEGMENT_AUTOTUNE, pack=0)  # (packed plans: tests/test_gpu_pack.py)
    try:
        F, G, D = plan.eval(us, diag=True)
        F2, G2 = plan.eval(us)
        S = plan.shapes()
    finally:
        plan.close()
    assert np.array_equal(F, F2) and all(np.array_equal(a, b) for a, b in zip(G, G2))
    print("autotuned segments per problem", D[:, 3].astype(int).tolist(), flush=True)
    print("shapes (slices/segment, segments, trace regs, split passes)", S.tolist(), flush=True)
    # the reported shape is the plan's: segments per member as in the diagnostics,
    # and segments x slices covering each problem's slices
    assert np.array_equal(S[:, 1], D[:, 3].astype(int))
    for (p, _), sh in zip(batch, S):
        assert sh[0] * (sh[1] - 1) < p.n_t <= sh[0] * sh[1]
        assert sh[2] in (0, 1) and (p.dim <= 16 or sh[2] == 0) and sh[3] == 0
    for i, (p, u) in enumerate(batch):
        F0, G0 = O.evaluate(p, u)
        assert abs(F[i] - F0) <= TOL_F, (i, p.dim)
        assert np.abs(G[i] - G0).max() / np.abs(G0).max() <= TOL_G, (i, p.dim)


def test_graph_replays_match_direct_launches():
    # evaluations replay a captured HIP graph; with per-kernel timing on, the
    # same evaluation launches its kernels directly. Several replays with new
    # controls, device-path replays and direct launches in between.
    p, u = W.four_qubit_ensemble(n_t=64)
    p, u = W.shrink(p, u, n_members=4)
    q, uq = W.three_qubit_lindblad(n_t=12, n_members=2)
    plan = native.Plan([p, q])
    rng = np.random.default_rng(8)
    try:
        dc, dF, dG = plan.buffers()
        for rep in range(4):
            us = [u * rng.uniform(0.5, 1.5), uq * rng.uniform(0.5, 1.5)]
            plan.set_timing(True)
            F0, G0 = plan.eval(us)
            Fn, _ = plan.eval(us, grad=False)
            plan.kernel_times()
            plan.set_timing(False)
            F1, G1 = plan.eval(us)
            assert np.array_equal(F1, F0)
            assert all(np.array_equal(a, b) for a, b in zip(G1, G0))
            F2, _ = plan.eval(us, grad=False)
            assert np.array_equal(F2, Fn)
            plan.upload_controls(us)
            plan.eval_device(dc, dF, dG)
            plan.synchronize()
        F3, G3 = plan.eval(us)
        assert np.array_equal(F3, F0) and all(np.array_equal(a, b) for a, b in zip(G3, G0))
        F4, G4 = O.evaluate(p, us[0])
        assert abs(F3[0] - F4) < TOL_F
    finally:
        plan.close()


def test_ensemble_weights_and_linearity():
    p, u = W.four_qubit_ensemble(n_t=30)
    p, u = W.shrink(p, u, n_members=6)
    F, G, _ = run(p, u)
    p2, _ = W.shrink(p, u)
    p2.weight = np.full(6, 2.0 / 6)
    F2, G2, _ = run(p2, u)
    assert abs(F2 - 2 * F) < 1e-15 and np.abs(G2 - 2 * G).max() <= 1e-15 * np.abs(G).max() * 4


@pytest.mark.parametrize("case", ["gate4", "gate16", "gate32", "pairs", "lindblad"])
def test_exact_gradient_matches_oracle(case):
    rng = np.random.default_rng(31)
    if case.startswith("gate"):
        d = int(case[4:])
        p = Problem(dim=d, n_ctrl=2, n_t=9, H0=W.random_hermitian(d, rng, 3.0),
                    Hc=np.stack([W.random_hermitian(d, rng, 1.0) for _ in range(2)]), dt=0.08,
                    objective=PR_GATE, n_members=2, ctrl_scale=rng.uniform(0.9, 1.1, (2, 2)),
                    U_target=W.haar_unitary(d, rng))
        u = rng.uniform(-2, 2, (2, 9))
    elif case == "pairs":
        d = 5
        p = Problem(dim=d, n_ctrl=2, n_t=11, H0=W.random_hermitian(d, rng, 3.0),
                    Hc=np.stack([W.random_hermitian(d, rng, 1.0) for _ in range(2)]), dt=0.1,
                    objective=PR_PAIRS_SQUARE, psi_init=rng.standard_normal((2, d)) + 0j,
                    psi_targ=rng.standard_normal((2, d)) + 1j * rng.standard_normal((2, d)))
        u = rng.uniform(-2, 2, (2, 11))
    else:
        p, u = W.three_qubit_lindblad(n_t=12, n_members=2, objective=PR_PAIRS_SQUARE)
    F0, G0 = O.evaluate(p, u, grad_mode=1)
    F, G, _ = run(p, u, grad_mode=1)
    assert abs(F - F0) <= TOL_F
    assert np.abs(G - G0).max() / np.abs(G0).max() <= TOL_G


def test_exact_gradient_finite_differences_uneven_subset():
    # 32 problems of the uneven batch, central differences (step 1e-6) on 8
    # sampled controls per problem, all problems in one plan per evaluation
    batch = W.uneven_batch(n_problems=32, seed=5)
    probs = [b[0] for b in batch]
    us = [b[1] for b in batch]
    plan_x = native.Plan(probs, grad_mode=1)
    plan_f = native.Plan(probs)
    try:
        _, Gx = plan_x.eval(us)
        rng = np.random.default_rng(0)
        picks = [[(int(rng.integers(p.n_ctrl)), int(rng.integers(p.n_t))) for _ in range(8)] for p in probs]
        eps = 1e-6
        fd = np.zeros((32, 8))
        for t in range(8):
            up = [u.copy() for u in us]
            um = [u.copy() for u in us]
            for i in range(32):
                j, k = picks[i][t]
                up[i][j, k] += eps
                um[i][j, k] -= eps
            Fp, _ = plan_f.eval(up, grad=False)
            Fm, _ = plan_f.eval(um, grad=False)
            fd[:, t] = (Fp - Fm) / (2 * eps)
        for i in range(32):
            gx = np.array([Gx[i][j, k] for j, k in picks[i]])
            scale = np.abs(Gx[i]).max()
            assert np.abs(gx - fd[i]).max() <= 1e-6 * scale + 1e-12, (i, probs[i].dim)
    finally:
        plan_x.close()
        plan_f.close()


# ------------------------------------------------------------------ 64 < d <= 128


def test_spin_chain_128_small():
    # the 7-qubit Heisenberg chain (d = 128), 2 members, 24 slices
    p, u = W.spin_chain_128(n_t=24, n_members=2)
    F, G, D = check(p, u)
    assert D[0] < 1e-12
    assert D[1] >= 1  # generators above the degree-9 bound: scaling and squaring


def test_spin_chain_128_quarter_of_the_ensemble():
    # the first 16 of the benchmarked 64 members (control scales 0.95-0.974),
    # all 500 slices of the d = 128 chain (about 30 s of oracle on 16 host
    # threads)
    p, u = W.spin_chain_128()
    q, uq = W.shrink(p, u, n_members=16)
    F, G, D = check(q, uq)
    assert D[0] < 1e-12


def test_spin_chain_128_two_members_500_slices():
    # the 7-qubit Heisenberg chain (d = 128) at 2 of its 64 members and all 500
    # slices, with stored U_k images and without them (store_props = 0: the
    # gradient walk recomputes U_k from segment checkpoints of <= 32 slices)
    p, u = W.spin_chain_128(n_members=2)
    print(f"case spin chain d={p.dim} M={p.n_members} n_t={p.n_t}", flush=True)
    F0, G0 = O.evaluate(p, u)
    for sp in (1, 0):
        F, G, D = run(p, u, store_props=sp)
        print("store_props", sp, "segments", D[3], "dF", abs(F - F0), flush=True)
        assert abs(F - F0) <= TOL_F
        assert np.abs(G - G0).max() / np.abs(G0).max() <= TOL_G
        assert D[0] < 1e-12
        if sp == 0:
            assert D[3] >= 16  # checkpoints every 32 slices at most


@pytest.mark.parametrize("seg", [32, 7])
def test_spin_chain_128_recomputed_propagators_equal_stored_images(seg):
    # same segmentation in both plans: the recomputed U_k are the stored ones
    # bit for bit, so F and G are identical
    p, u = W.spin_chain_128(n_members=2, n_t=70)
    out = []
    for sp in (1, 0):
        plan = native.Plan([p], store_props=sp, segment=seg)
        try:
            out.append(plan.eval([u]))
        finally:
            plan.close()
    (F1, G1), (F0, G0) = out
    assert F0[0] == F1[0]
    assert np.array_equal(G0[0], G1[0])


@pytest.mark.parametrize("d,seg", [(65, 0), (80, 3), (100, 0), (128, 5)])
def test_random_gate_problems_beyond_64(d, seg):
    rng = np.random.default_rng(200 + d)
    H0 = W.random_hermitian(d, rng, 3.0)
    Hc = np.stack([W.random_hermitian(d, rng, 1.0) for _ in range(2)])
    u = rng.uniform(-1, 1, (2, 11))
    p = Problem(dim=d, n_ctrl=2, n_t=11, H0=H0, Hc=Hc, dt=0.05, objective=PR_GATE, n_members=2,
                ctrl_scale=rng.uniform(0.9, 1.1, (2, 2)), U_target=W.haar_unitary(d, rng))
    check(p, u, segment=seg)


def test_mixed_dimensions_up_to_128_in_one_plan():
    rng = np.random.default_rng(5)
    probs, us = [], []
    for d in (16, 64, 96):
        H0 = W.random_hermitian(d, rng, 3.0)
        Hc = np.stack([W.random_hermitian(d, rng, 1.0) for _ in range(2)])
        probs.append(Problem(dim=d, n_ctrl=2, n_t=7, H0=H0, Hc=Hc, dt=0.05, objective=PR_GATE,
                             U_target=W.haar_unitary(d, rng)))
        us.append(rng.uniform(-1, 1, (2, 7)))
    plan = native.Plan(probs)
    try:
        F, G = plan.eval(us)
    finally:
        plan.close()
    for q, uu, f, g in zip(probs, us, F, G):
        F0, G0 = O.evaluate(q, uu)
        assert abs(f - F0) <= TOL_F
        assert np.abs(g - G0).max() <= TOL_G * np.abs(G0).max()


@pytest.mark.parametrize("d,nc", [(4, 40), (8, 33), (16, 70), (24, 5), (40, 12), (40, 20)])
def test_many_controls_general_contraction(d, nc):
    # more controls than one lane group holds (32 groups), dense controls, and
    # two- and three-block problems: the general trace path and its reduction
    # widths (d = 40 with 16 or 32 lane groups: 576 threads, groups of 32 or 16)
    rng = np.random.default_rng(400 + d + nc)
    H0 = W.random_hermitian(d, rng, 2.0)
    Hc = np.stack([W.random_hermitian(d, rng, 0.3) for _ in range(nc)])
    u = rng.uniform(-1, 1, (nc, 9))
    p = Problem(dim=d, n_ctrl=nc, n_t=9, H0=H0, Hc=Hc, dt=0.05, objective=PR_GATE, n_members=2,
                U_target=W.haar_unitary(d, rng))
    check(p, u)
    check(p, u, segment=1)


def test_mixed_plan_with_large_generators_falls_back_per_slice():
    # every slice above the degree-6 bound: the mixed kernels take their FP64
    # fallback for each slice and keep FP64 accuracy
    from pulsar_rocm.problem import PR_MIXED
    rng = np.random.default_rng(77)
    d = 12
    p = Problem(dim=d, n_ctrl=3, n_t=30, H0=W.random_hermitian(d, rng, 20.0),
                Hc=np.stack([W.random_hermitian(d, rng, 5.0) for _ in range(3)]), dt=0.02,
                objective=PR_GATE, n_members=3, ctrl_scale=rng.uniform(0.9, 1.1, (3, 3)),
                U_target=W.haar_unitary(d, rng))
    u = rng.uniform(-1, 1, (3, 30))
    check(p, u, precision=PR_MIXED)
    check(p, u, precision=2)  # PR_SPLIT16


def _random_state(d, rng):
    v = rng.standard_normal(d) + 1j * rng.standard_normal(d)
    return v / np.linalg.norm(v)


def _random_gate(d, nc, n_t, M, seed, objective=PR_GATE, n_pairs=0):
    rng = np.random.default_rng(seed)
    if objective != PR_GATE:
        kw = dict(psi_init=np.stack([_random_state(d, rng) for _ in range(n_pairs)]),
                  psi_targ=np.stack([_random_state(d, rng) for _ in range(n_pairs)]))
    else:
        kw = dict(U_target=W.haar_unitary(d, rng))
    p = Problem(dim=d, n_ctrl=nc, n_t=n_t, H0=W.random_hermitian(d, rng, 2.0),
                Hc=np.stack([W.random_hermitian(d, rng, 0.5) for _ in range(nc)]), dt=0.05,
                objective=objective, n_members=M, ctrl_scale=rng.uniform(0.9, 1.1, (M, nc)), **kw)
    return p, rng.uniform(-1, 1, (nc, n_t))


def _same(plans_args):
    """F and G of the same problems under store_props = 1 and 0."""
    out = []
    for sp in (1, 0):
        probs, us, kw = plans_args
        plan = native.Plan(probs, store_props=sp, **kw)
        try:
            out.append(plan.eval(us))
        finally:
            plan.close()
    (F1, G1), (F0, G0) = out
    for a, b in zip(F0, F1):
        assert a == b
    for a, b in zip(G0, G1):
        assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()


@pytest.mark.parametrize("d,objective,grad_mode", [(16, PR_GATE, 0), (24, PR_PAIRS_SQUARE, 0),
                                                    (40, PR_GATE, 1), (7, PR_PAIRS_REAL, 1)])
def test_recomputed_propagators_match_stored(d, objective, grad_mode):
    # store_props = 0: no U_k images; the gradient walk recomputes U_k with the
    # segment kernel's own calls, so F and G equal the stored-image plan's and
    # the oracle's
    p, u = _random_gate(d, 3, 37, 3, 500 + d, objective=objective, n_pairs=2)
    print(f"case d={d} objective={objective} grad_mode={grad_mode}", flush=True)
    _same(([p], [u], dict(grad_mode=grad_mode)))
    F0, G0, _ = run(p, u, store_props=0, grad_mode=grad_mode)
    Fo, Go = O.evaluate(p, u, grad_mode=grad_mode)
    assert abs(F0 - Fo) <= TOL_F
    assert np.abs(G0 - Go).max() <= TOL_G * np.abs(Go).max()


def test_recompute_with_other_problems_in_the_plan():
    # under store_props = 0 the closed d = 20 problem, the Lindblad problem and
    # the FP32-form d <= 16 problem all recompute; results equal the
    # stored-image plans
    from pulsar_rocm.problem import PR_MIXED
    lind = Problem(dim=4, n_ctrl=2, n_t=4, H0=FIX["lind_L0"], Hc=FIX["lind_Lc"], dt_slices=FIX["lind_dts"],
                   dynamics=PR_LINDBLAD, objective=PR_PAIRS_SQUARE, n_members=1,
                   ctrl_scale=np.full((1, 2), float(FIX["lind_pwr"])), psi_init=FIX["lind_s0"],
                   psi_targ=FIX["lind_st"], weight=np.ones(1))
    pg, ug = _random_gate(20, 2, 25, 2, 92)
    pm, um = _random_gate(8, 2, 25, 2, 93)
    _same(([pg, lind], [ug, FIX["lind_u"]], dict()))
    _same(([pg, pm], [ug, um], dict(precision=PR_MIXED)))


def _one_qubit_lindblad(n_members, n_t, seed):
    """The fixture's one-qubit Liouvillians (4 x 4) over n_t slices of random
    length, members with their own drift scale and control scale."""
    rng = np.random.default_rng(seed)
    L0 = np.stack([FIX["lind_L0"] * rng.uniform(0.9, 1.1) for _ in range(n_members)])
    p = Problem(dim=4, n_ctrl=2, n_t=n_t, H0=L0, Hc=FIX["lind_Lc"], dt_slices=rng.uniform(1e-3, 3e-3, n_t),
                dynamics=PR_LINDBLAD, objective=PR_PAIRS_SQUARE, n_members=n_members,
                ctrl_scale=rng.uniform(0.9, 1.1, (n_members, 2)) * float(FIX["lind_pwr"]),
                psi_init=FIX["lind_s0"], psi_targ=FIX["lind_st"])
    return p, rng.uniform(-1, 1, (2, n_t))


@pytest.mark.parametrize("case", ["d64_windows", "d4_windows", "d64_exact"])
def test_lindblad_recompute_matches_stored(case):
    # store_props = 0 for Lindblad dynamics: no U_k images; the gradient walk
    # recomputes U_k on its way forward into a scratch slot per work item and
    # is launched over windows of consecutive work items (512 at d = 64, 8192
    # at d <= 16), so the slots are reused. With the same segments F and G
    # equal the stored-image plan's bit for bit.
    kw = {}
    if case == "d64_windows":
        p, u = W.three_qubit_lindblad(n_members=20, n_t=1000)
        seg = 8  # 20 x 125 = 2500 work items: 5 windows
    elif case == "d4_windows":
        p, u = _one_qubit_lindblad(64, 1000, seed=5)
        seg = 4  # 64 x 250 = 16000 work items: 2 windows
    else:
        p, u = W.three_qubit_lindblad(n_members=2, n_t=60)
        seg, kw = 5, dict(grad_mode=1)
    print(f"case {case}: d={p.dim} M={p.n_members} n_t={p.n_t} segment={seg} {kw}", flush=True)
    out = []
    for sp in (1, 0):
        plan = native.Plan([p], store_props=sp, segment=seg, **kw)
        try:
            out.append(plan.eval([u]))
        finally:
            plan.close()
    (F1, G1), (F0, G0) = out
    assert F0[0] == F1[0]
    assert np.array_equal(G0[0], G1[0])
    if case == "d64_exact":
        Fo, Go = O.evaluate(p, u, grad_mode=1)
        assert abs(F0[0] - Fo) <= TOL_F
        assert np.abs(G0[0] - Go).max() <= TOL_G * np.abs(Go).max()


def test_lindblad_recompute_default_segments_match_oracle():
    # store_props = 0 with the library's own segmentation (at most 32 slices
    # per segment without images): the three-qubit Lindblad ensemble at 2 of
    # its members and all 1000 slices against the oracle
    p, u = W.three_qubit_lindblad(n_members=2)
    check(p, u, store_props=0)


def test_closed_and_lindblad_problems_in_one_single_block_group():
    # both d <= 16: the group runs the general segment kernel (Lindblad
    # generators), where the closed problem takes the anti-Hermitian exp and
    # the LDS transpose for its images; under store_props = 0 it keeps them
    lind = Problem(dim=4, n_ctrl=2, n_t=4, H0=FIX["lind_L0"], Hc=FIX["lind_Lc"], dt_slices=FIX["lind_dts"],
                   dynamics=PR_LINDBLAD, objective=PR_PAIRS_SQUARE, n_members=1,
                   ctrl_scale=np.full((1, 2), float(FIX["lind_pwr"])), psi_init=FIX["lind_s0"],
                   psi_targ=FIX["lind_st"], weight=np.ones(1))
    pg, ug = _random_gate(8, 3, 29, 2, 94)
    for sp in (1, 0):
        plan = native.Plan([pg, lind], store_props=sp)
        try:
            F, G = plan.eval([ug, FIX["lind_u"]])
        finally:
            plan.close()
        F0, G0 = O.evaluate(pg, ug)
        assert abs(F[0] - F0) <= TOL_F
        assert np.abs(G[0] - G0).max() <= TOL_G * np.abs(G0).max()
        assert abs(F[1] - float(FIX["lind_F"])) < 1e-13
        assert np.abs(G[1] - FIX["lind_G"]).max() < 1e-12


@pytest.mark.parametrize("d,M,nt,npairs", [(2, 1, 100, 1), (7, 3, 40, 2), (16, 4, 60, 3), (24, 2, 30, 2), (40, 1, 12, 1)])
def test_state_objective_compute_grape_gradient_convention(d, M, nt, npairs):
    # PR_STATE: compute_grape_gradient with a state target (Gradients.jl:357-399),
    # the first-order term after U_k; the same states under the MR convention
    # (PR_PAIRS_SQUARE) give the same F and a different gradient
    from pulsar_rocm.problem import PR_STATE
    rng = np.random.default_rng(d + nt)
    H0 = np.stack([W.random_hermitian(d, rng, 3.0) for _ in range(M)])
    Hc = np.stack([W.random_hermitian(d, rng, 1.0) for _ in range(3)])
    psi_i = np.stack([_random_state(d, rng) for _ in range(npairs)])
    psi_t = np.stack([_random_state(d, rng) for _ in range(npairs)])
    u = rng.uniform(-2, 2, (3, nt))
    p = Problem(dim=d, n_ctrl=3, n_t=nt, H0=H0, Hc=Hc, dt=0.1, objective=PR_STATE, n_members=M,
                ctrl_scale=rng.uniform(0.9, 1.1, (M, 3)), psi_init=psi_i, psi_targ=psi_t)
    F, G, _ = check(p, u)
    q = Problem(**{**p.__dict__, "objective": PR_PAIRS_SQUARE, "_keep": []})
    Fq, Gq, _ = run(q, u)
    assert abs(Fq - F) <= 1e-13 and np.abs(Gq - G).max() > 1e-6 * np.abs(G).max()


def test_state_objective_mixed_precision():
    from pulsar_rocm.problem import PR_MIXED, PR_STATE
    p, u = W.single_qubit_state(psi_init=[1, 0])
    assert p.objective == PR_STATE
    F0, G0 = O.evaluate(p, u)
    for prec in (PR_MIXED, 2):  # FP32 and FP16x2 split products
        F, G, _ = run(p, u, precision=prec)
        assert abs(F - F0) <= 1e-9 and np.abs(G - G0).max() <= 1e-6 * np.abs(G0).max(), prec


def test_state_objective_exact_gradient_matches_oracle():
    from pulsar_rocm.problem import PR_GRAD_EXACT, PR_STATE
    rng = np.random.default_rng(77)
    d = 8
    p = Problem(dim=d, n_ctrl=2, n_t=30, H0=W.random_hermitian(d, rng, 3.0),
                Hc=np.stack([W.random_hermitian(d, rng, 1.0) for _ in range(2)]), dt=0.1, objective=PR_STATE,
                psi_init=_random_state(d, rng)[None], psi_targ=_random_state(d, rng)[None])
    u = rng.uniform(-2, 2, (2, 30))
    F0, G0 = O.evaluate(p, u, grad_mode=PR_GRAD_EXACT)
    F, G, _ = run(p, u, grad_mode=PR_GRAD_EXACT)
    assert abs(F - F0) <= TOL_F and np.abs(G - G0).max() <= TOL_G * np.abs(G0).max()


def test_propagator_helper_objective_fp64_mixed_and_recompute():
    from pulsar_rocm.problem import PR_GATE_PROPAGATORS, PR_MIXED
    p, u = W.shrink(*W.four_qubit_ensemble(n_t=60, n_det=2, n_amp=2))
    p.objective = PR_GATE_PROPAGATORS
    F0, G0 = O.evaluate(p, u)
    for kw, tol in (({}, TOL_G), ({"store_props": 0}, TOL_G), ({"precision": PR_MIXED}, 1e-6),
                    ({"precision": 2}, 1e-6)):
        F, G, _ = run(p, u, **kw)
        assert abs(F - F0) <= 1e-9 and np.abs(G - G0).max() <= tol * np.abs(G0).max(), kw


@pytest.mark.parametrize("d,grad_mode,precision", [(5, 0, 0), (20, 0, 0), (5, 1, 0), (12, 0, 1), (16, 0, 1), (12, 0, 2), (16, 0, 2)])
def test_modulus_objective(d, grad_mode, precision):
    # :modulus (Objectives.jl:222, 356-361): F = sum |z_s|, reference prefactor
    # dt Im(inner / |z|); exact mode the derivative of |z|
    from pulsar_rocm.problem import PR_GRAD_EXACT, PR_MIXED, PR_PAIRS_MODULUS
    rng = np.random.default_rng(40 + d)
    npairs, nt = 3, 30
    p = Problem(dim=d, n_ctrl=3, n_t=nt, H0=np.stack([W.random_hermitian(d, rng, 2.0) for _ in range(2)]),
                Hc=np.stack([W.random_hermitian(d, rng, 1.0) for _ in range(3)]), dt_slices=rng.uniform(0.005, 0.02, nt),
                objective=PR_PAIRS_MODULUS, n_members=2, ctrl_scale=np.array([[1.0, 0.9, 1.1], [1.1, 1.0, 0.95]]),
                psi_init=np.stack([W.random_state(d, rng) for _ in range(npairs)]),
                psi_targ=np.stack([W.random_state(d, rng) for _ in range(npairs)]))
    u = rng.uniform(-1, 1, (3, nt))
    print(f"case modulus d={d} grad_mode={grad_mode} precision={precision}", flush=True)
    F0, G0 = O.evaluate(p, u, grad_mode=grad_mode)
    F, G, _ = run(p, u, grad_mode=grad_mode, precision=precision)
    tf, tg = (TOL_F, TOL_G) if precision == 0 else (1e-9, 1e-6)
    assert abs(F - F0) <= tf, (F, F0)
    assert np.abs(G - G0).max() / np.abs(G0).max() <= tg
