"""pr_eval_device_static (the entry point that allocates nothing) against
pr_eval_device on the same plans and inputs, through the C-ABI.

Both run the same kernels on the same inputs; only the scratch memory differs
(the library's static array instead of the plan's buffers), so F and G must
agree bit for bit. The calls come in a row: a 16-dimensional ensemble, then a
smaller two-problem plan (state pairs and Lindblad), then the first plan again,
so each call runs on what the previous one left in the static array.
"""
import numpy as np
import pytest

from pulsar_rocm import native, workloads as W
from pulsar_rocm.problem import PR_LINDBLAD, PR_PAIRS_SQUARE, Problem, controls_cm

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def device():
    native.init(0)


def _plans():
    p, u = W.four_qubit_ensemble(n_t=96)
    p, u = W.shrink(p, u, n_members=8)
    rng = np.random.default_rng(3)
    d = 6
    q = Problem(dim=d, n_ctrl=2, n_t=20, H0=np.stack([W.random_hermitian(d, rng, 2.0) for _ in range(2)]),
                Hc=np.stack([W.random_hermitian(d, rng, 1.0) for _ in range(2)]), dt=0.05,
                objective=PR_PAIRS_SQUARE, n_members=2, ctrl_scale=np.array([[1.0, 1.0], [1.1, 0.9]]),
                psi_init=np.stack([W.random_state(d, rng) for _ in range(2)]),
                psi_targ=np.stack([W.random_state(d, rng) for _ in range(2)]))
    uq = rng.uniform(-1, 1, (2, 20))
    lq, ul = W.three_qubit_lindblad(n_t=6, n_members=2, objective=PR_PAIRS_SQUARE)
    return [([p], [u]), ([q, lq], [uq, ul])]


def _eval(plan, us, static):
    """F and the gradients of every problem, evaluated on torch device buffers."""
    dc = torch.tensor(np.concatenate([controls_cm(u) for u in us]), dtype=torch.float64, device="cuda")
    dF = torch.full((len(us),), np.nan, dtype=torch.float64, device="cuda")
    dG = torch.full((plan.n_params,), np.nan, dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    torch.cuda.synchronize()
    f = plan.eval_device_static if static else plan.eval_device
    f(dc.data_ptr(), dF.data_ptr(), dG.data_ptr(), st)
    torch.cuda.synchronize()
    return dF.cpu().numpy(), dG.cpu().numpy()


def test_static_scratch_matches_plan_scratch_on_consecutive_calls():
    (pa, ua), (pb, ub) = _plans()
    plan_a, plan_b = native.Plan(pa), native.Plan(pb)
    try:
        for step, (plan, us) in enumerate(((plan_a, ua), (plan_b, ub), (plan_a, ua))):
            print(f"static call {step}: {len(us)} problem(s), {plan.n_params} controls", flush=True)
            F0, G0 = _eval(plan, us, static=False)
            F1, G1 = _eval(plan, us, static=True)
            assert np.all(np.isfinite(F1)) and np.all(np.isfinite(G1))
            assert np.array_equal(F0, F1), (F0, F1)
            assert np.array_equal(G0, G1)
        # the plan-owned path still gives the same answer after static calls
        F2, G2 = _eval(plan_b, ub, static=False)
        assert np.array_equal(F2, _eval(plan_b, ub, static=True)[0])
    finally:
        plan_a.close()
        plan_b.close()


def test_static_scratch_rejects_plans_beyond_its_size():
    # 64 members x 2000 slices keep 0.5 GB of U_k images: beyond the 64 MiB array
    p, u = W.four_qubit_ensemble()
    p, u = W.shrink(p, u, n_members=64)
    plan = native.Plan([p])
    try:
        with pytest.raises(native.RocmError) as e:
            _eval(plan, [u], static=True)
        assert e.value.code == -5 and "64" in str(e.value)
    finally:
        plan.close()
