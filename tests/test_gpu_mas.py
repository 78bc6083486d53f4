"""MAS gate gradients on the device (PR_GATE_PROPAGATORS with rotor channels)
against the MAS oracle (MRPhysics.jl:150-188, 312-336 restated).
FP64 on both sides: max|dG| / max|G| <= 1e-10."""
import numpy as np
import pytest

import oracle as O
from pulsar_rocm import mas as MAS, native
from pulsar_rocm.api import ControlSequence, unitary_target

pytestmark = pytest.mark.gpu
TWO_PI = 2 * np.pi


@pytest.fixture(scope="module", autouse=True)
def device():
    native.init(0)


def orientations(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.uniform(0, 2 * np.pi, n)
    b = np.arccos(rng.uniform(-1, 1, n))
    g = rng.uniform(0, 2 * np.pi, n)
    w = rng.uniform(0.5, 1.5, n)
    return [(float(a[i]), float(b[i]), float(g[i]), float(w[i] / w.sum())) for i in range(n)]


@pytest.mark.parametrize("isotopes,csa,dip,nt,n_ori", [
    (["13C"], [(TWO_PI * 2e3, TWO_PI * 9e3, 0.0, (0.2, 0.7, 0.0))], [], 60, 8),
    (["1H", "13C"], [(TWO_PI * 1e3, TWO_PI * 8e3, 0.0, (0.3, 0.9, 0.1)),
                     (-TWO_PI * 2e3, TWO_PI * 5e3, 0.0, (1.1, 0.4, 0.2))], [(1, 2, TWO_PI * 15e3, (0.2, 1.2, 0.5))], 40, 12),
    (["14N", "1H"], [(TWO_PI * 3e3, TWO_PI * 6e3, 0.0, (0.5, 1.0, 0.0))], [(1, 2, TWO_PI * 8e3, (0.0, 0.6, 0.9))], 30, 5),
])
def test_mas_powder_gradient_matches_oracle(isotopes, csa, dip, nt, n_ori):
    print(f"mas case {isotopes} n_t={nt} orientations={n_ori}", flush=True)
    rng = np.random.default_rng(nt)
    base = MAS.mr_system(isotopes)
    s = MAS.mas_spin_system(base, 12.5e3, [MAS.CSATensor(*c) for c in csa], [MAS.DipolarCoupling(*d) for d in dip])
    dt = 2e-6
    u = rng.uniform(-TWO_PI * 3e4, TWO_PI * 3e4, (s.n_controls, nt))
    ctrl = ControlSequence.from_controls(u, dt, nt * dt, nt)
    from pulsar_rocm.workloads import haar_unitary
    Ut = haar_unitary(s.dim, np.random.default_rng(3))
    ori = orientations(n_ori, n_ori)
    G = MAS.compute_grape_gradient_powder(s, ctrl, unitary_target(Ut), ori)
    G0 = O.mas_powder_gradient(isotopes, csa, dip, 12.5e3, u, dt, Ut, ori)
    assert G.shape == G0.shape
    assert np.abs(G - G0).max() <= 1e-10 * np.abs(G0).max(), np.abs(G - G0).max() / np.abs(G0).max()


def test_mas_single_orientation_with_rotor_phase():
    csa = [(TWO_PI * 1e3, TWO_PI * 8e3, 0.0, (0.3, 0.9, 0.1))]
    s = MAS.mas_spin_system(MAS.mr_system(["1H"]), 10e3, [MAS.CSATensor(*csa[0])])
    rng = np.random.default_rng(5)
    u = rng.uniform(-TWO_PI * 3e4, TWO_PI * 3e4, (2, 50))
    ctrl = ControlSequence.from_controls(u, 1e-6, 50e-6, 50)
    from pulsar_rocm.workloads import haar_unitary
    Ut = haar_unitary(2, np.random.default_rng(11))
    G = MAS.compute_grape_gradient_mas(s, ctrl, unitary_target(Ut), rotor_phase_0=0.9)
    Hf = O.mas_fourier(["1H"], csa, [])
    Ix, Iy, _ = O.spin_operators(["1H"])
    Us = O.mas_propagators(Hf, [Ix[0], Iy[0]], 10e3, u, 1e-6, rotor_phase_0=0.9)
    G0 = O.propagator_gradient(Us, [Ix[0], Iy[0]], Ut, 1e-6)
    assert np.abs(G - G0).max() <= 1e-10 * np.abs(G0).max()


def test_mas_state_target_matches_oracle():
    from pulsar_rocm.api import state_target
    csa = [(TWO_PI * 1e3, TWO_PI * 8e3, 0.0, (0.3, 0.9, 0.1)), (-TWO_PI * 2e3, TWO_PI * 5e3, 0.0, (1.1, 0.4, 0.2))]
    dip = [(1, 2, TWO_PI * 15e3, (0.2, 1.2, 0.5))]
    s = MAS.mas_spin_system(MAS.mr_system(["1H", "13C"]), 12.5e3, [MAS.CSATensor(*c) for c in csa],
                            [MAS.DipolarCoupling(*d) for d in dip])
    rng = np.random.default_rng(21)
    u = rng.uniform(-TWO_PI * 3e4, TWO_PI * 3e4, (4, 40))
    ctrl = ControlSequence.from_controls(u, 2e-6, 80e-6, 40)
    pt = np.array([0, 1, 1j, 0], complex)  # unnormalised: the helper uses states as given
    pi = np.array([1, 0, 0, 1], complex)
    ori = orientations(6, 6)
    G = MAS.compute_grape_gradient_powder(s, ctrl, state_target(pt, psi_init=pi), ori)
    G0 = O.mas_powder_gradient(["1H", "13C"], csa, dip, 12.5e3, u, 2e-6, None, ori, psi_init=pi, psi_targ=pt)
    assert np.abs(G - G0).max() <= 1e-10 * np.abs(G0).max()


def test_per_orientation_problems_in_one_plan():
    # the Julia glue's layout: one PR_GATE_PROPAGATORS problem per orientation
    from pulsar_rocm.problem import PR_GATE_PROPAGATORS, Problem
    from pulsar_rocm.workloads import haar_unitary
    csa = [(TWO_PI * 1e3, TWO_PI * 8e3, 0.0, (0.3, 0.9, 0.1)), (-TWO_PI * 2e3, TWO_PI * 5e3, 0.0, (1.1, 0.4, 0.2))]
    dip = [(1, 2, TWO_PI * 15e3, (0.2, 1.2, 0.5))]
    rng = np.random.default_rng(31)
    u = rng.uniform(-TWO_PI * 3e4, TWO_PI * 3e4, (4, 30))
    dt, w_r = 2e-6, TWO_PI * 12.5e3
    Ut = haar_unitary(4, np.random.default_rng(2))
    Ix, Iy, Iz = O.spin_operators(["1H", "13C"])
    Hc = [Ix[0], Iy[0], Ix[1], Iy[1]]
    ori = orientations(5, 5)
    probs, us = [], []
    for a, b, g, _ in ori:
        Hf = O.mas_fourier(["1H", "13C"], csa, dip, (a, b, g))
        chans = Hc + [Hf[1] + Hf[-1], 1j * (Hf[1] - Hf[-1]), Hf[2] + Hf[-2], 1j * (Hf[2] - Hf[-2])]
        th = w_r * (np.arange(30) + 0.5) * dt
        us.append(np.vstack([u, np.cos(th), np.sin(th), np.cos(2 * th), np.sin(2 * th)]))
        probs.append(Problem(dim=4, n_ctrl=8, n_t=30, H0=Hf[0], Hc=np.stack(chans), dt=dt,
                             objective=PR_GATE_PROPAGATORS, U_target=Ut))
    plan = native.Plan(probs)
    try:
        _, Gs = plan.eval(us)
    finally:
        plan.close()
    G = sum(o[3] * g[:4] for o, g in zip(ori, Gs))
    G0 = O.mas_powder_gradient(["1H", "13C"], csa, dip, 12.5e3, u, dt, Ut, ori)
    assert np.abs(G - G0).max() <= 1e-10 * np.abs(G0).max()
