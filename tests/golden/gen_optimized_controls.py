"""Writes tests/golden/c2_optimized.npz: a GRAPE-optimized control point of
the four-qubit ensemble (d=16, 256 members, 8 controls, 2000 slices), where
the mixed- and split-precision forms are checked and timed near F = 1.

The ensemble's Haar-random target cannot be reached (the Ising couplings give
entangling phases of a few tenths of a radian over 2 ms), so this point uses a
reachable target: the nominal system's propagator (no detuning, amplitude
scale 1) at the ensemble's reference controls u_ref. From u_ref the device's
ensemble L-BFGS (pr_lbfgs_optimize, FP64) maximises the ensemble mean of
|Tr(U_t^dag U_m)/16|^2 over the 16 x 16 detuning x amplitude grid. The file
holds the controls, the target, the fidelity history and the settings.

    python tests/golden/gen_optimized_controls.py [--iters 400]     (needs the GPU)
"""
import argparse
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "pulsar.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
from pulsar_rocm import native, workloads as W  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=400)
    ap.add_argument("--out", default=os.path.join(HERE, "c2_optimized.npz"))
    args = ap.parse_args()
    p, u_ref = W.four_qubit_ensemble()
    nom, _ = W.four_qubit_nominal()
    Ut = O.total_propagator(nom, u_ref)  # FP64 oracle: the target is reachable by construction
    p.U_target = Ut
    native.init(0)
    plan = native.Plan([p])
    try:
        F0, _ = plan.eval([u_ref], grad=False)
        cfg = native.LbfgsConfig(max_iter=args.iters, convergence_tol=1e-12, step_size=0.05 * np.abs(u_ref).max(),
                                 gradient_norm_tol=1e-12, lbfgs_memory=10, amplitude_limit=0.0)
        best, fh, gh, res = native.lbfgs_optimize_device(plan, u_ref, cfg)
        F1, _ = plan.eval([best], grad=False)
    finally:
        plan.close()
    print(f"ensemble F at u_ref {F0[0]:.6f} -> {F1[0]:.6f} after {res.n_iterations} iterations "
          f"({native.PR_STOP.get(res.reason)})", flush=True)
    np.savez_compressed(args.out, controls=best, target=Ut, fidelity_history=np.array(fh),
                        iterations=res.n_iterations, F=F1[0], F_ref=F0[0])
    print("wrote", args.out)


if __name__ == "__main__":
    main()
