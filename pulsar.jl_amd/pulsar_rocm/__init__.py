"""pulsar_rocm — the GRAPE inner loop of Pulsar.jl on MI355X (gfx950).

The compute path is libpulsar_rocm.so (hand-written HIP, C-ABI in
include/pulsar_rocm.h), loaded through `native`. `api` and `mr` mirror the
reference's Julia interface for this path.
"""
from . import native, problem, qc, workloads  # noqa: F401
from .api import (BatchedEnsemble, ControlSequence, EnsembleObjective, GRAPEConfig,  # noqa: F401
                  OptimizationResult, QuantumSystem, QuantumTarget, adapt_grape_step_size,
                  available_devices, build_ensemble_from_systems, check_grape_convergence,
                  compute_fidelity, compute_grape_gradient, ensemble_value, ensemble_value_and_grad,
                  evaluate_fidelity, get_device, grape_gradient, grape_optimize, grape_optimize_ensemble,
                  grape_step, quantum_system, random_controls, set_device, state_target, unitary_target,
                  with_device, zero_controls)
from .ensemble_map import MemberObjectives, ensemble_fobj, ensemble_grad, ensemble_map  # noqa: F401
from .verify import (AutoDiffConfig, compute_gradient_autodiff, finite_diff_gradient_ad,  # noqa: F401
                     finite_difference_gradient, grape_gradient_check, select_autodiff_backend,
                     verify_gradient, verify_gradient_autodiff)
from .mr import (FidelityForward, LindbladMRControl, MRControl, TrackingPoint,  # noqa: F401
                 build_control_liouvillian, build_drift_liouvillian, fidelity_forward, grape_lindblad_kernel,
                 grape_state_kernel, grape_tracking_kernel, optimcon, pure_state_to_vec_rho, vec_rho)
from .problem import (PR_CLOSED, PR_FP64, PR_GATE, PR_GRAD_EXACT, PR_GRAD_REFERENCE,  # noqa: F401
                      PR_LINDBLAD, PR_MIXED, PR_PAIRS_REAL, PR_PAIRS_SQUARE, PR_STATE, PR_TRACKING, Problem)
