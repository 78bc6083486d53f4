# rocm_hooks.jl — :rocm dispatch for the GRAPE path, included after
# src/Optimization (it reaches Physics/Optimization/Application types), e.g.
# from src/Pulsar.jl right after the Application layer.
#
# Each hook keeps the reference signature and falls back to the CPU code on a
# library error, as CUDABackend.jl:144-147 does.

# compute_fidelity(system, controls, target)  (src/Physics/Objectives.jl:632-639)
function _rocm_compute_fidelity(system::AbstractQuantumSystem, controls::ControlSequence,
                                target::QuantumTarget)::Float64
    prob, roots = rocm_problem(system.H_drift, system.H_controls, controls.n_steps, controls.dt, target;
                               state_normalize = false)
    plan = ROCmPlan([prob], roots)
    F, _ = rocm_eval(plan, [Matrix(controls.controls)]; grad = false)
    return F[1]
end

# compute_grape_gradient(system, controls, target)  (src/Physics/Gradients.jl:80-127)
function _rocm_compute_grape_gradient(system::AbstractQuantumSystem, controls::ControlSequence,
                                      target::QuantumTarget)::Matrix{Float64}
    prob, roots = rocm_problem(system.H_drift, system.H_controls, controls.n_steps, controls.dt, target)
    plan = ROCmPlan([prob], roots)
    _, G = rocm_eval(plan, [Matrix(controls.controls)])
    return G[1]
end

# The reference's entry points gain one branch each, e.g. in Objectives.jl:
#
#   function compute_fidelity(system::AbstractQuantumSystem, controls::ControlSequence,
#                             target::QuantumTarget)::Float64
#       if get_device() === :rocm && is_rocm_available()
#           try
#               return _rocm_compute_fidelity(system, controls, target)
#           catch e
#               @warn "Pulsar: ROCm compute_fidelity failed — falling back to CPU" exception=e maxlog=1
#           end
#       end
#       ... existing CPU body ...
#   end

# build_ensemble_from_systems → batched_eval  (src/Optimization/Ensemble/SystemBuilder.jl:49-87)
function _rocm_ensemble_objective(systems::Vector{<:AbstractQuantumSystem}, target::QuantumTarget,
                                  controls::ControlSequence)
    n_c = size(controls.controls, 1)
    n_t = controls.n_steps
    ref = systems[1].H_controls
    # members share the control operators up to per-channel scale factors (_scale_controls)
    scales = zeros(length(systems), n_c)
    for (i, s) in enumerate(systems), j in 1:n_c
        a = ref[j]; b = s.H_controls[j]
        c = real(dot(a, b) / dot(a, a))
        norm(b - c * a) <= 1e-13 * max(norm(b), eps()) ||
            error("ensemble members differ in control operators beyond a scale factor")
        scales[i, j] = c
    end
    prob, roots = rocm_problem([s.H_drift for s in systems], ref, n_t, controls.dt, target;
                               scales = scales, fixed_point = true)
    plan = ROCmPlan([prob], roots)
    batched = θ -> begin
        F, G = rocm_eval(plan, [reshape(Vector{Float64}(θ), n_c, n_t)])
        (F[1], vec(G[1]))
    end
    return EnsembleObjective(; batched_eval = batched, n_samples = length(systems))
end

# grape_state_kernel / grape_lindblad_kernel  (src/Application/MR/GRAPEState.jl:73-85,
# GRAPELindblad.jl:83-92) gain `elseif ctrl.backend == :rocm` branches that build
# one pr_problem per call: members = [(H, p) for H in ctrl.drifts for p in ctrl.pwr_levels]
# (H0_per_member = 1, ctrl_scale[m, :] .= p), pairs = ctrl.rho_init / ctrl.rho_targ,
# dt_slices = ctrl.pulse_dt, objective = :square → PR_PAIRS_SQUARE, :real → PR_PAIRS_REAL;
# the Lindblad branch passes ctrl._L_drifts / ctrl._L_controls with dynamics = PR_LINDBLAD.
