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

# one device plan for the whole ensemble (members share the control operators
# up to per-channel scale factors, _scale_controls)
function _rocm_ensemble_plan(systems::Vector{<:AbstractQuantumSystem}, target::QuantumTarget,
                             controls::ControlSequence)
    n_c = size(controls.controls, 1)
    n_t = controls.n_steps
    ref = systems[1].H_controls
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
    return ROCmPlan([prob], roots)
end

# build_ensemble_from_systems → batched_eval  (src/Optimization/Ensemble/SystemBuilder.jl:49-87)
function _rocm_ensemble_objective(systems::Vector{<:AbstractQuantumSystem}, target::QuantumTarget,
                                  controls::ControlSequence)
    n_c = size(controls.controls, 1)
    n_t = controls.n_steps
    plan = _rocm_ensemble_plan(systems, target, controls)
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

# grape_tracking_kernel(waveform, ctrl)  (src/Application/MR/GRAPETracking.jl:145-157)
# gains `elseif ctrl.backend == :rocm` → `_rocm_grape_tracking_kernel(waveform, ctrl)`:
# one PR_TRACKING problem over the members [(H, p) for H in ctrl.drifts for p in
# ctrl.pwr_levels], initial states ctrl.rho_init and checkpoints ctrl.tracking; the
# penalties are applied afterwards, as the CPU path does (:260).
function _rocm_grape_tracking_kernel(waveform::Matrix{Float64}, ctrl)
    isempty(ctrl.tracking) && error("grape_tracking_kernel: ctrl.tracking is empty. " *
                                    "Add TrackingPoint entries or call grape_state_kernel instead.")
    n_ctrl, n_t = size(waveform)
    d = size(ctrl.drifts[1], 1)
    members = [(H, p) for H in ctrl.drifts for p in ctrl.pwr_levels]
    M = length(members)
    H0 = cat([ComplexF64.(H) for (H, _) in members]...; dims = 3)
    HcA = cat([ComplexF64.(o) for o in ctrl.operators]...; dims = 3)
    sc = Matrix{Float64}(undef, n_ctrl, M)                        # C layout [M][n_ctrl]
    for (m, (_, p)) in enumerate(members)
        sc[:, m] .= p
    end
    ψ0 = reduce(hcat, [ComplexF64.(vec(r)) for r in ctrl.rho_init])   # C layout [n_pairs][d]
    dts = Vector{Float64}(ctrl.pulse_dt)
    steps = Cint[tp.step for tp in ctrl.tracking]
    tw = Float64[tp.weight for tp in ctrl.tracking]
    tt = cat([ComplexF64.(tp.target_dm) for tp in ctrl.tracking]...; dims = 3)
    roots = Any[H0, HcA, sc, ψ0, dts, steps, tw, tt]
    prob = PRProblem(d, n_ctrl, n_t, 0.0, pointer(dts), PR_CLOSED, PR_TRACKING, M, _cptr(H0), 1,
                     _cptr(HcA), _cptr(sc), C_NULL, C_NULL, length(ctrl.rho_init), _cptr(ψ0), C_NULL,
                     length(steps), pointer(steps), pointer(tw), _cptr(tt))
    plan = ROCmPlan([prob], roots)
    F, G = rocm_eval(plan, [waveform])
    fidelity = _apply_penalties!(F[1], G[1], waveform, ctrl)
    return fidelity, G[1]
end

# compute_grape_gradient(sys::MASSpinSystem, ctrl, target; rotor_phase_0) and
# compute_grape_gradient_powder(sys, ctrl, target, orientations)
# (src/Physics/MRPhysics.jl:155-188) gain a `get_device() === :rocm` branch calling
# these. One problem per orientation: the drift is H_0 of the rotated system and the
# rotor modulation enters as four channels K_m^c = H_m + H_-m, K_m^s = i (H_m - H_-m)
# (m = 1, 2) with the fixed waveform rows cos(m w_r t_k), sin(m w_r t_k); the objective
# is the propagator helper's (:312-351), whose trace has no U_target^dag.
function _rocm_mas_problem(sys, ctrl, target, rotor_phase_0::Float64)
    n_c, n_t = size(ctrl.controls)
    d = sys.dim
    Hf = sys.H_fourier
    chans = [ComplexF64.(h) for h in sys.H_controls]
    for m in 1:2
        push!(chans, Hf[m] .+ Hf[-m])
        push!(chans, 1im .* (Hf[m] .- Hf[-m]))
    end
    t = ((1:n_t) .- 0.5) .* ctrl.dt .+ rotor_phase_0 / sys.omega_r
    θ = sys.omega_r .* t
    u = vcat(Matrix{Float64}(ctrl.controls), permutedims(hcat(cos.(θ), sin.(θ), cos.(2θ), sin.(2θ))))
    H0 = ComplexF64.(Hf[0])
    HcA = cat(chans...; dims = 3)
    roots = Any[H0, HcA, u]
    if target.type == "unitary"
        Ut = ComplexF64.(target.target_unitary)
        push!(roots, Ut)
        prob = PRProblem(d, length(chans), n_t, ctrl.dt, C_NULL, PR_CLOSED, PR_GATE_PROPAGATORS, 1, _cptr(H0), 0,
                         _cptr(HcA), C_NULL, C_NULL, _cptr(Ut), 0, C_NULL, C_NULL, 0, C_NULL, C_NULL, C_NULL)
    else
        ψt = ComplexF64.(target.target_state)
        ψi = target.initial_state === nothing ? ψt : ComplexF64.(target.initial_state)   # as given
        w1 = ones(1)
        push!(roots, ψi, ψt, w1)
        prob = PRProblem(d, length(chans), n_t, ctrl.dt, C_NULL, PR_CLOSED, PR_STATE_PROPAGATORS, 1, _cptr(H0), 0,
                         _cptr(HcA), C_NULL, _cptr(w1), C_NULL, 1, _cptr(ψi), _cptr(ψt), 0, C_NULL, C_NULL, C_NULL)
    end
    return prob, roots, u
end

function _rocm_compute_grape_gradient_powder(sys, ctrl, target, orientations)
    probs, roots, us = PRProblem[], Any[], Matrix{Float64}[]
    for (α, β, γ, _) in orientations
        p, r, u = _rocm_mas_problem(rotate_spin_system(sys, α, β, γ), ctrl, target, 0.0)
        push!(probs, p); push!(roots, r); push!(us, u)
    end
    plan = ROCmPlan(probs, roots)
    _, Gs = rocm_eval(plan, us)
    n_c = size(ctrl.controls, 1)
    G = zeros(Float64, size(ctrl.controls))
    for (i, ori) in enumerate(orientations)
        G .+= ori[4] .* Gs[i][1:n_c, :]
    end
    return G
end

function _rocm_compute_grape_gradient_mas(sys, ctrl, target; rotor_phase_0::Float64 = 0.0)
    p, r, u = _rocm_mas_problem(sys, ctrl, target, rotor_phase_0)
    plan = ROCmPlan([p], Any[r])
    _, Gs = rocm_eval(plan, [u])
    return Gs[1][1:size(ctrl.controls, 1), :]
end

# grape_optimize(system, target, controls_init; config)  (src/Optimization/GRAPE.jl:469-643)
# The whole gradient-ascent loop in the library (pr_grape_optimize): controls,
# gradient and best controls stay in HBM; the step adaptation, update and
# convergence rules are those of GRAPE.jl:261-402. Used when no callback or
# printing is requested (those need the Julia loop).
struct PRGrapeConfig
    max_iter::Cint
    convergence_tol::Cdouble
    step_size::Cdouble
    adapt_step_size::Cint
    min_step_size::Cdouble
    max_step_size::Cdouble
    gradient_norm_tol::Cdouble
end

mutable struct PRGrapeResult
    n_iterations::Cint
    converged::Cint
    reason::Cint
    n_fidelity_history::Cint
    n_fidelity_evaluations::Cint
    n_gradient_evaluations::Cint
    fidelity::Cdouble
    step_size_final::Cdouble
    PRGrapeResult() = new(0, 0, 0, 0, 0, 0, 0.0, 0.0)
end

function _rocm_grape_optimize(system::AbstractQuantumSystem, target::QuantumTarget,
                              controls_init::ControlSequence, config::GRAPEConfig)::OptimizationResult
    t0 = time()
    n_c, n_t, dt = system.n_controls, controls_init.n_timesteps, controls_init.dt
    fprob, froots = rocm_problem(system.H_drift, system.H_controls, n_t, dt, target; state_normalize = false)
    fplan = ROCmPlan([fprob], froots)
    gplan = nothing
    if target.type != "unitary"   # the gradient uses target_state as init (Gradients.jl:156-157)
        gprob, groots = rocm_problem(system.H_drift, system.H_controls, n_t, dt, target; fixed_point = true)
        gplan = ROCmPlan([gprob], groots)
    end
    cfg = PRGrapeConfig(config.max_iter, config.convergence_tol, config.step_size, config.adapt_step_size,
                        config.min_step_size, config.max_step_size, config.gradient_norm_tol)
    u0 = Matrix{Float64}(controls_init.controls)
    best = similar(u0)
    fh = zeros(config.max_iter + 2)
    gh = zeros(config.max_iter)
    res = PRGrapeResult()
    GC.@preserve u0 best fh gh begin
        _rocm_check(ccall((:pr_grape_optimize, _ROCM_LIB[]), Cint,
                          (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ref{PRGrapeConfig}, Ptr{Float64}, Ptr{Float64},
                           Ptr{Float64}, Ref{PRGrapeResult}),
                          fplan.handle, gplan === nothing ? C_NULL : gplan.handle, u0, cfg, best, fh, gh, res),
                    "pr_grape_optimize")
    end
    reasons = ("maximum iterations reached", "perfect fidelity", "gradient norm below tolerance",
               "fidelity change below tolerance")
    return OptimizationResult(best, res.fidelity, fh[1:res.n_fidelity_history], gh[1:res.n_iterations],
                              Int(res.n_iterations), res.converged != 0, reasons[res.reason + 1], time() - t0,
                              Int(res.n_fidelity_evaluations), Int(res.n_gradient_evaluations),
                              Dict{String, Any}("step_size_final" => res.step_size_final, "algorithm" => "GRAPE",
                                                "n_controls" => n_c, "n_timesteps" => n_t, "dt" => dt,
                                                "total_time" => dt * n_t))
end

# grape_optimize_ensemble(systems, target, controls_init; config, lbfgs_memory, amplitude_limit)
# (src/Optimization/GRAPE.jl:1012-1230). The L-BFGS loop in the library
# (pr_lbfgs_optimize): two-loop recursion and line search of GRAPE.jl:1062-1104
# with controls, gradient and curvature pairs resident on the device. Used when
# no callback or printing is requested (those need the Julia loop).
struct PRLbfgsConfig
    max_iter::Cint
    convergence_tol::Cdouble
    step_size::Cdouble
    gradient_norm_tol::Cdouble
    lbfgs_memory::Cint
    amplitude_limit::Cdouble
end

function _rocm_grape_optimize_ensemble(systems::Vector{<:AbstractQuantumSystem}, target::QuantumTarget,
                                       controls_init::ControlSequence, config::GRAPEConfig;
                                       lbfgs_memory::Int = 10, amplitude_limit::Real = Inf)::OptimizationResult
    t0 = time()
    plan = _rocm_ensemble_plan(systems, target, controls_init)
    cfg = PRLbfgsConfig(config.max_iter, config.convergence_tol, config.step_size, config.gradient_norm_tol,
                        lbfgs_memory, isfinite(amplitude_limit) ? amplitude_limit : 0.0)
    u0 = Matrix{Float64}(controls_init.controls)
    best = similar(u0)
    fh = zeros(config.max_iter + 1)
    gh = zeros(config.max_iter)
    res = PRGrapeResult()
    GC.@preserve u0 best fh gh begin
        _rocm_check(ccall((:pr_lbfgs_optimize, _ROCM_LIB[]), Cint,
                          (Ptr{Cvoid}, Ptr{Float64}, Ref{PRLbfgsConfig}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                           Ref{PRGrapeResult}), plan.handle, u0, cfg, best, fh, gh, res),
                    "pr_lbfgs_optimize")
    end
    reasons = ("maximum iterations reached", "near-perfect fidelity", "gradient norm below tolerance",
               "fidelity change below tolerance", "zero gradient direction")
    return OptimizationResult(best, res.fidelity, fh[1:res.n_fidelity_history], gh[1:res.n_iterations],
                              Int(res.n_fidelity_history), res.converged != 0, reasons[res.reason + 1], time() - t0,
                              Int(res.n_fidelity_evaluations), Int(res.n_gradient_evaluations),
                              Dict{String, Any}("algorithm" => "Ensemble L-BFGS-GRAPE",
                                                "n_systems" => length(systems), "lbfgs_memory" => lbfgs_memory))
end

