# rocm_hooks.jl — :rocm dispatch for the GRAPE path, included after
# src/Optimization (it reaches Physics/Optimization/Application types), e.g.
# from src/Pulsar.jl right after the Application layer.
#
# Each hook keeps the reference signature and falls back to the CPU code on a
# library error, as CUDABackend.jl:144-147 does.

# Plans are cached per (system, target, shape, objective form): a GRAPE loop that
# calls compute_fidelity / compute_grape_gradient on the same system and target
# every iteration validates and uploads the operators once, and at C2 sizes
# allocates the U_k images once. Keys are content hashes; a hit is confirmed by
# comparing the stored inputs, so a hash collision cannot return a wrong plan.
const _ROCM_PLAN_CACHE = Vector{Tuple{UInt, Any, ROCmPlan}}()
const _ROCM_PLAN_CACHE_MAX = 8

function _rocm_cached_plan(build::Function, key_inputs::Tuple)
    h = hash(key_inputs)
    for (kh, kin, plan) in _ROCM_PLAN_CACHE
        kh == h && isequal(kin, key_inputs) && plan.handle != C_NULL && return plan
    end
    plan = build()
    length(_ROCM_PLAN_CACHE) >= _ROCM_PLAN_CACHE_MAX && popfirst!(_ROCM_PLAN_CACHE)
    push!(_ROCM_PLAN_CACHE, (h, deepcopy(key_inputs), plan))
    return plan
end

_target_key(t::QuantumTarget) = (t.type, t.target_unitary, t.target_state, t.initial_state)

# compute_fidelity(system, controls, target)  (src/Physics/Objectives.jl:632-639)
function _rocm_compute_fidelity(system::AbstractQuantumSystem, controls::ControlSequence,
                                target::QuantumTarget)::Float64
    key = (:fidelity, system.H_drift, system.H_controls, controls.n_steps, controls.dt, _target_key(target))
    plan = _rocm_cached_plan(key) do
        prob, roots = rocm_problem(system.H_drift, system.H_controls, controls.n_steps, controls.dt, target;
                                   state_normalize = false)
        ROCmPlan([prob], roots)
    end
    F, _ = rocm_eval(plan, [Matrix(controls.controls)]; grad = false)
    return F[1]
end

# compute_grape_gradient(system, controls, target)  (src/Physics/Gradients.jl:80-127)
function _rocm_compute_grape_gradient(system::AbstractQuantumSystem, controls::ControlSequence,
                                      target::QuantumTarget)::Matrix{Float64}
    key = (:gradient, system.H_drift, system.H_controls, controls.n_steps, controls.dt, _target_key(target))
    plan = _rocm_cached_plan(key) do
        prob, roots = rocm_problem(system.H_drift, system.H_controls, controls.n_steps, controls.dt, target)
        ROCmPlan([prob], roots)
    end
    _, G = rocm_eval(plan, [Matrix(controls.controls)])
    return G[1]
end

# _fidelity_flat(u_flat, system, target, n_controls, n_timesteps, dt)
# (src/Physics/AutoDiff.jl:254-280) with its exact gradient: the primal and
# pullback of the ChainRulesCore rrule (rocm_zygote_rrule.jl). The objective is
# compute_fidelity(U_total, target) (Objectives.jl:609-631): the gate fidelity,
# or |<psi_t|U psi_i>|^2 with the states as given (psi_i = target_state when no
# initial state is set), hence PR_GATE or PR_PAIRS_SQUARE with weight 1.
function _rocm_fidelity_flat_and_gradient(u_flat::AbstractVector, system::AbstractQuantumSystem,
                                          target::QuantumTarget, n_controls::Int, n_timesteps::Int,
                                          dt::Float64; grad::Bool = true)
    key = (:flat, system.H_drift, system.H_controls, n_timesteps, dt, _target_key(target))
    plan = _rocm_cached_plan(key) do
        H0 = ComplexF64.(system.H_drift)
        HcA = cat([ComplexF64.(h) for h in system.H_controls]...; dims = 3)
        d = size(H0, 1)
        if target.type == "unitary"
            Ut = ComplexF64.(target.target_unitary)
            roots = Any[H0, HcA, Ut]
            prob = PRProblem(d, n_controls, n_timesteps, dt, C_NULL, PR_CLOSED, PR_GATE, 1, _cptr(H0), 0,
                             _cptr(HcA), C_NULL, C_NULL, _cptr(Ut), 0, C_NULL, C_NULL, 0, C_NULL, C_NULL, C_NULL)
        else
            ψt = ComplexF64.(target.target_state)
            ψi = target.initial_state === nothing ? ψt : ComplexF64.(target.initial_state)
            w1 = ones(1)
            roots = Any[H0, HcA, ψi, ψt, w1]
            prob = PRProblem(d, n_controls, n_timesteps, dt, C_NULL, PR_CLOSED, PR_PAIRS_SQUARE, 1, _cptr(H0), 0,
                             _cptr(HcA), C_NULL, _cptr(w1), C_NULL, 1, _cptr(ψi), _cptr(ψt), 0, C_NULL, C_NULL,
                             C_NULL)
        end
        ROCmPlan([prob], roots; grad_mode = PR_GRAD_EXACT)
    end
    F, G = rocm_eval(plan, [reshape(real.(Vector(u_flat)), n_controls, n_timesteps)]; grad = grad)
    return F[1], vec(G[1])
end

# The device form of _fidelity_flat (same arguments, same value). The Zygote
# extension differentiates this function instead of _fidelity_flat when
# get_device() === :rocm; its rrule (rocm_zygote_rrule.jl) takes F and dF/du
# from one device evaluation.
function _rocm_fidelity_flat(u_flat::AbstractVector, system::AbstractQuantumSystem, target::QuantumTarget,
                             n_controls::Int, n_timesteps::Int, dt::Float64)::Float64
    F, _ = _rocm_fidelity_flat_and_gradient(u_flat, system, target, n_controls, n_timesteps, dt; grad = false)
    return F
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
# GRAPELindblad.jl:83-92) gain `elseif ctrl.backend == :rocm` branches calling these:
# one problem whose members are [(H, p) for H in ctrl.drifts for p in ctrl.pwr_levels]
# (per-member drifts, ctrl_scale[m, :] .= p), pairs ctrl.rho_init / ctrl.rho_targ,
# dt_slices ctrl.pulse_dt; the penalties are applied afterwards as the CPU paths do.
function _rocm_mr_objective(fidelity::Symbol, lindblad::Bool)
    fidelity === :square && return PR_PAIRS_SQUARE
    fidelity === :real && return PR_PAIRS_REAL
    !lindblad && fidelity === :modulus && return PR_PAIRS_MODULUS
    lindblad && throw(ArgumentError("lindblad_grad_prefactor: ':$fidelity' not supported. Use :square or :real."))
    throw(ArgumentError("ctrl.fidelity ':$fidelity' is not GRAPE-compatible. Use one of: real, square, modulus"))
end

function _rocm_mr_plan(ctrl, drifts, operators, rho_init, rho_targ, lindblad::Bool)
    key = (:mr, lindblad, drifts, operators, rho_init, rho_targ, ctrl.pwr_levels, ctrl.pulse_dt, ctrl.fidelity)
    return _rocm_cached_plan(key) do
        members = [(H, p) for H in drifts for p in ctrl.pwr_levels]
        M = length(members)
        n_ctrl = length(operators)
        d = size(operators[1], 1)
        H0 = cat([ComplexF64.(H) for (H, _) in members]...; dims = 3)
        HcA = cat([ComplexF64.(o) for o in operators]...; dims = 3)
        sc = Matrix{Float64}(undef, n_ctrl, M)                        # C layout [M][n_ctrl]
        for (m, (_, p)) in enumerate(members)
            sc[:, m] .= p
        end
        ψ0 = reduce(hcat, [ComplexF64.(vec(r)) for r in rho_init])   # C layout [n_pairs][d]
        ψt = reduce(hcat, [ComplexF64.(vec(r)) for r in rho_targ])
        dts = Vector{Float64}(ctrl.pulse_dt)
        roots = Any[H0, HcA, sc, ψ0, ψt, dts]
        prob = PRProblem(d, n_ctrl, length(dts), 0.0, pointer(dts), lindblad ? PR_LINDBLAD : PR_CLOSED,
                         _rocm_mr_objective(ctrl.fidelity, lindblad), M, _cptr(H0), 1, _cptr(HcA), _cptr(sc),
                         C_NULL, C_NULL, length(rho_init), _cptr(ψ0), _cptr(ψt), 0, C_NULL, C_NULL, C_NULL)
        ROCmPlan([prob], roots)
    end
end

function _rocm_grape_state_kernel(waveform::Matrix{Float64}, ctrl)
    plan = _rocm_mr_plan(ctrl, ctrl.drifts, ctrl.operators, ctrl.rho_init, ctrl.rho_targ, false)
    F, G = rocm_eval(plan, [waveform])
    fidelity = _apply_penalties!(F[1], G[1], waveform, ctrl)
    return fidelity, G[1]
end

# the Liouvillians and vectorised states the reference precomputes in the
# LindbladMRControl constructor (OptControl.jl:855-870)
function _rocm_grape_lindblad_kernel(waveform::Matrix{Float64}, ctrl)
    plan = _rocm_mr_plan(ctrl, ctrl._L_drifts, ctrl._L_controls, ctrl._sigma_init, ctrl._sigma_targ, true)
    F, G = rocm_eval(plan, [waveform])
    fidelity = _apply_penalties!(F[1], G[1], waveform, ctrl)
    return fidelity, G[1]
end

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


# optimcon(ctrl::MRControl / LindbladMRControl, guess) (src/Application/MR/OptControl.jl:310-367)
# gains, before its dispatch on ctrl.method,
#   ctrl.backend == :rocm && _rocm_optimcon_applies(ctrl) && return _rocm_optimcon(ctrl, guess)
# The :lbfgs and :grape loops (_optimcon_lbfgs :380-484, _optimcon_grape :488-561) then
# run in the library (pr_mr_optimcon): waveform, gradient and curvature pairs stay in
# device memory and the line search evaluates F only. Penalties, non-default
# parameterizations, tracking and verbose printing keep the Julia loops (over the
# :rocm kernels above).
struct PRMrConfig
    method::Cint
    max_iter::Cint
    grad_tol::Cdouble
    lbfgs_memory::Cint
    l_bound::Cdouble
    u_bound::Cdouble
    tile::Cint
end
const PR_MR_LBFGS = Cint(0)
const PR_MR_GRAPE = Cint(1)

_rocm_optimcon_applies(ctrl) =
    ctrl.method in (:lbfgs, :grape) && all(==(:none), ctrl.penalties) && !ctrl.verbose &&
    ctrl.parameterization isa PiecewiseConstant && (!hasproperty(ctrl, :tracking) || isempty(ctrl.tracking))

function _rocm_optimcon(ctrl, guess::Matrix{Float64})::OptimizationResult
    t0 = time()
    lindblad = hasproperty(ctrl, :_L_drifts)
    plan = lindblad ?
        _rocm_mr_plan(ctrl, ctrl._L_drifts, ctrl._L_controls, ctrl._sigma_init, ctrl._sigma_targ, true) :
        _rocm_mr_plan(ctrl, ctrl.drifts, ctrl.operators, ctrl.rho_init, ctrl.rho_targ, false)
    cfg = PRMrConfig(ctrl.method === :lbfgs ? PR_MR_LBFGS : PR_MR_GRAPE, ctrl.max_iter, ctrl.grad_tol,
                     ctrl.lbfgs_memory, ctrl.l_bound, ctrl.u_bound, 1)
    best = similar(guess)
    fh = zeros(ctrl.max_iter + 1)
    gh = zeros(ctrl.max_iter + 1)
    res = PRGrapeResult()
    GC.@preserve guess best fh gh begin
        _rocm_check(ccall((:pr_mr_optimcon, _ROCM_LIB[]), Cint,
                          (Ptr{Cvoid}, Ptr{Float64}, Ref{PRMrConfig}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                           Ref{PRGrapeResult}), plan.handle, guess, cfg, best, fh, gh, res),
                    "pr_mr_optimcon")
    end
    n = Int(res.n_fidelity_history)
    reason = res.reason == 2 ? "gradient norm < $(ctrl.grad_tol)" :
             ("max iterations reached", "near-perfect fidelity", "", "fidelity stalled",
              "zero gradient direction")[res.reason + 1]
    algo = (lindblad ? "Lindblad " : "MR ") * (ctrl.method === :lbfgs ? "L-BFGS-GRAPE" : "GRAPE")
    meta = Dict{String, Any}("algorithm" => algo, "on_device" => true)
    ctrl.method === :lbfgs && (meta["lbfgs_memory"] = ctrl.lbfgs_memory)
    return OptimizationResult(best, res.fidelity, fh[1:n], gh[1:n], Int(res.n_iterations), res.converged != 0,
                              reason, time() - t0, Int(res.n_fidelity_evaluations),
                              Int(res.n_gradient_evaluations), meta)
end

# robust_optimcon_qs (src/Application/QuantumComputing/NoiseModels/QuasiStatic.jl:196-216)
# builds its quasi-static drift ensemble and calls grape_optimize_ensemble, whose
# :rocm branch (_rocm_grape_optimize_ensemble above) makes the whole ensemble one
# plan with the L-BFGS loop in the library: no edit of QuasiStatic.jl is needed.
#
# LindbladQCControl (src/Application/QuantumComputing/LindbladQCControl.jl:237-284)
# reaches grape_lindblad_kernel through _LindbladQCAdapter, whose getproperty answers
# :backend with :cpu. The one-line edit
#     name === :backend && return get_device() === :rocm ? :rocm : :cpu
# sends _qc_kernel to _rocm_grape_lindblad_kernel: the adapter's _L_drifts,
# _L_controls, _sigma_init, _sigma_targ, pwr_levels = [1.0], pulse_dt and :square
# fidelity build one PR_LINDBLAD problem (_rocm_mr_plan), and the adapter's own
# _apply_penalties! subtracts the QC penalties afterwards.

# ensemble_map / ensemble_grad! / ensemble_fobj  (src/Computation/EnsembleMap.jl:365-607)
# gain a `device === :rocm` branch. Arbitrary closures keep the CPU-thread map (as
# under :cuda/:metal, :375-415); a member function with a device description is one
# device evaluation for all members:
#
#   elseif device === :rocm
#       f isa RocmMemberFidelity ? (out .= _rocm_member_values(f.objs, f.θ, indices)) :
#                                  (@threadsif true for k in 1:n; out[k] = Float64(f(indices[k])); end)
#
# and in ensemble_grad!, before the threaded loop:
#
#   if device === :rocm && grad_member! isa RocmMemberGradient
#       ws = w === nothing ? fill(1.0 / n, n) : w ./ sum(w)
#       g_out .= _rocm_member_grad(grad_member!.objs, θ, indices, ws); return g_out
#   end
#
# Member objectives are the sample closures of build_ensemble_from_systems
# (SystemBuilder.jl:61-77): F_i = compute_fidelity(systems[i], ctrl(θ), target),
# g_i = vec(compute_grape_gradient(systems[i], ctrl(θ), target)).
struct RocmMemberObjectives
    systems::Vector{AbstractQuantumSystem}
    target::QuantumTarget
    n_t::Int
    dt::Float64
end
struct RocmMemberFidelity      # i -> F_i(θ), the closure of the ensemble_map example
    objs::RocmMemberObjectives
    θ::Vector{Float64}
end
struct RocmMemberF             # (θ, i) -> F_i(θ), ensemble_fobj's f_member
    objs::RocmMemberObjectives
end
struct RocmMemberGradient      # (g, θ, i) -> g_i(θ), grad_member!
    objs::RocmMemberObjectives
end
(f::RocmMemberFidelity)(i) = _rocm_member_values(f.objs, f.θ, [i])[1]
(f::RocmMemberF)(θ, i) = _rocm_member_values(f.objs, θ, [i])[1]
(g::RocmMemberGradient)(gv, θ, i) = (gv .= _rocm_member_grad(g.objs, θ, [i], [1.0]); gv)

function _rocm_member_values(o::RocmMemberObjectives, θ::AbstractVector, indices::Vector)
    n_c = length(o.systems[1].H_controls)
    key = (:member_F, [(o.systems[i].H_drift, o.systems[i].H_controls) for i in indices], o.n_t, o.dt,
           _target_key(o.target))
    plan = _rocm_cached_plan(key) do
        probs = PRProblem[]; roots = Any[]
        for i in indices
            p, r = rocm_problem(o.systems[i].H_drift, o.systems[i].H_controls, o.n_t, o.dt, o.target;
                                state_normalize = false)
            push!(probs, p); append!(roots, r)
        end
        ROCmPlan(probs, roots)
    end
    u = reshape(Vector{Float64}(θ), n_c, o.n_t)
    F, _ = rocm_eval(plan, fill(u, length(indices)); grad = false)
    return F
end

# the weighted mean gradient, reduced on the device: one problem per group of
# members that share the control operators up to per-channel scales, member
# weights ws (summing to 1)
function _rocm_member_grad(o::RocmMemberObjectives, θ::AbstractVector, indices::Vector, ws::Vector{Float64})
    n_c = length(o.systems[1].H_controls)
    key = (:member_G, [(o.systems[i].H_drift, o.systems[i].H_controls) for i in indices], ws, o.n_t, o.dt,
           _target_key(o.target))
    plan = _rocm_cached_plan(key) do
        groups = Vector{Tuple{Vector, Vector{Int}, Vector{Vector{Float64}}}}()
        for (k, i) in enumerate(indices)
            hc = o.systems[i].H_controls
            placed = false
            for (ref, ks, scs) in groups
                sc = [real(dot(a, b) / dot(a, a)) for (a, b) in zip(ref, hc)]
                if all(norm(b - c * a) <= 1e-13 * max(norm(b), eps()) for (a, b, c) in zip(ref, hc, sc))
                    push!(ks, k); push!(scs, sc); placed = true; break
                end
            end
            placed || push!(groups, (hc, [k], [ones(n_c)]))
        end
        probs = PRProblem[]; roots = Any[]
        for (ref, ks, scs) in groups
            p, r = rocm_problem([o.systems[indices[k]].H_drift for k in ks], ref, o.n_t, o.dt, o.target;
                                scales = permutedims(hcat(scs...)), weights = ws[ks])
            push!(probs, p); append!(roots, r)
        end
        ROCmPlan(probs, roots)
    end
    u = reshape(Vector{Float64}(θ), n_c, o.n_t)
    _, G = rocm_eval(plan, fill(u, length(plan.n_ctrl)))
    return vec(sum(G))
end
