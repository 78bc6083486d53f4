# ROCmBackend.jl — the :rocm device for Pulsar.jl (AMD MI355X, gfx950).
#
# Drop-in beside src/Backend/Hardware/CUDABackend.jl and MetalBackend.jl.
# Binds libpulsar_rocm.so (include/pulsar_rocm.h) with ccall; no Julia GPU
# package is needed. Layering: this file uses only Types (ControlSequence,
# QuantumTarget, AbstractQuantumSystem) and defines the stubs that the Physics,
# Optimization and Application layers dispatch to when get_device() == :rocm
# (see rocm_hooks.jl), so Backend never includes a higher layer
# (test/Architecture/LayerDependencyTests.jl:20-26).
#
# Error behaviour mirrors CUDABackend.jl:144-147: a non-zero status from the
# library is reported once with @warn and the caller reruns the CPU path.

const _ROCM_LOADED = Ref{Bool}(false)
const _ROCM_LIB = Ref{String}("")

const PR_CLOSED = Cint(0)
const PR_LINDBLAD = Cint(1)
const PR_GATE = Cint(0)
const PR_PAIRS_SQUARE = Cint(1)
const PR_PAIRS_REAL = Cint(2)
const PR_TRACKING = Cint(3)
const PR_GATE_PROPAGATORS = Cint(4)
const PR_STATE = Cint(5)   # compute_grape_gradient with a state target (Gradients.jl:357-399)
const PR_STATE_PROPAGATORS = Cint(6)   # state branch of the MAS/DNP propagator helper
const PR_PAIRS_MODULUS = Cint(7)       # :modulus state pairs (Objectives.jl:222, 356-361)
const PR_FP64 = Cint(0)
const PR_MIXED = Cint(1)
const PR_SPLIT16 = Cint(2)
const PR_GRAD_REFERENCE = Cint(0)
const PR_GRAD_EXACT = Cint(1)
const PR_SEGMENT_AUTOTUNE = Cint(-1)   # PROptions.segment: the library times candidate segment lengths

# Memory layout identical to `pr_problem` (include/pulsar_rocm.h).
struct PRProblem
    dim::Cint
    n_ctrl::Cint
    n_t::Cint
    dt::Cdouble
    dt_slices::Ptr{Cdouble}
    dynamics::Cint
    objective::Cint
    n_members::Cint
    H0::Ptr{Cdouble}
    H0_per_member::Cint
    Hc::Ptr{Cdouble}
    ctrl_scale::Ptr{Cdouble}
    weight::Ptr{Cdouble}
    U_target::Ptr{Cdouble}
    n_pairs::Cint
    psi_init::Ptr{Cdouble}
    psi_targ::Ptr{Cdouble}
    n_track::Cint
    track_step::Ptr{Cint}
    track_weight::Ptr{Cdouble}
    track_target::Ptr{Cdouble}
end

struct PROptions
    precision::Cint
    grad_mode::Cint
    segment::Cint
    store_props::Cint
    pack::Cint
end

"""
    rocm_init!(; lib = get(ENV, "PULSAR_ROCM_LIB", "libpulsar_rocm.so"), device = 0) -> Bool

Load libpulsar_rocm.so and select a gfx950 device. Called from `__init__`
in the style of `_detect_and_load_backends` (src/Pulsar.jl:806-836).
"""
function rocm_init!(; lib::AbstractString = get(ENV, "PULSAR_ROCM_LIB", "libpulsar_rocm.so"),
                    device::Integer = 0)::Bool
    _ROCM_LOADED[] = false
    h = Libc.Libdl.dlopen(lib; throw_error = false)
    h === nothing && return false
    _ROCM_LIB[] = lib
    ccall(Libc.Libdl.dlsym(h, :pr_abi_version), Cint, ()) == 2 || return false
    ccall(Libc.Libdl.dlsym(h, :pr_device_count), Cint, ()) > 0 || return false
    ccall(Libc.Libdl.dlsym(h, :pr_init), Cint, (Cint,), device) == 0 || return false
    _ROCM_LOADED[] = true
    return true
end

is_rocm_available() = _ROCM_LOADED[]

function _rocm_last_error()
    buf = Vector{UInt8}(undef, 512)
    ccall((:pr_last_error, _ROCM_LIB[]), Cint, (Ptr{UInt8}, Csize_t), buf, length(buf))
    return unsafe_string(pointer(buf))
end

_rocm_check(rc::Cint, what::String) =
    rc == 0 || error("libpulsar_rocm $what failed (status $rc): $(_rocm_last_error())")

"""
    ROCmPlan(probs, roots; grad_mode = 0, precision = 0, segment = 0, pack = 1)

The keywords are pr_options' fields (pack = 1: small problems and the members
of small ensembles run as diagonal blocks of 16 x 16 problems).
A device plan for one or more problems; owns device memory and is freed by a
finalizer (pr_plan_destroy). One Julia task at a time per plan.
"""
mutable struct ROCmPlan
    handle::Ptr{Cvoid}
    n_ctrl::Vector{Int}
    n_t::Vector{Int}
    function ROCmPlan(probs::Vector{PRProblem}, roots::Vector{Any};
                      grad_mode::Integer = 0, precision::Integer = 0, segment::Integer = 0,
                      pack::Integer = 1)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        opt = Ref(PROptions(precision, grad_mode, segment, -1, pack))
        GC.@preserve roots probs begin
            _rocm_check(ccall((:pr_plan_create, _ROCM_LIB[]), Cint,
                              (Ptr{PRProblem}, Cint, Ptr{PROptions}, Ptr{Ptr{Cvoid}}),
                              probs, length(probs), opt, h), "pr_plan_create")
        end
        p = new(h[], [Int(q.n_ctrl) for q in probs], [Int(q.n_t) for q in probs])
        finalizer(p) do x
            x.handle == C_NULL || ccall((:pr_plan_destroy, _ROCM_LIB[]), Cint, (Ptr{Cvoid},), x.handle)
            x.handle = C_NULL
        end
        return p
    end
end

"""
    rocm_eval(plan, controls::Vector{<:AbstractMatrix}; grad=true) -> (F::Vector, G::Vector{Matrix})

`controls[p]` is the `[n_ctrl × n_t]` control matrix of problem p (the
`ControlSequence.controls` view, materialised column-major).
"""
function rocm_eval(plan::ROCmPlan, controls::Vector{<:AbstractMatrix}; grad::Bool = true)
    n = length(plan.n_ctrl)
    cs = [Matrix{Float64}(c) for c in controls]
    F = zeros(Float64, n)
    Gs = [zeros(Float64, plan.n_ctrl[i], plan.n_t[i]) for i in 1:n]
    cptr = [pointer(c) for c in cs]
    gptr = grad ? [pointer(g) for g in Gs] : Ptr{Float64}[]
    GC.@preserve cs Gs cptr gptr begin
        _rocm_check(ccall((:pr_eval, _ROCM_LIB[]), Cint,
                          (Ptr{Cvoid}, Ptr{Ptr{Float64}}, Ptr{Float64}, Ptr{Ptr{Float64}}, Ptr{Float64}),
                          plan.handle, cptr, F, grad ? pointer(gptr) : C_NULL, C_NULL), "pr_eval")
    end
    return F, Gs
end

# -- building pr_problem structs from Pulsar types (zero-copy for ComplexF64 arrays) --------

_cptr(x::Array{ComplexF64}) = Ptr{Cdouble}(pointer(x))
_cptr(x::Array{Float64}) = pointer(x)
_cptr(::Nothing) = Ptr{Cdouble}(C_NULL)

"""
    rocm_problem(H0s, Hc, n_t, dt, target; scales=nothing, weights=nothing, state_normalize=true, fixed_point=false)

H0s: one drift (Matrix) or a Vector of member drifts; Hc: Vector of control
matrices. Returns `(PRProblem, roots)`; keep `roots` alive while the struct is used.
"""
function rocm_problem(H0s, Hc::Vector{<:AbstractMatrix}, n_t::Int, dt::Float64, target;
                      scales = nothing, weights = nothing, state_normalize::Bool = true,
                      fixed_point::Bool = false)
    d = size(Hc[1], 1)
    permember = H0s isa AbstractVector
    H0 = permember ? cat([ComplexF64.(h) for h in H0s]...; dims = 3) : ComplexF64.(H0s)
    M = permember ? length(H0s) : 1
    HcA = cat([ComplexF64.(h) for h in Hc]...; dims = 3)
    sc = scales === nothing ? nothing : Matrix{Float64}(permutedims(scales))   # [n_ctrl × M] = C [M][n_ctrl]
    w = weights === nothing ? nothing : Vector{Float64}(weights)
    roots = Any[H0, HcA, sc, w]
    if target.type == "unitary"
        Ut = ComplexF64.(target.target_unitary)
        push!(roots, Ut)
        prob = PRProblem(d, length(Hc), n_t, dt, C_NULL, PR_CLOSED, PR_GATE, M, _cptr(H0), permember,
                         _cptr(HcA), _cptr(sc), _cptr(w), _cptr(Ut), 0, C_NULL, C_NULL, 0, C_NULL, C_NULL, C_NULL)
    else
        ψt = ComplexF64.(target.target_state)
        ψi = (fixed_point || target.initial_state === nothing) ? ψt : ComplexF64.(target.initial_state)
        if state_normalize
            ψt = ψt / norm(ψt); ψi = ψi / norm(ψi)
        end
        push!(roots, ψi, ψt)
        w === nothing && (w = ones(M) ./ M; push!(roots, w))
        prob = PRProblem(d, length(Hc), n_t, dt, C_NULL, PR_CLOSED, PR_STATE, M, _cptr(H0), permember,
                         _cptr(HcA), _cptr(sc), _cptr(w), C_NULL, 1, _cptr(ψi), _cptr(ψt), 0, C_NULL, C_NULL, C_NULL)
    end
    return prob, roots
end
