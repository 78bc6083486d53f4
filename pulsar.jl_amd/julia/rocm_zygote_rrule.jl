# rocm_zygote_rrule.jl — the ChainRulesCore rrule of the :rocm fidelity, included
# at the end of ext/PulsarZygoteExt.jl (which already has `using Pulsar` and
# `using Zygote`):
#
#   include(joinpath(pkgdir(Pulsar), "src", "Backend", "Hardware", "rocm_zygote_rrule.jl"))
#
# reverse_diff_gradient (ext/PulsarZygoteExt.jl:7-36) differentiates
# u -> _fidelity_flat(u, system, target, nc, nt, dt) (src/Physics/AutoDiff.jl:254-280)
# with Zygote.gradient. On the :rocm device it differentiates the device form
# Pulsar._rocm_fidelity_flat (same arguments, same value) instead — a one-line
# change there:
#
#   fflat = (Pulsar.get_device() === :rocm && Pulsar.is_rocm_available()) ?
#           Pulsar._rocm_fidelity_flat : Pulsar._fidelity_flat
#   f = u_flat -> fflat(u_flat, system, target, nc, nt, dt)
#
# and this rule gives Zygote the primal F and the pullback's gradient from one
# pr_eval with PR_GRAD_EXACT: the Frechet derivative of every slice's exp, the
# derivative Zygote computes through the eigen-exp, instead of a trace through
# n_t eigendecompositions. A library error propagates, so reverse_diff_gradient's
# own catch falls back to finite differences as it does for a Zygote failure
# (ext/PulsarZygoteExt.jl:19-25). The gradient is dF/du for real controls
# (_fidelity_flat reads real(u_flat[i])).

const _CRC = Zygote.ChainRulesCore

function _CRC.rrule(::typeof(Pulsar._rocm_fidelity_flat), u_flat::AbstractVector{<:Real},
                    system::Pulsar.AbstractQuantumSystem, target::Pulsar.QuantumTarget,
                    n_controls::Int, n_timesteps::Int, dt::Float64)
    F, g = Pulsar._rocm_fidelity_flat_and_gradient(u_flat, system, target, n_controls, n_timesteps, dt)
    function _rocm_fidelity_flat_pullback(ΔF)
        Δ = _CRC.unthunk(ΔF)
        return (_CRC.NoTangent(), Δ .* g, _CRC.NoTangent(), _CRC.NoTangent(), _CRC.NoTangent(),
                _CRC.NoTangent(), _CRC.NoTangent())
    end
    return F, _rocm_fidelity_flat_pullback
end
