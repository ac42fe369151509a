# ParticleInCellChainRulesExt.jl -- a package extension (loaded when
# ChainRulesCore is present) that puts libpicmi355x's discrete adjoint
# (pic_adjoint) behind a ChainRulesCore rrule, Julia's counterpart of a
# custom autograd function.
#
# field_energy(p.array, mesh, kernel, dt, nsteps) is J = sum(ex^2 + ey^2) dx dy
# of the last step of the hot-path loop (examples/landau_damping_poisson_cic_m6.jl
# :70-83); its pullback returns dJ/dp.array (rows x, y, vx, vy; the weights
# carry no gradient).  The gradient with respect to the Landau amplitude
# alpha follows through the sampler: x0 = newton(r) with P(x) = x +
# alpha sin(k x)/k gives dx0/dalpha = -(sin(k x0)/k) / (1 + alpha cos(k x0)).
module ParticleInCellChainRulesExt

using ChainRulesCore, ParticleInCell
import ParticleInCell.MI355X: DeviceParticles, lib, code, check, set_fp32!

"J = sum(ex^2 + ey^2) dx dy after nsteps of the Landau loop, as a function of p.array."
function field_energy(array::Matrix{Float64}, mesh, kernel, dt, nsteps; sort_every = 4,
                      fp32 = false)
    d = DeviceParticles(array, mesh)
    set_fp32!(d, fp32)
    e = zeros(2, nsteps)
    check(ccall((:pic_run, lib), Cint, (Ptr{Cvoid}, Cint, Float64, Cint, Cint, Ptr{Float64}),
                d.h, code(kernel), dt, nsteps, sort_every, e))
    e[1, end] + e[2, end]
end

function ChainRulesCore.rrule(::typeof(field_energy), array, mesh, kernel, dt, nsteps;
                              sort_every = 4, fp32 = false, ckpt = 20)
    d = DeviceParticles(array, mesh)
    set_fp32!(d, fp32)
    n = size(array, 2)
    J = Ref(0.0)
    gx = zeros(2, n)
    gv = zeros(2, n)
    check(ccall((:pic_adjoint, lib), Cint,
                (Ptr{Cvoid}, Cint, Float64, Cint, Cint, Cint, Ref{Float64}, Ptr{Float64},
                 Ptr{Float64}),
                d.h, code(kernel), dt, nsteps, ckpt, sort_every, J, gx, gv))
    function pullback(Jbar)
        g = zeros(5, n)
        g[1:2, :] .= Jbar .* gx
        g[3:4, :] .= Jbar .* gv
        (NoTangent(), g, NoTangent(), NoTangent(), NoTangent(), NoTangent())
    end
    J[], pullback
end

end  # module ParticleInCellChainRulesExt
