# PTStokesAMDChainRules.jl -- a ChainRulesCore rule for the top-boundary
# velocity misfit of a converged 2D viscous Stokes solve,
#     J(ln μ) = ½ Σ_k (Vx[I_k, ny] - obs_k)²        (Vx[:, ny]: free-slip top row)
# whose pullback is the libpts adjoint (pts_adjoint_solve + pts_grad_logmu).
# Load after PTStokesAMD with ChainRulesCore in the environment.
module PTStokesAMDChainRules

using ChainRulesCore
using ..PTStokesAMD
const P = PTStokesAMD

"""
    Misfit(c, p, points, obs, setup!; ε=1e-10, nout=200, iterMax=10^7)

`setup!(c, logμ)` must reset the state to the initial conditions and set Mus,
Musτ, Gdτ and dτ_Rho for the given ln μ (as the scripts' setup does);
`points` are 1-based Vx column indices on the top row.
"""
struct Misfit
    c::P.Ctx; p::P.PtParams; points::Vector{Int}; obs::Vector{Float64}; setup!::Function
    ε::Float64; nout::Int; iterMax::Int
end
Misfit(c, p, points, obs, setup!; ε=1e-10, nout=200, iterMax=10^7) =
    Misfit(c, p, points, obs, setup!, ε, nout, iterMax)

function forward(m::Misfit, logμ)
    m.setup!(m.c, logμ)
    P.pt_solve!(m.c, m.p, m.ε, m.nout, m.iterMax; norm_kind=1)
    Vx = P.download(m.c, :Vx)
    r = Vx[m.points, end] .- m.obs
    return 0.5 * sum(abs2, r), Vx, r
end

(m::Misfit)(logμ) = forward(m, logμ)[1]

function ChainRulesCore.rrule(m::Misfit, logμ)
    J, Vx, r = forward(m, logμ)
    g = zeros(size(Vx))
    for (k, i) in enumerate(m.points)
        g[i, end] += r[k]
    end
    P.adjoint_set_source!(m.c, g, nothing)
    P.adjoint_solve!(m.c, m.p, m.ε, m.nout, m.iterMax)
    grad = P.grad_logμ(m.c, m.p, size(logμ)...)
    misfit_pullback(ΔJ) = (NoTangent(), unthunk(ΔJ) .* grad)
    return J, misfit_pullback
end

end # module
