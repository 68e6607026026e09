# common.jl -- shared driver of the script entry points in this folder: the
# reference scripts' ENV knobs, their set-up pipeline and PT loops, run on
# libpts through PTStokesAMD (one GPU; the multi-XPU scripts as one rank).
using Printf
include(joinpath(@__DIR__, "..", "PTStokesAMD.jl"))
using .PTStokesAMD
const PT = PTStokesAMD

# ENV knob with a default (README.md:71-79: USE_RETURN, DO_SAVE, DO_SAVE_VIZ, NX, NY, NZ)
knob(name, ::Type{T}, default) where {T} = haskey(ENV, name) ? parse(T, ENV[name]) : default

Base.@kwdef struct Case
    ndim::Int; nx::Int; ny::Int; nz::Int = 1
    ve::Bool; fd::Int32; igg::Bool = false
    lx::Float64 = 10.0; ly::Float64 = 10.0; lz::Float64 = 10.0
    μs0::Float64 = 1.0; μsi::Float64 = 1e-3; rad2::Float64 = 1.0; εbg::Float64 = 1.0
    G::Float64 = 1.0; dt::Float64 = 1.0
    nt::Int; iterMax::Int; nout::Int
    Re::Float64; r::Float64; CFL::Float64; ε::Float64 = 1e-8; nsmooth::Int = 10
end

cellsizes(k::Case) = (k.lx / k.nx, k.ly / k.ny, k.ndim == 3 ? k.lz / k.nz : 1.0)
function centres(k::Case)
    dx, dy, dz = cellsizes(k)
    xc = LinRange(dx / 2, k.lx - dx / 2, k.nx)
    yc = LinRange(dy / 2, k.ly - dy / 2, k.ny)
    return k.ndim == 2 ? (xc, yc) : (xc, yc, LinRange(dz / 2, k.lz - dz / 2, k.nz))
end
params(k::Case) = PT.PtParams(k.r, k.G, k.dt, cellsizes(k)..., k.lx)

# ICs with the inclusion, smooth! with swap, Musτ = maxloc(Mus) + free-slip
# copies, iteration parameters (Stokes2D_ve.jl:131-147, Stokes3D_ve_perf.jl:167-192)
function setup(k::Case)
    c = PT.Ctx(k.ndim, k.nx, k.ny, k.nz; ve=k.ve, fd_mode=k.fd)
    PT.init_inclusion!(c, k.lx, k.ly, k.lz, k.μs0, k.μsi, k.rad2, k.εbg; igg=k.igg)
    PT.copy!(c, :Mus2, :Mus)
    for _ in 1:k.nsmooth
        PT.smooth!(c, :Mus2, :Mus, 1.0)
        PT.swap!(c, :Mus, :Mus2)
    end
    PT.copy!(c, :Musτ, :Mus)
    PT.compute_maxloc!(c, :Musτ, :Mus)
    PT.bc_x!(c, :Musτ); PT.bc_y!(c, :Musτ)
    k.ndim == 3 && PT.bc_z!(c, :Musτ)
    dx, dy, dz = cellsizes(k)
    h = k.ndim == 3 ? min(dx, dy, dz) : min(dx, dy)
    PT.compute_iter_params!(c, h * k.CFL, k.G, k.dt, k.Re, k.r, k.ndim == 3 ? max(k.lx, k.ly, k.lz) : max(k.lx, k.ly))
    return c
end

# nt physical steps, each PT loop to err < ε on the device (checks every
# nout); the clock starts with the loop (t_it = t/(ittot-10) as in the scripts)
function run_converged(c, k::Case)
    p = params(k); ittot = 0
    PT.sync(c); t0 = time()
    for it in 1:k.nt
        k.ve && PT.assign_τ!(c)
        res = PT.pt_solve!(c, p, k.ε, k.nout, k.iterMax)
        ittot += res.iters
        @printf("Step = %d, iter = %d, err = %1.3e \n", it, res.iters, res.err)
    end
    PT.sync(c)
    return ittot, time() - t0
end

# the perf scripts: one step of iterMax+1 iterations, the clock from iteration 11
function run_fixed(c, k::Case)
    p = params(k)
    k.ve && PT.assign_τ!(c)
    total = k.iterMax + 1
    PT.pt_iterate!(c, p, 10); PT.sync(c)
    t0 = time()
    PT.pt_iterate!(c, p, total - 10); PT.sync(c)
    return total, time() - t0
end

report(ittot, nt, t, T_eff) =
    @printf("Total iters = %d (%d steps), time = %1.3e sec (@ T_eff = %1.2f GB/s) \n", ittot, nt, t, T_eff)

# DO_SAVE: one line appended to ../output/<file>, the fields space-separated
function save_line(file, vals...)
    !ispath("../output") && mkdir("../output")
    open(joinpath("../output", file), "a") do io
        println(io, join(string.(vals), " "))
    end
end

# DO_SAVE_VIZ: the scripts' .mat keys (needs MAT.jl, as the reference does)
function save_viz(file, d::Dict)
    @eval using MAT
    !ispath("../out_visu") && mkdir("../out_visu")
    Base.invokelatest(matwrite, joinpath("../out_visu", file), d; compress=true)
end
