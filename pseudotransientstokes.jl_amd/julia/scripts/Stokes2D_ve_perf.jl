# Entry point of scripts/Stokes2D_ve_perf.jl (2D Maxwell visco-elastic, fused
# form, one step of iterMax+1 = 101 iterations, timed from iteration 11) on
# libpts.  ENV: NX, NY, USE_RETURN, DO_SAVE, DO_SAVE_VIZ (Stokes2D_ve_perf.jl:1-7).
include(joinpath(@__DIR__, "common.jl"))

function Stokes2D_ve_perf_()
    # Stokes2D_ve_perf.jl:82-102
    k = Case(ndim=2, nx=knob("NX", Int, 255), ny=knob("NY", Int, 255), ve=true, fd=PT.FD_MUL,
             nt=1, iterMax=100, nout=2000, Re=5π, r=1.0, CFL=0.8 / sqrt(2))
    c = setup(k)
    ittot, t_toc = run_fixed(c, k)
    A_eff = (6 * 2 + 3 * 1 + 3) / 1e9 * k.nx * k.ny * 8   # Stokes2D_ve_perf.jl:173
    t_it = t_toc / (ittot - 10)
    T_eff = A_eff / t_it
    report(ittot, k.nt, t_toc, round(T_eff, sigdigits=3))
    knob("DO_SAVE", Bool, false) && save_line("out_Stokes2D_ve_perf.txt", k.nx, k.ny, ittot, t_toc, A_eff, t_it, T_eff)
    dx, dy, _ = cellsizes(k)
    knob("DO_SAVE_VIZ", Bool, false) && save_viz("Stokes_2D_ve_perf.mat", Dict("Pt_2D" => PT.download(c, :Pt),
        "Mus_2D" => PT.download(c, :Mus), "Txy_2D" => PT.download(c, :τxy), "Vy_2D" => PT.download(c, :Vy),
        "dx_2D" => dx, "dy_2D" => dy))
    return (centres(k)..., PT.download(c, :Pt))
end

knob("USE_RETURN", Bool, false) || Stokes2D_ve_perf_()
