# Entry point of scripts/Stokes2D.jl (2D viscous, 4-kernel form, converged)
# on libpts.  ENV: NX, NY, USE_RETURN, DO_SAVE, DO_SAVE_VIZ (Stokes2D.jl:1-7).
include(joinpath(@__DIR__, "common.jl"))

function Stokes2D_()
    # Stokes2D.jl:74-89
    k = Case(ndim=2, nx=knob("NX", Int, 255), ny=knob("NY", Int, 255), ve=false, fd=PT.FD_DIV,
             nt=1, iterMax=100000, nout=500, Re=5π, r=1.0, CFL=0.9 / sqrt(2))
    c = setup(k)
    ittot, t = run_converged(c, k)
    A_eff = (3 * 2) / 1e9 * k.nx * k.ny * 8          # Stokes2D.jl:150
    T_eff = A_eff / (t / (ittot - 10))
    report(ittot, k.nt, t, round(T_eff, sigdigits=2))
    knob("DO_SAVE", Bool, false) && save_line("out_Stokes2D.txt", k.nx, k.ny, ittot)
    dx, dy, _ = cellsizes(k)
    knob("DO_SAVE_VIZ", Bool, false) && save_viz("Stokes_2D.mat", Dict("Pt_2D" => PT.download(c, :Pt),
        "Mus_2D" => PT.download(c, :Mus), "Txy_2D" => PT.download(c, :τxy), "Vy_2D" => PT.download(c, :Vy),
        "dx_2D" => dx, "dy_2D" => dy))
    return (centres(k)..., PT.download(c, :Pt))
end

knob("USE_RETURN", Bool, false) || Stokes2D_()
