# Entry point of scripts/stokes_3D/Stokes3D_ve_perf.jl (3D Maxwell
# visco-elastic, fused form, one step of iterMax+1 = 201 iterations, timed
# from iteration 11) on libpts.  ENV: NX, NY, NZ, USE_RETURN, DO_SAVE,
# DO_SAVE_VIZ (Stokes3D_ve_perf.jl:1-8).
include(joinpath(@__DIR__, "common.jl"))

function Stokes3D_ve_perf_()
    # Stokes3D_ve_perf.jl:118-138
    k = Case(ndim=3, nx=knob("NX", Int, 63), ny=knob("NY", Int, 63), nz=knob("NZ", Int, 63), ve=true,
             fd=PT.FD_MUL, nt=1, iterMax=200, nout=2000, Re=6.0π, r=1.0, CFL=0.8 / sqrt(3))
    c = setup(k)
    ittot, wtime = run_fixed(c, k)
    A_eff = (10 * 2 + 6 * 1 + 3) / 1e9 * k.nx * k.ny * k.nz * 8   # Stokes3D_ve_perf.jl:229
    wtime_it = wtime / (ittot - 10)
    T_eff = A_eff / wtime_it
    report(ittot, k.nt, wtime, round(T_eff, sigdigits=3))
    knob("DO_SAVE", Bool, false) &&
        save_line("out_Stokes3D_ve_perf.txt", 1, k.nx, k.ny, k.nz, ittot, wtime, A_eff, wtime_it, T_eff)
    dx, dy, dz = cellsizes(k)
    knob("DO_SAVE_VIZ", Bool, false) && save_viz("Stokes_3D_ve_perf.mat", Dict("Pt_3D" => PT.download(c, :Pt),
        "Mus_3D" => PT.download(c, :Mus), "Txz_3D" => PT.download(c, :τxz), "Vz_3D" => PT.download(c, :Vz),
        "dx_3D" => dx, "dy_3D" => dy, "dz_3D" => dz))
    return (centres(k)..., PT.download(c, :Pt))
end

knob("USE_RETURN", Bool, false) || Stokes3D_ve_perf_()
