# Entry point of scripts/stokes_3D/Stokes3D_ve_multixpu.jl on one rank (the
# local block with physical boundaries; ImplicitGlobalGrid's x_g coordinates):
# 3D Maxwell visco-elastic, 4-kernel form, nt converged steps, on libpts.
# ENV: NX, NY, NZ, USE_RETURN, DO_SAVE, DO_SAVE_VIZ (Stokes3D_ve_multixpu.jl:1-8).
include(joinpath(@__DIR__, "common.jl"))

function Stokes3D_ve_multixpu_()
    # Stokes3D_ve_multixpu.jl:120-141
    k = Case(ndim=3, nx=knob("NX", Int, 63), ny=knob("NY", Int, 63), nz=knob("NZ", Int, 63), ve=true,
             fd=PT.FD_DIV, igg=true, nt=5, iterMax=200000, nout=200, Re=6.0π, r=1.0, CFL=0.8 / sqrt(3))
    c = setup(k)
    ittot, t = run_converged(c, k)
    A_eff = (10 * 2 + 1) / 1e9 * k.nx * k.ny * k.nz * 8   # Stokes3D_ve_multixpu.jl:261
    T_eff = A_eff / (t / (ittot - 10))
    report(ittot, k.nt, t, round(T_eff, sigdigits=2))
    knob("DO_SAVE", Bool, false) && save_line("out_Stokes3D_ve.txt", k.nx, k.ny, k.nz, ittot, k.nt)
    if knob("DO_SAVE_VIZ", Bool, false)
        # the inner / averaged arrays the script gathers (Stokes3D_ve_multixpu.jl:25-28,267-271)
        Pt, Vz, Mus, τxz = PT.download(c, :Pt), PT.download(c, :Vz), PT.download(c, :Mus), PT.download(c, :τxz)
        inn(A) = A[2:end-1, 2:end-1, 2:end-1]
        av_zi(A) = (A[2:end-1, 2:end-1, 2:end-2] .+ A[2:end-1, 2:end-1, 3:end-1]) .* 0.5
        av_xza(A) = (A[1:end-1, :, 1:end-1] .+ A[1:end-1, :, 2:end] .+ A[2:end, :, 1:end-1] .+ A[2:end, :, 2:end]) .* 0.25
        dx, dy, dz = cellsizes(k)
        save_viz("Stokes_3D_ve.mat", Dict("Pt_3D" => inn(Pt), "Mus_3D" => inn(Mus), "Txz_3D" => av_xza(τxz),
                                          "Vz_3D" => av_zi(Vz), "dx_3D" => dx, "dy_3D" => dy, "dz_3D" => dz))
    end
    return (centres(k)..., PT.download(c, :Pt))
end

knob("USE_RETURN", Bool, false) || Stokes3D_ve_multixpu_()
