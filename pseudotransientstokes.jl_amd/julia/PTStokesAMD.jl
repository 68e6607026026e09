# PTStokesAMD.jl -- Julia binding of libpts (include/pts.h) for the reference
# scripts of PseudoTransientStokes.jl.
#
# Each function replaces one `@parallel kernel!(arrays..., scalars...)` call of
# the reference with a `ccall` into libpts; the arrays live in the context.
# Julia is not available where this repository is built and tested, so this
# file is exercised only through the same C ABI from Python (ptstokes).
module PTStokesAMD

const libpts = get(ENV, "PTS_LIB", joinpath(@__DIR__, "..", "libpts.so"))

const FD_DIV = Int32(0)   # 4-kernel scripts: (a-b)/dx
const FD_MUL = Int32(1)   # *_perf scripts:   _dx*(a-b)

const FIELDS = (:Pt, :Vx, :Vy, :Vz, :τxx, :τyy, :τzz, :τxy, :τxz, :τyz,
                :τxx_o, :τyy_o, :τzz_o, :τxy_o, :τxz_o, :τyz_o,
                :Mus, :Mus2, :Musτ, :Gdτ, :dτ_Rho, :∇V, :Rx, :Ry, :Rz, :dVx, :dVy, :dVz, :ρg)
fid(s::Symbol) = Int32(findfirst(==(s), FIELDS) - 1)

struct Grid
    ndim::Int32; nx::Int32; ny::Int32; nz::Int32
    ve::Int32; fd_mode::Int32; body_force::Int32; device::Int32
end

struct PtParams
    r::Float64; G::Float64; dt::Float64; dx::Float64; dy::Float64; dz::Float64; lx::Float64
end

struct SolveParams
    eps::Float64; nout::Int64; iter_max::Int64; norm_kind::Int32; use_graph::Int32
end

mutable struct SolveResult
    iters::Int64; err::Float64; comp::NTuple{4,Float64}; converged::Int32; nchecks::Int32
end
SolveResult() = SolveResult(0, 0.0, (0.0, 0.0, 0.0, 0.0), 0, 0)

mutable struct Ctx
    h::Ptr{Cvoid}
    ndim::Int
end

lasterr(c) = unsafe_string(ccall((:pts_last_error, libpts), Cstring, (Ptr{Cvoid},), c === nothing ? C_NULL : c.h))
chk(c, rc) = rc == 0 ? nothing : error("libpts error $rc: $(lasterr(c))")

function Ctx(ndim, nx, ny, nz=1; ve=true, fd_mode=FD_MUL, body_force=false, device=0)
    g = Ref(Grid(ndim, nx, ny, nz, ve, fd_mode, body_force, device))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    chk(nothing, ccall((:pts_create, libpts), Cint, (Ref{Grid}, Ref{Ptr{Cvoid}}), g, h))
    c = Ctx(h[], ndim)
    finalizer(x -> ccall((:pts_destroy, libpts), Cint, (Ptr{Cvoid},), x.h), c)
    return c
end

function shape(c::Ctx, f::Symbol)
    s = zeros(Int64, 3)
    chk(c, ccall((:pts_field_shape, libpts), Cint, (Ptr{Cvoid}, Cint, Ptr{Int64}), c.h, fid(f), s))
    return Tuple(s[1:c.ndim])
end
# Data.Array(A) / Array(A): host <-> device, column-major as in the scripts
upload!(c::Ctx, f::Symbol, A::Array{Float64}) =
    chk(c, ccall((:pts_upload, libpts), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Int64), c.h, fid(f), A, length(A)))
function download(c::Ctx, f::Symbol)
    A = zeros(Float64, shape(c, f))
    chk(c, ccall((:pts_download, libpts), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Int64), c.h, fid(f), A, length(A)))
    return A
end
copy!(c::Ctx, dst::Symbol, src::Symbol) = chk(c, ccall((:pts_copy, libpts), Cint, (Ptr{Cvoid}, Cint, Cint), c.h, fid(dst), fid(src)))
swap!(c::Ctx, a::Symbol, b::Symbol) = chk(c, ccall((:pts_swap, libpts), Cint, (Ptr{Cvoid}, Cint, Cint), c.h, fid(a), fid(b)))
# @zeros / @ones: A .= v on the device
fill!(c::Ctx, f::Symbol, v) = chk(c, ccall((:pts_fill, libpts), Cint, (Ptr{Cvoid}, Cint, Float64), c.h, fid(f), v))
sync(c::Ctx) = chk(c, ccall((:pts_sync, libpts), Cint, (Ptr{Cvoid},), c.h))
# the context's hipStream_t, for callers that order their own work against it
function stream(c::Ctx)
    s = Ref{Ptr{Cvoid}}(C_NULL)
    chk(c, ccall((:pts_get_stream, libpts), Cint, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}), c.h, s))
    return s[]
end
function device_bytes(c::Ctx)
    b = Ref{Int64}(0)
    chk(c, ccall((:pts_device_bytes, libpts), Cint, (Ptr{Cvoid}, Ref{Int64}), c.h, b))
    return b[]
end
# iteration times (ms) of the field placements tried at the first solve
function placement_times(c::Ctx, n=32)
    ms = zeros(Float32, n)
    chk(c, ccall((:pts_placement_times, libpts), Cint, (Ptr{Cvoid}, Ptr{Float32}, Int32), c.h, ms, n))
    return ms
end
# launch-configuration knob by its environment-variable name (no counterpart in the reference)
set_knob!(c::Ctx, name::AbstractString, value::Integer) =
    chk(c, ccall((:pts_set_knob, libpts), Cint, (Ptr{Cvoid}, Cstring, Int64), c.h, name, value))

# setup kernels
init_inclusion!(c::Ctx, lx, ly, lz, μs0, μsi, rad2, εbg, ρg_in=0.0, ρg_out=0.0; igg=false) =
    chk(c, ccall((:pts_init_inclusion, libpts), Cint,
                 (Ptr{Cvoid}, Float64, Float64, Float64, Float64, Float64, Float64, Float64, Float64, Float64, Int32),
                 c.h, lx, ly, lz, μs0, μsi, rad2, εbg, ρg_in, ρg_out, igg))
# seeded Gaussian random field for log10 μ: modes is a 4×n matrix of (kx, ky, kz, φ)
init_grf_viscosity!(c::Ctx, modes::Matrix{Float64}, σ, lo, hi, lx, ly, lz, μsi, rad2) =
    chk(c, ccall((:pts_init_grf_viscosity, libpts), Cint,
                 (Ptr{Cvoid}, Int32, Ptr{Float64}, Float64, Float64, Float64, Float64, Float64, Float64, Float64, Float64),
                 c.h, size(modes, 2), modes, σ, lo, hi, lx, ly, lz, μsi, rad2))
smooth!(c::Ctx, A2::Symbol, A::Symbol, fact) =
    chk(c, ccall((:pts_smooth, libpts), Cint, (Ptr{Cvoid}, Cint, Cint, Float64), c.h, fid(A2), fid(A), fact))
compute_maxloc!(c::Ctx, A2::Symbol, A::Symbol) =
    chk(c, ccall((:pts_maxloc, libpts), Cint, (Ptr{Cvoid}, Cint, Cint), c.h, fid(A2), fid(A)))
bc_x!(c::Ctx, A::Symbol) = chk(c, ccall((:pts_bc_x, libpts), Cint, (Ptr{Cvoid}, Cint), c.h, fid(A)))
bc_y!(c::Ctx, A::Symbol) = chk(c, ccall((:pts_bc_y, libpts), Cint, (Ptr{Cvoid}, Cint), c.h, fid(A)))
bc_z!(c::Ctx, A::Symbol) = chk(c, ccall((:pts_bc_z, libpts), Cint, (Ptr{Cvoid}, Cint), c.h, fid(A)))
compute_iter_params!(c::Ctx, Vpdτ, G, dt, Re, r, max_l) =
    chk(c, ccall((:pts_iter_params, libpts), Cint, (Ptr{Cvoid}, Float64, Float64, Float64, Float64, Float64, Float64),
                 c.h, Vpdτ, G, dt, Re, r, max_l))
assign_τ!(c::Ctx) = chk(c, ccall((:pts_assign_tau, libpts), Cint, (Ptr{Cvoid},), c.h))

# 4-kernel form (Stokes2D.jl, Stokes2D_ve.jl, Stokes3D_ve_multixpu.jl)
compute_P!(c::Ctx, r, dx, dy, dz=1.0) =
    chk(c, ccall((:pts_compute_P, libpts), Cint, (Ptr{Cvoid}, Float64, Float64, Float64, Float64), c.h, r, dx, dy, dz))
compute_τ!(c::Ctx, G, dt, dx, dy, dz=1.0) =
    chk(c, ccall((:pts_compute_tau, libpts), Cint, (Ptr{Cvoid}, Float64, Float64, Float64, Float64, Float64), c.h, G, dt, dx, dy, dz))
compute_dV!(c::Ctx, dx, dy, dz=1.0) =
    chk(c, ccall((:pts_compute_dV, libpts), Cint, (Ptr{Cvoid}, Float64, Float64, Float64), c.h, dx, dy, dz))
compute_V!(c::Ctx) = chk(c, ccall((:pts_compute_V, libpts), Cint, (Ptr{Cvoid},), c.h))
compute_Res!(c::Ctx, dx, dy, dz=1.0) =
    chk(c, ccall((:pts_compute_Res, libpts), Cint, (Ptr{Cvoid}, Float64, Float64, Float64), c.h, dx, dy, dz))

# fused perf form (Stokes2D_ve_perf.jl, Stokes3D_ve_perf.jl)
compute_Pt_τ!(c::Ctx, r, G, dt, _dx, _dy, _dz=1.0) =
    chk(c, ccall((:pts_compute_Pt_tau, libpts), Cint, (Ptr{Cvoid}, Float64, Float64, Float64, Float64, Float64, Float64),
                 c.h, r, G, dt, _dx, _dy, _dz))
compute_V_perf!(c::Ctx, _dx, _dy, _dz=1.0) =
    chk(c, ccall((:pts_compute_V_perf, libpts), Cint, (Ptr{Cvoid}, Float64, Float64, Float64), c.h, _dx, _dy, _dz))
compute_Res_perf!(c::Ctx, _dx, _dy, _dz=1.0) =
    chk(c, ccall((:pts_compute_Res_perf, libpts), Cint, (Ptr{Cvoid}, Float64, Float64, Float64), c.h, _dx, _dy, _dz))

# residual norms of the stored Rx, Ry, Rz, ∇V and the ranges of Vx and Pt
# (Stokes2D_ve.jl:164-168); err from them (kind 0: the scripts' scaled RMS,
# kind 1: max |R|)
mutable struct Norms
    Vmin::Float64; Vmax::Float64; Pmin::Float64; Pmax::Float64
    sumsq::NTuple{4,Float64}; maxabs::NTuple{4,Float64}; len::NTuple{4,Int64}
end
Norms() = Norms(0.0, 0.0, 0.0, 0.0, (0.0, 0.0, 0.0, 0.0), (0.0, 0.0, 0.0, 0.0), (0, 0, 0, 0))
function norms(c::Ctx)
    n = Ref(Norms())
    chk(c, ccall((:pts_norms_stored, libpts), Cint, (Ptr{Cvoid}, Ref{Norms}), c.h, n))
    return n[]
end
# The same reductions over caller-owned device arrays (reference shapes), for
# callers that cannot allocate (no counterpart in the reference): enqueued on
# `stream`, 12 doubles into the device buffer `out`.
function norms_dense_noalloc!(out, Vx, Pt, Rx, Ry, Rz, ∇V, ndim, nx, ny, nz; stream=C_NULL)
    rc = ccall((:pts_norms_dense_noalloc, libpts), Cint,
               (Int32, Int32, Int32, Int32, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble},
                Ptr{Cdouble}, Ptr{Cvoid}, Ptr{Cdouble}),
               ndim, nx, ny, nz, Vx, Pt, Rx, Ry, Rz, ∇V, stream, out)
    rc == 0 || error("libpts error $rc: " * unsafe_string(ccall((:pts_last_error, libpts), Cstring, (Ptr{Cvoid},), C_NULL)))
    return out
end
function err_from_norms(n::Norms, ndim, lx; kind=0)
    comp = zeros(Float64, 4)
    err = ccall((:pts_err_from_norms, libpts), Float64, (Ref{Norms}, Int32, Float64, Int32, Ptr{Float64}),
                Ref(n), ndim, lx, kind, comp)
    return err, comp
end

# device-resident PT loop of one physical step
function pt_solve!(c::Ctx, p::PtParams, ε, nout, iterMax; norm_kind=0, use_graph=true)
    sp = Ref(SolveParams(ε, nout, iterMax, norm_kind, use_graph))
    res = Ref(SolveResult())
    rc = ccall((:pts_pt_solve, libpts), Cint, (Ptr{Cvoid}, Ref{PtParams}, Ref{SolveParams}, Ref{SolveResult}),
               c.h, Ref(p), sp, res)
    rc == -3 && error("NaN")   # the scripts' `if isnan(err) error("NaN") end`
    rc in (0, -4) || chk(c, rc)
    return res[]
end
pt_iterate!(c::Ctx, p::PtParams, n) =
    chk(c, ccall((:pts_pt_iterate, libpts), Cint, (Ptr{Cvoid}, Ref{PtParams}, Int64), c.h, Ref(p), n))

# FP32 PT sweeps inside FP64 defect correction (viscous grids)
struct DcParams
    eps::Float64; norm_kind::Int32; max_outer::Int32; rtol_inner::Float64; nout_inner::Int64; iter_max_inner::Int64
end
mutable struct DcResult
    outer::Int32; converged::Int32; inner_iters::Int64; err::Float64; err_outer::NTuple{32,Float64}
end
DcResult() = DcResult(0, 0, 0, 0.0, ntuple(_ -> 0.0, 32))
function dc_solve!(c::Ctx, p::PtParams, ε; norm_kind=1, max_outer=20, rtol_inner=1e-4, nout_inner=200,
                   iter_max_inner=10^7)
    dp = Ref(DcParams(ε, norm_kind, max_outer, rtol_inner, nout_inner, iter_max_inner))
    res = Ref(DcResult())
    rc = ccall((:pts_dc_solve, libpts), Cint, (Ptr{Cvoid}, Ref{PtParams}, Ref{DcParams}, Ref{DcResult}),
               c.h, Ref(p), dp, res)
    rc in (0, -4) || chk(c, rc)
    return res[]
end

# adjoint of the converged 2D viscous iteration and ∂J/∂ln μ
function adjoint_set_source!(c::Ctx, dJ_dVx::Union{Nothing,Array{Float64}}, dJ_dVy::Union{Nothing,Array{Float64}})
    px = dJ_dVx === nothing ? C_NULL : pointer(dJ_dVx)
    py = dJ_dVy === nothing ? C_NULL : pointer(dJ_dVy)
    GC.@preserve dJ_dVx dJ_dVy chk(c, ccall((:pts_adjoint_set_source, libpts), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}, Int64), c.h, px,
        dJ_dVx === nothing ? 0 : length(dJ_dVx), py, dJ_dVy === nothing ? 0 : length(dJ_dVy)))
end
# λ of :Pt, :τxx, :τyy, :τxy, :Vx or :Vy (forward shapes)
adjoint_upload!(c::Ctx, f::Symbol, A::Array{Float64}) =
    chk(c, ccall((:pts_adjoint_upload, libpts), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}, Int64), c.h, fid(f), A, length(A)))
function adjoint_download(c::Ctx, f::Symbol)
    A = zeros(Float64, shape(c, f))
    chk(c, ccall((:pts_adjoint_download, libpts), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}, Int64), c.h, fid(f), A, length(A)))
    return A
end
# n applications of λ <- Φᵀλ + ∂J/∂x
adjoint_iterate!(c::Ctx, p::PtParams, n) =
    chk(c, ccall((:pts_adjoint_iterate, libpts), Cint, (Ptr{Cvoid}, Ref{PtParams}, Int64), c.h, Ref(p), n))
function adjoint_solve!(c::Ctx, p::PtParams, ε, nout, iterMax)
    sp = Ref(SolveParams(ε, nout, iterMax, 0, 0))
    res = Ref(SolveResult())
    rc = ccall((:pts_adjoint_solve, libpts), Cint, (Ptr{Cvoid}, Ref{PtParams}, Ref{SolveParams}, Ref{SolveResult}),
               c.h, Ref(p), sp, res)
    rc == -3 && error("NaN")
    rc in (0, -4) || chk(c, rc)
    return res[]
end
function grad_logμ(c::Ctx, p::PtParams, nx, ny)
    g = zeros(Float64, nx, ny)
    chk(c, ccall((:pts_grad_logmu, libpts), Cint, (Ptr{Cvoid}, Ref{PtParams}, Ptr{Float64}, Int64),
                 c.h, Ref(p), g, length(g)))
    return g
end

# visco-elastic time loop: checkpoint each step's converged τ, then the
# backward pass over all steps (τ of step n is τ_o of step n+1)
checkpoint!(c::Ctx, slot) = chk(c, ccall((:pts_checkpoint, libpts), Cint, (Ptr{Cvoid}, Int32), c.h, slot))
function adjoint_ve_solve!(c::Ctx, p::PtParams, nsteps, ε, nout, iterMax, nx, ny)
    sp = Ref(SolveParams(ε, nout, iterMax, 0, 0))
    res = Ref(SolveResult())
    g = zeros(Float64, nx, ny)
    rc = ccall((:pts_adjoint_ve_solve, libpts), Cint,
               (Ptr{Cvoid}, Ref{PtParams}, Int32, Ref{SolveParams}, Ref{SolveResult}, Ptr{Float64}, Int64),
               c.h, Ref(p), nsteps, sp, res, g, length(g))
    rc == -3 && error("NaN")
    rc in (0, -4) || chk(c, rc)
    return g, res[]
end

end # module
