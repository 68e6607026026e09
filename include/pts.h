/*
 * pts.h -- C ABI of libpts, the MI355X (gfx950) pseudo-transient Stokes
 * library.
 *
 * The reference (PTsolvers/PseudoTransientStokes.jl) has no FFI: its hot
 * path is a set of ParallelStencil kernels called from Julia scripts with
 * the convention "@parallel kernel!(arrays..., scalars...)", every kernel
 * mutating its arrays in place.  This header replaces that call surface
 * 1:1.  The context owns every device array (the scripts' @zeros arrays);
 * scalars are passed by value exactly as the reference passes them.
 *
 * Conventions
 *  - Every entry point returns an int status: PTS_OK (0) or a negative
 *    PTS_ERR_* code; pts_last_error() gives the message.  The library never
 *    aborts or exits.
 *  - Host arrays are Julia column-major (x fastest) with exactly the
 *    reference shapes (pts_field_shape()).  Host pointers are borrowed for the
 *    duration of a call only.
 *  - One HIP stream per context, default priority.  Kernel entry points are
 *    asynchronous; pts_download, pts_norms, pts_pt_solve and pts_sync
 *    synchronise (the reference's @parallel synchronises after each launch,
 *    so results are observed in the same order).
 *  - A context is not thread-safe: the caller serialises calls, as the
 *    reference's single Julia task does.
 */
#ifndef PTS_H
#define PTS_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PTS_OK 0
#define PTS_ERR_ARG (-1)      /* bad argument (null, range, shape) */
#define PTS_ERR_HIP (-2)      /* HIP runtime error */
#define PTS_ERR_NAN (-3)      /* err is NaN: the reference's error("NaN") (Stokes2D_ve.jl:171) */
#define PTS_ERR_NOCONV (-4)   /* iterMax reached with err > eps (the reference ends silently) */
#define PTS_ERR_OOM (-5)      /* device allocation failed */
#define PTS_ERR_STATE (-6)    /* call not valid for this context (e.g. 3D field in 2D) */

/* Device arrays, named after the reference's variables (Stokes2D_ve.jl:111-130,
 * Stokes3D_ve_perf.jl:141-172). */
typedef enum pts_field {
    PTS_PT = 0, PTS_VX, PTS_VY, PTS_VZ,
    PTS_TXX, PTS_TYY, PTS_TZZ, PTS_TXY, PTS_TXZ, PTS_TYZ,
    PTS_TXX_O, PTS_TYY_O, PTS_TZZ_O, PTS_TXY_O, PTS_TXZ_O, PTS_TYZ_O,
    PTS_MUS, PTS_MUS2, PTS_MUSTAU, PTS_GDTAU, PTS_DTAU_RHO,
    PTS_DIVV, PTS_RX, PTS_RY, PTS_RZ, PTS_DVX, PTS_DVY, PTS_DVZ, PTS_RHOG,
    PTS_NFIELDS
} pts_field;

/* Finite-difference flavour.  The 4-kernel scripts divide by dx
 * (Stokes2D.jl, Stokes2D_ve.jl, Stokes3D_ve_multixpu.jl); the fused *_perf
 * scripts multiply by _dx = 1/dx (Stokes2D_ve_perf.jl:47-66,
 * Stokes3D_ve_perf.jl:69-96).  Both are reproduced bit for bit. */
#define PTS_FD_DIV 0
#define PTS_FD_MUL 1

typedef struct pts_grid {
    int32_t ndim;       /* 2 or 3 */
    int32_t nx, ny, nz; /* cell counts (nz ignored in 2D); each >= 3 */
    int32_t ve;         /* 1: Maxwell visco-elastic (uses τ_o, G, dt); 0: viscous */
    int32_t fd_mode;    /* PTS_FD_DIV or PTS_FD_MUL */
    int32_t body_force; /* 1: y- (2D) or z- (3D) momentum carries -avg(ρg)
                           (Stokes2D_param.jl:59) */
    int32_t device;     /* HIP device ordinal */
} pts_grid;

typedef struct pts_ctx pts_ctx;

/* ---- lifecycle and fields ------------------------------------------------ */
/* A context owns the arrays each script allocates with @zeros and
 * Data.Array (Stokes2D_ve.jl:111-136, Stokes3D_ve_perf.jl:141-172), addressed
 * by field id with the same shapes; fields are allocated on first use.
 * pts_last_error(NULL) reports a failed pts_create. */
int pts_create(const pts_grid* grid, pts_ctx** out);
int pts_destroy(pts_ctx* ctx);
const char* pts_last_error(const pts_ctx* ctx);
/* Reference shape of a field (3 entries; shape[2] = 1 in 2D). */
int pts_field_shape(const pts_ctx* ctx, int field, int64_t shape[3]);
/* Host <-> device copies of one field (count must equal the shape product). */
int pts_upload(pts_ctx* ctx, int field, const double* host, int64_t count);
int pts_download(pts_ctx* ctx, int field, double* host, int64_t count);
/* A .= value (the scripts' @zeros / ones initialisation). */
int pts_fill(pts_ctx* ctx, int field, double value);
/* dst .= src for two fields of equal shape (e.g. Musτ .= Mus). */
int pts_copy(pts_ctx* ctx, int dst, int src);
/* Swap two fields of equal shape ("Mus, Mus2 = Mus2, Mus", Stokes2D_ve.jl:140). */
int pts_swap(pts_ctx* ctx, int a, int b);
/* Wait for every call enqueued on this context.  No separate call in the
 * reference scripts: there a kernel's results are complete when @parallel
 * returns; here this is the point where the caller may rely on that. */
int pts_sync(pts_ctx* ctx);
/* No counterpart in the reference.  The HIP stream (hipStream_t) every kernel
 * of this context is enqueued on, so a caller can record events or order its
 * own work against it. */
int pts_get_stream(const pts_ctx* ctx, void** stream);
/* No counterpart in the reference.  Bytes of device memory held by the
 * context. */
int pts_device_bytes(const pts_ctx* ctx, int64_t* bytes);
/* No counterpart in the reference (placement of this library's device
 * blocks).  Ahead of its first long run -- a pts_pt_solve whose iter_max + 1,
 * or a pts_pt_iterate whose niter, is at least PTS_PLACE_BUDGET_ITERS, or a
 * pts_pt_iterate with niter = 0, which asks for it -- a single-pass context
 * whose field blocks total at least PTS_PLACE_MIN_MB (default 4096) times
 * real iterations on copies of its state in its own blocks and in up to
 * PTS_3D_PLACE_TRIES (default 5) / PTS_2D_PLACE_TRIES (default 31) fresh sets
 * of blocks, within a time budget of PTS_PLACE_BUDGET_ITERS (default 5000)
 * iterations (at least 1 s), and keeps the fastest set with the unchanged
 * state (the iteration speed depends on the blocks' physical placement).
 * Fills ms[0..n) with the iteration times of the sets tried: ms[0] the
 * original blocks, ms[t] try t, 0 for sets not tried (at most 32). */
int pts_placement_times(const pts_ctx* ctx, float* ms, int32_t n);

/* No counterpart in the reference.  Sets one launch-configuration knob of
 * the single-pass kernels by the name of the environment variable that sets
 * it at pts_create (PTS_PLANES_PER_BLOCK, PTS_3D_VARIANT, PTS_3D_MAP_FLAT,
 * PTS_3D_W64, PTS_ROWS_PER_BLOCK, PTS_2D_VARIANT, PTS_2D_PERSIST), so
 * that configurations can be compared on one set of device blocks.  Results
 * are bit-identical for every value; PTS_ERR_ARG for an unknown name. */
int pts_set_knob(pts_ctx* ctx, const char* name, int64_t value);

/* ---- setup kernels ------------------------------------------------------- */
/* Initial conditions of every reference script (Stokes2D_ve.jl:131-136,
 * Stokes3D_ve_perf.jl:167-172; ic_igg=1 uses ImplicitGlobalGrid's x_g()
 * coordinates, Stokes3D_ve_multixpu.jl:174-176): Mus = mus0 with musi where
 * Rad2 < rad2, ρg = rhog_in/rhog_out likewise, pure-shear Vx, Vy (2D) or
 * Vx, Vz (3D x-z shear), all other state fields zero. */
int pts_init_inclusion(pts_ctx* ctx, double lx, double ly, double lz, double mus0, double musi,
                       double rad2, double ebg, double rhog_in, double rhog_out, int32_t ic_igg);
/* Viscosity from a Gaussian random field (no counterpart in the reference):
 * log10 Mus = clip(sigma*sqrt(2/n)*sum_m cos(kx*x + ky*y + kz*z + phi), lo, hi)
 * at the cell centres, from n modes given as (kx, ky, kz, phi) quadruples;
 * Mus = musi where Rad2 < rad2 (the centred inclusion). */
int pts_init_grf_viscosity(pts_ctx* ctx, int32_t nmodes, const double* modes, double sigma, double lo, double hi,
                           double lx, double ly, double lz, double musi, double rad2);
/* smooth! (Stokes2D_ve.jl:18-21, Stokes3D_ve_perf.jl:24-27): dst[inn] = src[inn] + 1/4.1/fact*(d2x+d2y)
 * (1/6.1 in 3D). */
int pts_smooth(pts_ctx* ctx, int dst, int src, double fact);
/* compute_maxloc! (Stokes2D_ve.jl:23-26): dst[inn] = max over the 5/7-point cross of src. */
int pts_maxloc(pts_ctx* ctx, int dst, int src);
/* bc_x!/bc_y!/bc_z! (Stokes2D_ve.jl:75-85, Stokes3D_ve_perf.jl:98-114). */
int pts_bc_x(pts_ctx* ctx, int field);
int pts_bc_y(pts_ctx* ctx, int field);
int pts_bc_z(pts_ctx* ctx, int field);
/* compute_iter_params! (Stokes2D.jl:27-31 viscous; Stokes2D_ve.jl:28-33 VE). */
int pts_iter_params(pts_ctx* ctx, double Vpdtau, double G, double dt, double Re, double r, double max_l);
/* assign_τ! (Stokes2D_ve.jl:35-40): τ_o .= τ. */
int pts_assign_tau(pts_ctx* ctx);

/* ---- per-iteration kernels, 4-kernel form (divide by dx) --------------------- */
/* compute_P! (Stokes2D_ve.jl:42-46, Stokes3D_ve_multixpu.jl:57-61): ∇V, Pt. */
int pts_compute_P(pts_ctx* ctx, double r, double dx, double dy, double dz);
/* compute_τ! (Stokes2D.jl:39-44 viscous, Stokes2D_ve.jl:48-55 VE,
 * Stokes3D_ve_multixpu.jl:63-77). G, dt ignored when viscous. */
int pts_compute_tau(pts_ctx* ctx, double G, double dt, double dx, double dy, double dz);
/* compute_dV! (Stokes2D.jl:46-52: also stores Rx, Ry; Stokes2D_ve.jl:57-61;
 * Stokes3D_ve_multixpu.jl:79-84). */
int pts_compute_dV(pts_ctx* ctx, double dx, double dy, double dz);
/* compute_V! (Stokes2D.jl:54-58): V[inn] += dV. */
int pts_compute_V(pts_ctx* ctx);
/* compute_Res! (Stokes2D_ve.jl:69-73, Stokes3D_ve_multixpu.jl:93-98): Rx, Ry(, Rz). */
int pts_compute_Res(pts_ctx* ctx, double dx, double dy, double dz);

/* ---- per-iteration kernels, fused perf form (multiply by _dx) ---------------- */
/* compute_Pt_τ! (Stokes2D_ve_perf.jl:43-53, Stokes3D_ve_perf.jl:69-78). */
int pts_compute_Pt_tau(pts_ctx* ctx, double r, double G, double dt, double _dx, double _dy, double _dz);
/* compute_V! of the perf scripts (Stokes2D_ve_perf.jl:55-59, Stokes3D_ve_perf.jl:83-88). */
int pts_compute_V_perf(pts_ctx* ctx, double _dx, double _dy, double _dz);
/* compute_Res! of the perf scripts (Stokes2D_ve_perf.jl:61-66, Stokes3D_ve_perf.jl:90-96): ∇V, R. */
int pts_compute_Res_perf(pts_ctx* ctx, double _dx, double _dy, double _dz);

/* ---- residual norms ----------------------------------------------------------- */
typedef struct pts_norms {
    double Vmin, Vmax, Pmin, Pmax;   /* over all of Vx and Pt (incl. boundaries) */
    double sumsq[4];                 /* Σ R² for Rx, Ry, Rz, ∇V */
    double maxabs[4];                /* max |R| for Rx, Ry, Rz, ∇V */
    int64_t len[4];                  /* element counts */
} pts_norms;
/* Reductions of the stored Rx, Ry, Rz, ∇V, Vx, Pt arrays (Stokes2D_ve.jl:164-168). */
int pts_norms_stored(pts_ctx* ctx, pts_norms* out);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops).  The reductions of
 * pts_norms_stored over caller-owned device arrays in the reference's dense
 * column-major shapes (pts_field_shape of a context of this grid): Vx, Pt,
 * Rx, Ry, Rz (3D; may be NULL in 2D), ∇V.  Enqueued on `stream` (a
 * hipStream_t; NULL: the default stream) of the current device without
 * synchronising; d_out (device memory, 12 doubles) receives Vmin, Vmax, Pmin,
 * Pmax, Σ R² (Rx, Ry, Rz, ∇V) and max|R| (same order).  It calls no
 * hipMalloc or hipFree: the partial results live in a 24 KiB file-scope
 * device array, one per GPU, shared by every call on that GPU, so calls that
 * overlap on one GPU race.  Arrays of more than 2^31-1 elements, a bad ndim
 * or extent, or a null array give PTS_ERR_ARG (message in
 * pts_last_error(NULL)). */
int pts_norms_dense_noalloc(int32_t ndim, int32_t nx, int32_t ny, int32_t nz, const double* Vx, const double* Pt,
                            const double* Rx, const double* Ry, const double* Rz, const double* divV, void* stream,
                            double* d_out);
/* err from norms: max over components of ‖R‖₂/ΔP·lx/√len (kind 0, the
 * reference's err) or max|R|/ΔP·lx (kind 1); ∇V scaled by ΔV = Vmax-Vmin. */
double pts_err_from_norms(const pts_norms* n, int32_t ndim, double lx, int32_t kind, double comp[4]);

/* ---- device-resident PT iteration -------------------------------------------- */
typedef struct pts_pt_params {
    double r, G, dt;       /* reference kernel scalars */
    double dx, dy, dz;     /* cell sizes; the MUL flavour uses 1.0/dx computed in FP64 */
    double lx;             /* length scale of the err normalisation */
} pts_pt_params;

/* niter single-pass PT iterations: Pt/τ update, V update and free-slip BCs in
 * one kernel per iteration (ping-pong state).  Bit-identical to the
 * reference's per-iteration kernel sequence in the context's fd_mode. */
int pts_pt_iterate(pts_ctx* ctx, const pts_pt_params* p, int64_t niter);

typedef struct pts_solve_params {
    double eps;            /* tolerance ε on err */
    int64_t nout;          /* check every nout iterations */
    int64_t iter_max;      /* stop when iter > iter_max (the reference runs iter_max+1) */
    int32_t norm_kind;     /* 0: reference scaled-RMS err; 1: scaled max-norm */
    int32_t use_graph;     /* 1: replay each nout-iteration chunk as one HIP graph */
} pts_solve_params;

typedef struct pts_solve_result {
    int64_t iters;         /* PT iterations done (the reference's `iter`) */
    double err;            /* err at the last check */
    double comp[4];        /* its components (Rx, Ry, Rz, ∇V) */
    int32_t converged;     /* err <= eps */
    int32_t nchecks;
} pts_solve_result;

/* One physical step's PT loop: while err > eps && iter <= iter_max, iterate;
 * every nout iterations evaluate the residual norms on the device
 * (Stokes2D_ve.jl:153-175).  Returns PTS_ERR_NAN on NaN and PTS_ERR_NOCONV
 * when iter_max was reached (result still filled). */
int pts_pt_solve(pts_ctx* ctx, const pts_pt_params* p, const pts_solve_params* sp, pts_solve_result* res);

/* ---- mixed precision: FP32 PT sweeps inside FP64 defect correction ---------- */
/* No counterpart in the reference (FP64 only).  Viscous grids, multiply form.
 * Each outer step projects τ = 2μ ε̇(V), evaluates the FP64 residual (the stop
 * test uses it), solves the correction with FP32 single-pass PT sweeps until
 * its residual dropped by rtol_inner, and adds it to P and V in FP64.  An
 * inner solve also ends when 40 consecutive checks bring its residual no 1 %
 * below the best value so far (the FP32 roundoff floor, which grows with the
 * grid size, can lie above rtol_inner); the next correction then starts from
 * the FP64 residual that remains. */
typedef struct pts_dc_params {
    double eps;            /* FP64 stop: err <= eps */
    int32_t norm_kind;     /* 0: scaled RMS (reference err); 1: scaled max-norm */
    int32_t max_outer;     /* at most this many corrections */
    double rtol_inner;     /* inner stop: correction residual <= rtol_inner * initial */
    int64_t nout_inner;    /* inner check interval */
    int64_t iter_max_inner;
} pts_dc_params;

typedef struct pts_dc_result {
    int32_t outer;         /* corrections applied */
    int32_t converged;
    int64_t inner_iters;   /* FP32 sweeps in total */
    double err;            /* final FP64 err */
    double err_outer[32];  /* FP64 err before each correction (first 32) */
} pts_dc_result;

int pts_dc_solve(pts_ctx* ctx, const pts_pt_params* p, const pts_dc_params* dp, pts_dc_result* res);

/* ---- adjoint of the converged 2D PT iteration -------------------------------
 * No counterpart in the reference (the scripts have no adjoint).  Let Φ be one
 * PT iteration (compute_Pt_τ!, compute_V!, bc_y!(Vx), bc_x!(Vy) of
 * Stokes2D_ve_perf.jl:47-59,77-78; viscous: without the elastic terms), affine
 * in the state (Pt, τxx, τyy, τxy, Vx, Vy) for fixed μ, Gdτ, dτ_Rho (and τ_o).
 * For an objective J(Vx, Vy) of the converged state the adjoint λ solves
 * λ = Φᵀλ + ∂J/∂x; its iteration contracts like the forward one, so no
 * iteration history is kept.  Dirichlet (never-updated) velocity entries carry
 * no adjoint.  Requires a 2D context whose Mus, Gdτ, dτ_Rho are set up and whose
 * state is the converged forward solution (for pts_grad_logmu).  On a
 * visco-elastic grid these act on the last physical step with τ_o fixed;
 * pts_adjoint_ve_solve differentiates through all steps. */

/* Set ∂J/∂Vx ((nx+1)*ny) and ∂J/∂Vy (nx*(ny+1)), host arrays, x fastest;
 * NULL = zero; entries on Dirichlet faces are ignored.  Resets λ to zero. */
int pts_adjoint_set_source(pts_ctx* ctx, const double* dJ_dVx, int64_t nvx, const double* dJ_dVy, int64_t nvy);
/* λ of field PT, TXX, TYY, TXY, VX or VY (forward shapes), host arrays. */
int pts_adjoint_upload(pts_ctx* ctx, int32_t field, const double* host, int64_t count);
int pts_adjoint_download(pts_ctx* ctx, int32_t field, double* host, int64_t count);
/* niter applications of λ <- Φᵀλ + ∂J/∂x (r, dx, dy of p are used). */
int pts_adjoint_iterate(pts_ctx* ctx, const pts_pt_params* p, int64_t niter);
/* Iterate until max|Δλ_V| <= eps * max|λ_V| over one iteration, checked every
 * nout iterations (norm_kind and use_graph of sp are ignored).
 * res->comp[0] = max|Δλ_V|, res->comp[1] = max|λ_V|. */
int pts_adjoint_solve(pts_ctx* ctx, const pts_pt_params* p, const pts_solve_params* sp, pts_solve_result* res);
/* ∂J/∂ln μ at the cell centres (μ = the Mus field), host array nx*ny, x
 * fastest, from the current λ and the converged forward stresses. */
int pts_grad_logmu(pts_ctx* ctx, const pts_pt_params* p, double* grad, int64_t count);
/* Visco-elastic time loop (no counterpart in the reference).  After the
 * converged PT solve of physical step n (0-based; assign_τ! at the start of
 * each step, Stokes2D_ve.jl:149-150) the caller stores its stresses with
 * pts_checkpoint(ctx, n).  pts_adjoint_ve_solve then runs the adjoint
 * backward over steps nsteps-1 .. 0 for J(V of the last step) (source set by
 * pts_adjoint_set_source): each step's converged τ is the next step's τ_o, so
 * step n's adjoint is driven by the τ_o-sensitivity of step n+1.  grad
 * (host, nx*ny) receives ∂J/∂ln μ summed over the steps; res sums the
 * iterations of the nsteps adjoint solves. */
int pts_checkpoint(pts_ctx* ctx, int32_t slot);
int pts_adjoint_ve_solve(pts_ctx* ctx, const pts_pt_params* p, int32_t nsteps, const pts_solve_params* sp,
                         pts_solve_result* res, double* grad, int64_t count);

#ifdef __cplusplus
}
#endif
#endif
