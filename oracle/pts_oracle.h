/*
 * pts_oracle.h -- CPU restatement of the reference PT Stokes scripts.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP library is checked
 * against; only tests/, the smoke check and the benchmark's CPU-baseline leg
 * may load it.  The product path never links or calls it.
 *
 * One state object mirrors the local variables of one reference script
 * (scripts/Stokes2D.jl, scripts/Stokes2D_ve.jl, scripts/Stokes2D_ve_perf.jl,
 * scripts/stokes_3D/Stokes3D_ve_perf.jl, scripts/stokes_3D/Stokes3D_ve_multixpu.jl
 * and ..._multixpu_perf.jl).  Arrays are Julia column-major with x fastest and
 * keep the reference shapes exactly.
 */
#ifndef PTS_ORACLE_H
#define PTS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Field ids (same order as the product's pts_field enum). */
enum {
    ORF_PT = 0, ORF_VX, ORF_VY, ORF_VZ,
    ORF_TXX, ORF_TYY, ORF_TZZ, ORF_TXY, ORF_TXZ, ORF_TYZ,
    ORF_TXX_O, ORF_TYY_O, ORF_TZZ_O, ORF_TXY_O, ORF_TXZ_O, ORF_TYZ_O,
    ORF_MUS, ORF_MUS2, ORF_MUSTAU, ORF_GDTAU, ORF_DTAU_RHO,
    ORF_DIVV, ORF_RX, ORF_RY, ORF_RZ, ORF_DVX, ORF_DVY, ORF_DVZ, ORF_RHOG,
    ORF_NFIELDS
};

typedef struct or_params {
    int32_t ndim;      /* 2 or 3 */
    int32_t ve;        /* 1 Maxwell visco-elastic, 0 viscous */
    int32_t perf;      /* 1: fused 2-kernel form multiplying by _dx (the *_perf.jl scripts);
                          0: 4-kernel form dividing by dx (Stokes2D.jl, Stokes2D_ve.jl, multixpu) */
    int32_t ic_igg;    /* 1: ImplicitGlobalGrid x_g() coordinates for the ICs (multixpu scripts) */
    int32_t nx, ny, nz;
    int32_t nt, nout;
    int64_t iter_max;
    int32_t nsmooth;   /* smoothing passes of the viscosity (10 in every script) */
    int32_t norm_inf;  /* 0: reference scaled-RMS err; 1: scaled max-norm (no reference
                          counterpart: parity unpinned beyond its definition) */
    int32_t threads;   /* OpenMP threads for the stencil loops (0 = runtime default) */
    int32_t pad_;
    double lx, ly, lz;
    double mus0, musi; /* matrix / inclusion viscosity */
    double rad2;       /* inclusion test Rad2 < rad2 (1.0 in every script) */
    double G, dt, ebg;
    double Re, r, CFL, eps;
    double rhog_in;    /* body force inside the inclusion (0 = none; Stokes2D_param.jl:59) */
    double rhog_out;   /* body force outside */
} or_params;

typedef struct or_state or_state;

or_state* or_create(const or_params* p);
void      or_destroy(or_state* s);
/* Setup: ICs, smoothing, maxloc, BCs on Musτ, iteration parameters. */
void      or_setup(or_state* s);
/* Copy a field out / in (column-major, exact reference shape). */
int64_t   or_field_len(const or_state* s, int field);
int       or_field_shape(const or_state* s, int field, int64_t shape[3]);
int       or_get(const or_state* s, int field, double* out);
int       or_set(or_state* s, int field, const double* in);
/* One PT iteration of the script's loop body (kernels + BCs), no check. */
void      or_iterate(or_state* s, int64_t n);
/* τ_o ← τ (assign_τ!). */
void      or_assign_tau(or_state* s);
/* Residual kernel used at checks (compute_Res!). */
void      or_compute_res(or_state* s);
/* Norms at a check: out[0..3] = norm_Rx, norm_Ry, norm_Rz, norm_divV (scaled as
 * selected by norm_inf), out[4] = err, out[5..8] = Vmin, Vmax, Pmin, Pmax. */
void      or_check(or_state* s, double out[9]);
/* Full time loop (for it = 1:nt ...).  Returns 0, or -1 on NaN (the reference's
 * error("NaN")).  ittot/err/iters-per-step written to the out pointers. */
int       or_solve(or_state* s, int64_t* ittot, double* err_last, int64_t* iters_per_step);
/* Run one reference kernel in isolation (per-kernel parity checks):
 * 0 compute_P!, 1 compute_τ!, 2 compute_dV!, 3 compute_V!, 4 compute_Res! (divide form);
 * 5 compute_Pt_τ!, 6 compute_V! (perf), 7 compute_Res! (perf);
 * 8 the script's free-slip calls on the velocities;
 * 10 bc_x!, 11 bc_y!, 12 bc_z! on field `arg`. Returns -1 for an unknown id. */
int       or_kernel(or_state* s, int which, int arg);
/* Scalars derived by the script: dx,dy,dz,_dx,_dy,_dz,max_l,Vpdτ. */
void      or_scalars(const or_state* s, double out[8]);

#ifdef __cplusplus
}
#endif
#endif
