/*
 * pts_oracle.c -- CPU restatement of the reference PT Stokes scripts.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).  Only tests/, the smoke check and the
 * benchmark's CPU-baseline leg load this library; the product never does.
 *
 * Every kernel below restates one ParallelStencil kernel of the reference with
 * explicit 0-based indices (reference indices are 1-based: A[ix,iy] here is
 * A[ix-1,iy-1]).  Each keeps the reference's association of + - * / and its
 * choice of dividing by dx (4-kernel scripts) or multiplying by _dx = 1/dx
 * (the *_perf scripts).  Build with -ffp-contract=off: Julia does not fuse
 * a*b+c into an FMA, so neither may we.
 *
 * Parity pinning (see tests/test_oracle_*.py):
 *   - all six reference golden files (test/reftest-files/<name>.bson, decoded under
 *     tests/golden/) are reproduced: 2 converged 2D solves, the 101-iteration
 *     2D VE perf trajectory, the converged 3D VE multixpu solve and the
 *     201-iteration 3D VE perf trajectory;
 *   - the reference's published iteration counts for the converged 2D VE runs
 *     (output/out_Stokes2D_ve.txt) and 3D VE runs (output/out_Stokes3D_ve.txt);
 *   - closed forms: homogeneous viscosity (exact linear V, tau = 2 mu edot and
 *     the backward-Euler Maxwell recursion for VE), x<->y antisymmetry of P in
 *     pure shear, and a direct solve of the steady discrete equations (2D, and
 *     3D viscous).
 *   The body-force term (Stokes2D_param.jl:59) and the max-norm stop criterion
 *   have no reference fixture: the body force is pinned by the direct solve and
 *   the hydrostatic closed form, and both err forms (scaled RMS and max-norm)
 *   by hand-computed values on planted residuals (tests/test_err_norms.py).
 */
#include "pts_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    double* a;
    int64_t n1, n2, n3;
} arr;

#define A2(A, i, j) ((A).a[(int64_t)(i) + (A).n1 * (int64_t)(j)])
#define A3(A, i, j, k) ((A).a[(int64_t)(i) + (A).n1 * ((int64_t)(j) + (A).n2 * (int64_t)(k))])

struct or_state {
    or_params p;
    double dx, dy, dz, _dx, _dy, _dz, max_l, Vpdt;
    arr f[ORF_NFIELDS];
};

static arr mk(int64_t n1, int64_t n2, int64_t n3) {
    arr A;
    if (n1 < 0) n1 = 0;
    if (n2 < 0) n2 = 0;
    if (n3 < 0) n3 = 0;
    A.n1 = n1; A.n2 = n2; A.n3 = n3;
    A.a = (double*)calloc((size_t)(n1 * n2 * n3 > 0 ? n1 * n2 * n3 : 1), sizeof(double));
    return A;
}

static int64_t len(const arr* A) { return A->n1 * A->n2 * A->n3; }

or_state* or_create(const or_params* p) {
    or_state* s = (or_state*)calloc(1, sizeof(or_state));
    s->p = *p;
    const int64_t nx = p->nx, ny = p->ny, nz = p->ndim == 3 ? p->nz : 1;
    if (p->threads > 0) {
#ifdef _OPENMP
        omp_set_num_threads(p->threads);
#endif
    }
    if (p->ndim == 2) {
        /* Stokes2D_ve.jl:111-130 */
        s->f[ORF_PT] = mk(nx, ny, 1);
        s->f[ORF_DTAU_RHO] = mk(nx, ny, 1);
        s->f[ORF_GDTAU] = mk(nx, ny, 1);
        s->f[ORF_DIVV] = mk(nx, ny, 1);
        s->f[ORF_TXX] = mk(nx, ny, 1);
        s->f[ORF_TYY] = mk(nx, ny, 1);
        s->f[ORF_TXY] = mk(nx - 1, ny - 1, 1);
        s->f[ORF_TXX_O] = mk(nx, ny, 1);
        s->f[ORF_TYY_O] = mk(nx, ny, 1);
        s->f[ORF_TXY_O] = mk(nx - 1, ny - 1, 1);
        s->f[ORF_DVX] = mk(nx - 1, ny - 2, 1);
        s->f[ORF_DVY] = mk(nx - 2, ny - 1, 1);
        s->f[ORF_RX] = mk(nx - 1, ny - 2, 1);
        s->f[ORF_RY] = mk(nx - 2, ny - 1, 1);
        s->f[ORF_MUS] = mk(nx, ny, 1);
        s->f[ORF_MUS2] = mk(nx, ny, 1);
        s->f[ORF_MUSTAU] = mk(nx, ny, 1);
        s->f[ORF_RHOG] = mk(nx, ny, 1);
        s->f[ORF_VX] = mk(nx + 1, ny, 1);
        s->f[ORF_VY] = mk(nx, ny + 1, 1);
        s->f[ORF_VZ] = mk(0, 0, 0);
        s->f[ORF_TZZ] = mk(0, 0, 0);
        s->f[ORF_TXZ] = mk(0, 0, 0);
        s->f[ORF_TYZ] = mk(0, 0, 0);
        s->f[ORF_TZZ_O] = mk(0, 0, 0);
        s->f[ORF_TXZ_O] = mk(0, 0, 0);
        s->f[ORF_TYZ_O] = mk(0, 0, 0);
        s->f[ORF_RZ] = mk(0, 0, 0);
        s->f[ORF_DVZ] = mk(0, 0, 0);
    } else {
        /* Stokes3D_ve_perf.jl:141-172, Stokes3D_ve_multixpu.jl:144-173 */
        s->f[ORF_PT] = mk(nx, ny, nz);
        s->f[ORF_DTAU_RHO] = mk(nx, ny, nz);
        s->f[ORF_GDTAU] = mk(nx, ny, nz);
        s->f[ORF_DIVV] = mk(nx, ny, nz);
        s->f[ORF_VX] = mk(nx + 1, ny, nz);
        s->f[ORF_VY] = mk(nx, ny + 1, nz);
        s->f[ORF_VZ] = mk(nx, ny, nz + 1);
        s->f[ORF_TXX] = mk(nx, ny - 2, nz - 2);
        s->f[ORF_TYY] = mk(nx - 2, ny, nz - 2);
        s->f[ORF_TZZ] = mk(nx - 2, ny - 2, nz);
        s->f[ORF_TXY] = mk(nx - 1, ny - 1, nz - 2);
        s->f[ORF_TXZ] = mk(nx - 1, ny - 2, nz - 1);
        s->f[ORF_TYZ] = mk(nx - 2, ny - 1, nz - 1);
        s->f[ORF_TXX_O] = mk(nx, ny - 2, nz - 2);
        s->f[ORF_TYY_O] = mk(nx - 2, ny, nz - 2);
        s->f[ORF_TZZ_O] = mk(nx - 2, ny - 2, nz);
        s->f[ORF_TXY_O] = mk(nx - 1, ny - 1, nz - 2);
        s->f[ORF_TXZ_O] = mk(nx - 1, ny - 2, nz - 1);
        s->f[ORF_TYZ_O] = mk(nx - 2, ny - 1, nz - 1);
        s->f[ORF_DVX] = mk(nx - 1, ny - 2, nz - 2);
        s->f[ORF_DVY] = mk(nx - 2, ny - 1, nz - 2);
        s->f[ORF_DVZ] = mk(nx - 2, ny - 2, nz - 1);
        s->f[ORF_RX] = mk(nx - 1, ny - 2, nz - 2);
        s->f[ORF_RY] = mk(nx - 2, ny - 1, nz - 2);
        s->f[ORF_RZ] = mk(nx - 2, ny - 2, nz - 1);
        s->f[ORF_MUS] = mk(nx, ny, nz);
        s->f[ORF_MUS2] = mk(nx, ny, nz);
        s->f[ORF_MUSTAU] = mk(nx, ny, nz);
        s->f[ORF_RHOG] = mk(nx, ny, nz);
    }
    /* Derived numerics: Stokes2D_ve.jl:106-108, Stokes3D_ve_perf.jl:135-138 */
    s->dx = p->lx / p->nx;
    s->dy = p->ly / p->ny;
    s->dz = p->ndim == 3 ? p->lz / p->nz : 1.0;
    s->_dx = 1.0 / s->dx;
    s->_dy = 1.0 / s->dy;
    s->_dz = 1.0 / s->dz;
    if (p->ndim == 2) {
        s->max_l = fmax(p->lx, p->ly);
        s->Vpdt = fmin(s->dx, s->dy) * p->CFL;
    } else {
        s->max_l = fmax(fmax(p->lx, p->ly), p->lz);
        s->Vpdt = fmin(fmin(s->dx, s->dy), s->dz) * p->CFL;
    }
    return s;
}

void or_destroy(or_state* s) {
    if (!s) return;
    for (int i = 0; i < ORF_NFIELDS; ++i) free(s->f[i].a);
    free(s);
}

void or_scalars(const or_state* s, double out[8]) {
    out[0] = s->dx; out[1] = s->dy; out[2] = s->dz;
    out[3] = s->_dx; out[4] = s->_dy; out[5] = s->_dz;
    out[6] = s->max_l; out[7] = s->Vpdt;
}

int64_t or_field_len(const or_state* s, int field) {
    if (field < 0 || field >= ORF_NFIELDS) return -1;
    return len(&s->f[field]);
}

int or_field_shape(const or_state* s, int field, int64_t shape[3]) {
    if (field < 0 || field >= ORF_NFIELDS) return -1;
    shape[0] = s->f[field].n1; shape[1] = s->f[field].n2; shape[2] = s->f[field].n3;
    return 0;
}

int or_get(const or_state* s, int field, double* out) {
    if (field < 0 || field >= ORF_NFIELDS) return -1;
    memcpy(out, s->f[field].a, sizeof(double) * (size_t)len(&s->f[field]));
    return 0;
}

int or_set(or_state* s, int field, const double* in) {
    if (field < 0 || field >= ORF_NFIELDS) return -1;
    memcpy(s->f[field].a, in, sizeof(double) * (size_t)len(&s->f[field]));
    return 0;
}

static void swap_fields(or_state* s, int a, int b) {
    arr t = s->f[a];
    s->f[a] = s->f[b];
    s->f[b] = t;
}

/* ------------------------------------------------------------------------ */
/* Boundary kernels: bc_x!/bc_y!/bc_z! (Stokes2D_ve.jl:75-85,             */
/* Stokes3D_ve_perf.jl:98-114).  Copy the first interior plane outward.     */
/* ------------------------------------------------------------------------ */
static void bc_x(arr A) {
    if (A.n1 < 2) return;
    for (int64_t k = 0; k < A.n3; ++k)
        for (int64_t j = 0; j < A.n2; ++j) {
            A3(A, 0, j, k) = A3(A, 1, j, k);
            A3(A, A.n1 - 1, j, k) = A3(A, A.n1 - 2, j, k);
        }
}
static void bc_y(arr A) {
    if (A.n2 < 2) return;
    for (int64_t k = 0; k < A.n3; ++k)
        for (int64_t i = 0; i < A.n1; ++i) {
            A3(A, i, 0, k) = A3(A, i, 1, k);
            A3(A, i, A.n2 - 1, k) = A3(A, i, A.n2 - 2, k);
        }
}
static void bc_z(arr A) {
    if (A.n3 < 2) return;
    for (int64_t j = 0; j < A.n2; ++j)
        for (int64_t i = 0; i < A.n1; ++i) {
            A3(A, i, j, 0) = A3(A, i, j, 1);
            A3(A, i, j, A.n3 - 1) = A3(A, i, j, A.n3 - 2);
        }
}

/* ------------------------------------------------------------------------ */
/* Setup kernels                                                              */
/* ------------------------------------------------------------------------ */
/* smooth!: Stokes2D_ve.jl:18-21 (2D), Stokes3D_ve_perf.jl:24-27 (3D).
 * @inn(A2) = @inn(A) + 1.0/4.1/fact*(@d2_xi(A) + @d2_yi(A)). */
static void smooth2(arr A2, arr A, double fact) {
    const double c = 1.0 / 4.1 / fact;
#pragma omp parallel for
    for (int64_t j = 0; j < A.n2 - 2; ++j)
        for (int64_t i = 0; i < A.n1 - 2; ++i) {
            double d2x = (A2(A, i + 2, j + 1) - A2(A, i + 1, j + 1)) - (A2(A, i + 1, j + 1) - A2(A, i, j + 1));
            double d2y = (A2(A, i + 1, j + 2) - A2(A, i + 1, j + 1)) - (A2(A, i + 1, j + 1) - A2(A, i + 1, j));
            A2(A2, i + 1, j + 1) = A2(A, i + 1, j + 1) + c * (d2x + d2y);
        }
}
static void smooth3(arr A2, arr A, double fact) {
    const double c = 1.0 / 6.1 / fact;
#pragma omp parallel for
    for (int64_t k = 0; k < A.n3 - 2; ++k)
        for (int64_t j = 0; j < A.n2 - 2; ++j)
            for (int64_t i = 0; i < A.n1 - 2; ++i) {
                double m = A3(A, i + 1, j + 1, k + 1);
                double d2x = (A3(A, i + 2, j + 1, k + 1) - m) - (m - A3(A, i, j + 1, k + 1));
                double d2y = (A3(A, i + 1, j + 2, k + 1) - m) - (m - A3(A, i + 1, j, k + 1));
                double d2z = (A3(A, i + 1, j + 1, k + 2) - m) - (m - A3(A, i + 1, j + 1, k));
                A3(A2, i + 1, j + 1, k + 1) = m + c * (d2x + d2y + d2z);
            }
}
/* compute_maxloc!: @inn(Musτ2) = @maxloc(Musτ) -- max over the 5-point (2D) or
 * 7-point (3D) cross (Stokes2D_ve.jl:23-26, Stokes3D_ve_perf.jl:29-32). */
static void maxloc2(arr B, arr A) {
#pragma omp parallel for
    for (int64_t j = 0; j < A.n2 - 2; ++j)
        for (int64_t i = 0; i < A.n1 - 2; ++i) {
            double m = fmax(fmax(fmax(fmax(A2(A, i, j + 1), A2(A, i + 2, j + 1)), A2(A, i + 1, j + 1)),
                                 A2(A, i + 1, j)), A2(A, i + 1, j + 2));
            A2(B, i + 1, j + 1) = m;
        }
}
static void maxloc3(arr B, arr A) {
#pragma omp parallel for
    for (int64_t k = 0; k < A.n3 - 2; ++k)
        for (int64_t j = 0; j < A.n2 - 2; ++j)
            for (int64_t i = 0; i < A.n1 - 2; ++i) {
                double m = fmax(A3(A, i, j + 1, k + 1), A3(A, i + 2, j + 1, k + 1));
                m = fmax(m, A3(A, i + 1, j + 1, k + 1));
                m = fmax(m, A3(A, i + 1, j, k + 1));
                m = fmax(m, A3(A, i + 1, j + 2, k + 1));
                m = fmax(m, A3(A, i + 1, j + 1, k));
                m = fmax(m, A3(A, i + 1, j + 1, k + 2));
                A3(B, i + 1, j + 1, k + 1) = m;
            }
}
/* compute_iter_params!: Stokes2D.jl:27-31 (viscous), Stokes2D_ve.jl:28-33 and
 * Stokes3D_ve_perf.jl:34-39 (VE: Mu_eff = 1/(1/Musτ + 1/(G*dt))). */
static void iter_params(or_state* s) {
    arr R = s->f[ORF_DTAU_RHO], Gd = s->f[ORF_GDTAU], M = s->f[ORF_MUSTAU];
    const double Vp = s->Vpdt, ml = s->max_l, Re = s->p.Re, r = s->p.r, G = s->p.G, dt = s->p.dt;
    const int64_t n = len(&M);
    for (int64_t q = 0; q < n; ++q) {
        double mu_eff = s->p.ve ? 1.0 / (1.0 / M.a[q] + 1.0 / (G * dt)) : M.a[q];
        R.a[q] = Vp * ml / Re / mu_eff;
        Gd.a[q] = Vp * Vp / R.a[q] / (r + 2.0);
    }
}

/* ICs + setup: Stokes2D_ve.jl:131-147, Stokes2D.jl:109-125,
 * Stokes3D_ve_perf.jl:167-182, Stokes3D_ve_multixpu.jl:174-195. */
void or_setup(or_state* s) {
    const or_params* p = &s->p;
    const double dx = s->dx, dy = s->dy, dz = s->dz, lx = p->lx, ly = p->ly, lz = p->lz, ebg = p->ebg;
    arr Mus = s->f[ORF_MUS], Vx = s->f[ORF_VX], Vy = s->f[ORF_VY], Vz = s->f[ORF_VZ], rg = s->f[ORF_RHOG];
    if (p->ndim == 2) {
        for (int64_t j = 0; j < Mus.n2; ++j)
            for (int64_t i = 0; i < Mus.n1; ++i) {
                double xr = ((double)i * dx + 0.5 * dx) - 0.5 * lx;
                double yr = ((double)j * dy + 0.5 * dy) - 0.5 * ly;
                double rad2 = xr * xr + yr * yr;
                int in = rad2 < p->rad2;
                A2(Mus, i, j) = in ? p->musi : p->mus0 * 1.0;
                A2(rg, i, j) = in ? p->rhog_in : p->rhog_out;
            }
        for (int64_t j = 0; j < Vx.n2; ++j)
            for (int64_t i = 0; i < Vx.n1; ++i) A2(Vx, i, j) = (-ebg) * ((double)i * dx - 0.5 * lx);
        for (int64_t j = 0; j < Vy.n2; ++j)
            for (int64_t i = 0; i < Vy.n1; ++i) A2(Vy, i, j) = ebg * ((double)j * dy - 0.5 * ly);
    } else {
        for (int64_t k = 0; k < Mus.n3; ++k)
            for (int64_t j = 0; j < Mus.n2; ++j)
                for (int64_t i = 0; i < Mus.n1; ++i) {
                    /* x_g(ix,dx,A) == (ix-1)*dx on one rank for an nx-sized array */
                    double xr = ((double)i * dx + 0.5 * dx) - 0.5 * lx;
                    double yr = ((double)j * dy + 0.5 * dy) - 0.5 * ly;
                    double zr = ((double)k * dz + 0.5 * dz) - 0.5 * lz;
                    double rad2 = xr * xr + yr * yr + zr * zr;
                    int in = rad2 < p->rad2;
                    A3(Mus, i, j, k) = in ? p->musi : p->mus0 * 1.0;
                    A3(rg, i, j, k) = in ? p->rhog_in : p->rhog_out;
                }
        for (int64_t k = 0; k < Vx.n3; ++k)
            for (int64_t j = 0; j < Vx.n2; ++j)
                for (int64_t i = 0; i < Vx.n1; ++i) {
                    double x;
                    if (p->ic_igg) /* x_g(ix,dx,Vx) = (ix-1)*dx - dx/2 for an (nx+1)-sized array */
                        x = (((double)i * dx + (-0.5 * dx)) + 0.5 * dx) - 0.5 * lx;
                    else
                        x = (double)i * dx - 0.5 * lx;
                    A3(Vx, i, j, k) = (-ebg) * x;
                }
        for (int64_t k = 0; k < Vz.n3; ++k)
            for (int64_t j = 0; j < Vz.n2; ++j)
                for (int64_t i = 0; i < Vz.n1; ++i) {
                    double z;
                    if (p->ic_igg)
                        z = (((double)k * dz + (-0.5 * dz)) + 0.5 * dz) - 0.5 * lz;
                    else
                        z = (double)k * dz - 0.5 * lz;
                    A3(Vz, i, j, k) = ebg * z;
                }
        /* Vy stays zero (x-z pure shear) */
    }
    /* Mus2 .= Mus; 10x smooth! with swap */
    memcpy(s->f[ORF_MUS2].a, Mus.a, sizeof(double) * (size_t)len(&Mus));
    for (int ism = 0; ism < p->nsmooth; ++ism) {
        if (p->ndim == 2) smooth2(s->f[ORF_MUS2], s->f[ORF_MUS], 1.0);
        else smooth3(s->f[ORF_MUS2], s->f[ORF_MUS], 1.0);
        swap_fields(s, ORF_MUS, ORF_MUS2);
    }
    /* Musτ .= Mus; compute_maxloc!(Musτ, Mus); bc_x!, bc_y! (, bc_z!) */
    memcpy(s->f[ORF_MUSTAU].a, s->f[ORF_MUS].a, sizeof(double) * (size_t)len(&Mus));
    if (p->ndim == 2) maxloc2(s->f[ORF_MUSTAU], s->f[ORF_MUS]);
    else maxloc3(s->f[ORF_MUSTAU], s->f[ORF_MUS]);
    bc_x(s->f[ORF_MUSTAU]);
    bc_y(s->f[ORF_MUSTAU]);
    if (p->ndim == 3) bc_z(s->f[ORF_MUSTAU]);
    iter_params(s);
}

void or_assign_tau(or_state* s) {
    /* assign_τ!: Stokes2D_ve.jl:35-40, Stokes3D_ve_perf.jl:41-49 */
    const int src[6] = {ORF_TXX, ORF_TYY, ORF_TZZ, ORF_TXY, ORF_TXZ, ORF_TYZ};
    const int dst[6] = {ORF_TXX_O, ORF_TYY_O, ORF_TZZ_O, ORF_TXY_O, ORF_TXZ_O, ORF_TYZ_O};
    for (int q = 0; q < 6; ++q)
        memcpy(s->f[dst[q]].a, s->f[src[q]].a, sizeof(double) * (size_t)len(&s->f[src[q]]));
}

/* ------------------------------------------------------------------------ */
/* 2D, 4-kernel form (divide by dx)                                           */
/* ------------------------------------------------------------------------ */
/* compute_P!: Stokes2D.jl:33-37, Stokes2D_ve.jl:42-46 */
static void p2_compute_P(or_state* s) {
    arr dV = s->f[ORF_DIVV], Pt = s->f[ORF_PT], Vx = s->f[ORF_VX], Vy = s->f[ORF_VY], Gd = s->f[ORF_GDTAU];
    const double dx = s->dx, dy = s->dy, r = s->p.r;
#pragma omp parallel for
    for (int64_t j = 0; j < Pt.n2; ++j)
        for (int64_t i = 0; i < Pt.n1; ++i) {
            double d = (A2(Vx, i + 1, j) - A2(Vx, i, j)) / dx + (A2(Vy, i, j + 1) - A2(Vy, i, j)) / dy;
            A2(dV, i, j) = d;
            A2(Pt, i, j) = A2(Pt, i, j) - r * A2(Gd, i, j) * d;
        }
}
/* compute_τ!: viscous Stokes2D.jl:39-44; VE Stokes2D_ve.jl:48-55 */
static void p2_compute_tau(or_state* s) {
    arr txx = s->f[ORF_TXX], tyy = s->f[ORF_TYY], txy = s->f[ORF_TXY];
    arr txxo = s->f[ORF_TXX_O], tyyo = s->f[ORF_TYY_O], txyo = s->f[ORF_TXY_O];
    arr Gd = s->f[ORF_GDTAU], Mu = s->f[ORF_MUS], Vx = s->f[ORF_VX], Vy = s->f[ORF_VY];
    const double dx = s->dx, dy = s->dy, Gdt = s->p.G * s->p.dt;
    const int ve = s->p.ve;
#pragma omp parallel for
    for (int64_t j = 0; j < txx.n2; ++j)
        for (int64_t i = 0; i < txx.n1; ++i) {
            double g = A2(Gd, i, j), m = A2(Mu, i, j);
            double exx = A2(Vx, i + 1, j) - A2(Vx, i, j);
            double eyy = A2(Vy, i, j + 1) - A2(Vy, i, j);
            if (ve) {
                double Gr = g / Gdt;
                A2(txx, i, j) = (A2(txx, i, j) + A2(txxo, i, j) * Gr + 2.0 * g * (exx / dx)) / (1.0 + g / m + Gr);
                A2(tyy, i, j) = (A2(tyy, i, j) + A2(tyyo, i, j) * Gr + 2.0 * g * (eyy / dy)) / (1.0 + g / m + Gr);
            } else {
                A2(txx, i, j) = (A2(txx, i, j) + 2.0 * g * exx / dx) / (g / m + 1.0);
                A2(tyy, i, j) = (A2(tyy, i, j) + 2.0 * g * eyy / dy) / (g / m + 1.0);
            }
        }
#pragma omp parallel for
    for (int64_t j = 0; j < txy.n2; ++j)
        for (int64_t i = 0; i < txy.n1; ++i) {
            double ag = (A2(Gd, i, j) + A2(Gd, i + 1, j) + A2(Gd, i, j + 1) + A2(Gd, i + 1, j + 1)) * 0.25;
            double am = (A2(Mu, i, j) + A2(Mu, i + 1, j) + A2(Mu, i, j + 1) + A2(Mu, i + 1, j + 1)) * 0.25;
            double dvx = A2(Vx, i + 1, j + 1) - A2(Vx, i + 1, j);
            double dvy = A2(Vy, i + 1, j + 1) - A2(Vy, i, j + 1);
            if (ve) {
                double aGr = ag / Gdt;
                A2(txy, i, j) = (A2(txy, i, j) + A2(txyo, i, j) * aGr + 2.0 * ag * (0.5 * (dvx / dy + dvy / dx))) /
                                (1.0 + ag / am + aGr);
            } else {
                A2(txy, i, j) = (A2(txy, i, j) + 2.0 * ag * (0.5 * (dvx / dy + dvy / dx))) / (ag / am + 1.0);
            }
        }
}
/* compute_dV!: Stokes2D.jl:46-52 (stores Rx, Ry too), Stokes2D_ve.jl:57-61;
 * body force as Stokes2D_param.jl:59 (Ry -= @av_yi(ρg)). */
static void p2_compute_dV(or_state* s, int store_R) {
    arr dVx = s->f[ORF_DVX], dVy = s->f[ORF_DVY], Rx = s->f[ORF_RX], Ry = s->f[ORF_RY];
    arr Pt = s->f[ORF_PT], txx = s->f[ORF_TXX], tyy = s->f[ORF_TYY], txy = s->f[ORF_TXY], Rh = s->f[ORF_DTAU_RHO];
    arr rg = s->f[ORF_RHOG];
    const double dx = s->dx, dy = s->dy;
    const int bf = s->p.rhog_in != 0.0 || s->p.rhog_out != 0.0;
#pragma omp parallel for
    for (int64_t j = 0; j < dVx.n2; ++j)
        for (int64_t i = 0; i < dVx.n1; ++i) {
            double R = (A2(txx, i + 1, j + 1) - A2(txx, i, j + 1)) / dx + (A2(txy, i, j + 1) - A2(txy, i, j)) / dy -
                       (A2(Pt, i + 1, j + 1) - A2(Pt, i, j + 1)) / dx;
            if (store_R) A2(Rx, i, j) = R;
            A2(dVx, i, j) = R * ((A2(Rh, i, j + 1) + A2(Rh, i + 1, j + 1)) * 0.5);
        }
#pragma omp parallel for
    for (int64_t j = 0; j < dVy.n2; ++j)
        for (int64_t i = 0; i < dVy.n1; ++i) {
            double R = (A2(tyy, i + 1, j + 1) - A2(tyy, i + 1, j)) / dy + (A2(txy, i + 1, j) - A2(txy, i, j)) / dx -
                       (A2(Pt, i + 1, j + 1) - A2(Pt, i + 1, j)) / dy;
            if (bf) R = R - (A2(rg, i + 1, j) + A2(rg, i + 1, j + 1)) * 0.5;
            if (store_R) A2(Ry, i, j) = R;
            A2(dVy, i, j) = R * ((A2(Rh, i + 1, j) + A2(Rh, i + 1, j + 1)) * 0.5);
        }
}
/* compute_V!: Stokes2D.jl:54-58 */
static void p2_compute_V(or_state* s) {
    arr Vx = s->f[ORF_VX], Vy = s->f[ORF_VY], dVx = s->f[ORF_DVX], dVy = s->f[ORF_DVY];
    for (int64_t j = 0; j < dVx.n2; ++j)
        for (int64_t i = 0; i < dVx.n1; ++i) A2(Vx, i + 1, j + 1) = A2(Vx, i + 1, j + 1) + A2(dVx, i, j);
    for (int64_t j = 0; j < dVy.n2; ++j)
        for (int64_t i = 0; i < dVy.n1; ++i) A2(Vy, i + 1, j + 1) = A2(Vy, i + 1, j + 1) + A2(dVy, i, j);
}
/* compute_Res!: Stokes2D_ve.jl:69-73 */
static void p2_compute_Res(or_state* s) {
    arr Rx = s->f[ORF_RX], Ry = s->f[ORF_RY], Pt = s->f[ORF_PT], txx = s->f[ORF_TXX], tyy = s->f[ORF_TYY],
        txy = s->f[ORF_TXY], rg = s->f[ORF_RHOG];
    const double dx = s->dx, dy = s->dy;
    const int bf = s->p.rhog_in != 0.0 || s->p.rhog_out != 0.0;
    for (int64_t j = 0; j < Rx.n2; ++j)
        for (int64_t i = 0; i < Rx.n1; ++i)
            A2(Rx, i, j) = (A2(txx, i + 1, j + 1) - A2(txx, i, j + 1)) / dx + (A2(txy, i, j + 1) - A2(txy, i, j)) / dy -
                           (A2(Pt, i + 1, j + 1) - A2(Pt, i, j + 1)) / dx;
    for (int64_t j = 0; j < Ry.n2; ++j)
        for (int64_t i = 0; i < Ry.n1; ++i) {
            double R = (A2(tyy, i + 1, j + 1) - A2(tyy, i + 1, j)) / dy + (A2(txy, i + 1, j) - A2(txy, i, j)) / dx -
                       (A2(Pt, i + 1, j + 1) - A2(Pt, i + 1, j)) / dy;
            if (bf) R = R - (A2(rg, i + 1, j) + A2(rg, i + 1, j + 1)) * 0.5;
            A2(Ry, i, j) = R;
        }
}

/* ------------------------------------------------------------------------ */
/* 2D, fused perf form (multiply by _dx): Stokes2D_ve_perf.jl:43-66          */
/* ------------------------------------------------------------------------ */
static void f2_compute_Pt_tau(or_state* s) {
    arr Pt = s->f[ORF_PT], txx = s->f[ORF_TXX], tyy = s->f[ORF_TYY], txy = s->f[ORF_TXY];
    arr txxo = s->f[ORF_TXX_O], tyyo = s->f[ORF_TYY_O], txyo = s->f[ORF_TXY_O];
    arr Gd = s->f[ORF_GDTAU], Mu = s->f[ORF_MUS], Vx = s->f[ORF_VX], Vy = s->f[ORF_VY];
    const double _dx = s->_dx, _dy = s->_dy, r = s->p.r, Gdt = s->p.G * s->p.dt;
    const int ve = s->p.ve;
    const int64_t nx = Pt.n1, ny = Pt.n2;
#pragma omp parallel for
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
            double g = A2(Gd, i, j), m = A2(Mu, i, j);
            double exx = _dx * (A2(Vx, i + 1, j) - A2(Vx, i, j));
            double eyy = _dy * (A2(Vy, i, j + 1) - A2(Vy, i, j));
            A2(Pt, i, j) = A2(Pt, i, j) - r * g * (exx + eyy);
            if (ve) {
                double Gr = g / Gdt;
                A2(txx, i, j) = (A2(txx, i, j) + A2(txxo, i, j) * Gr + 2.0 * g * exx) / (1.0 + g / m + Gr);
                A2(tyy, i, j) = (A2(tyy, i, j) + A2(tyyo, i, j) * Gr + 2.0 * g * eyy) / (1.0 + g / m + Gr);
            } else {
                A2(txx, i, j) = (A2(txx, i, j) + 2.0 * g * exx) / (1.0 + g / m);
                A2(tyy, i, j) = (A2(tyy, i, j) + 2.0 * g * eyy) / (1.0 + g / m);
            }
            if (i < nx - 1 && j < ny - 1) {
                double ag = 0.25 * (A2(Gd, i, j) + A2(Gd, i + 1, j) + A2(Gd, i, j + 1) + A2(Gd, i + 1, j + 1));
                double am = 0.25 * (A2(Mu, i, j) + A2(Mu, i + 1, j) + A2(Mu, i, j + 1) + A2(Mu, i + 1, j + 1));
                double sh = _dy * (A2(Vx, i + 1, j + 1) - A2(Vx, i + 1, j)) + _dx * (A2(Vy, i + 1, j + 1) - A2(Vy, i, j + 1));
                if (ve) {
                    double aGr = ag / Gdt;
                    A2(txy, i, j) = (A2(txy, i, j) + A2(txyo, i, j) * aGr + 2.0 * ag * 0.5 * sh) / (1.0 + ag / am + aGr);
                } else {
                    A2(txy, i, j) = (A2(txy, i, j) + 2.0 * ag * 0.5 * sh) / (1.0 + ag / am);
                }
            }
        }
}
static void f2_compute_V(or_state* s) {
    arr Vx = s->f[ORF_VX], Vy = s->f[ORF_VY], Pt = s->f[ORF_PT], txx = s->f[ORF_TXX], tyy = s->f[ORF_TYY],
        txy = s->f[ORF_TXY], Rh = s->f[ORF_DTAU_RHO], rg = s->f[ORF_RHOG];
    const double _dx = s->_dx, _dy = s->_dy;
    const int64_t nx = Pt.n1, ny = Pt.n2;
    const int bf = s->p.rhog_in != 0.0 || s->p.rhog_out != 0.0;
#pragma omp parallel for
    for (int64_t j = 0; j < ny - 1; ++j)
        for (int64_t i = 0; i < nx - 1; ++i) {
            if (j < ny - 2) {
                double R = _dx * (A2(txx, i + 1, j + 1) - A2(txx, i, j + 1)) + _dy * (A2(txy, i, j + 1) - A2(txy, i, j)) -
                           _dx * (A2(Pt, i + 1, j + 1) - A2(Pt, i, j + 1));
                A2(Vx, i + 1, j + 1) = A2(Vx, i + 1, j + 1) + R * (0.5 * (A2(Rh, i, j + 1) + A2(Rh, i + 1, j + 1)));
            }
            if (i < nx - 2) {
                double R = _dy * (A2(tyy, i + 1, j + 1) - A2(tyy, i + 1, j)) + _dx * (A2(txy, i + 1, j) - A2(txy, i, j)) -
                           _dy * (A2(Pt, i + 1, j + 1) - A2(Pt, i + 1, j));
                if (bf) R = R - 0.5 * (A2(rg, i + 1, j) + A2(rg, i + 1, j + 1));
                A2(Vy, i + 1, j + 1) = A2(Vy, i + 1, j + 1) + R * (0.5 * (A2(Rh, i + 1, j) + A2(Rh, i + 1, j + 1)));
            }
        }
}
static void f2_compute_Res(or_state* s) {
    arr dV = s->f[ORF_DIVV], Rx = s->f[ORF_RX], Ry = s->f[ORF_RY], Pt = s->f[ORF_PT], Vx = s->f[ORF_VX],
        Vy = s->f[ORF_VY], txx = s->f[ORF_TXX], tyy = s->f[ORF_TYY], txy = s->f[ORF_TXY], rg = s->f[ORF_RHOG];
    const double _dx = s->_dx, _dy = s->_dy;
    const int bf = s->p.rhog_in != 0.0 || s->p.rhog_out != 0.0;
    for (int64_t j = 0; j < dV.n2; ++j)
        for (int64_t i = 0; i < dV.n1; ++i)
            A2(dV, i, j) = _dx * (A2(Vx, i + 1, j) - A2(Vx, i, j)) + _dy * (A2(Vy, i, j + 1) - A2(Vy, i, j));
    for (int64_t j = 0; j < Rx.n2; ++j)
        for (int64_t i = 0; i < Rx.n1; ++i)
            A2(Rx, i, j) = _dx * (A2(txx, i + 1, j + 1) - A2(txx, i, j + 1)) + _dy * (A2(txy, i, j + 1) - A2(txy, i, j)) -
                           _dx * (A2(Pt, i + 1, j + 1) - A2(Pt, i, j + 1));
    for (int64_t j = 0; j < Ry.n2; ++j)
        for (int64_t i = 0; i < Ry.n1; ++i) {
            double R = _dy * (A2(tyy, i + 1, j + 1) - A2(tyy, i + 1, j)) + _dx * (A2(txy, i + 1, j) - A2(txy, i, j)) -
                       _dy * (A2(Pt, i + 1, j + 1) - A2(Pt, i + 1, j));
            if (bf) R = R - 0.5 * (A2(rg, i + 1, j) + A2(rg, i + 1, j + 1));
            A2(Ry, i, j) = R;
        }
}

/* ------------------------------------------------------------------------ */
/* 3D, 4-kernel form (divide by dx): Stokes3D_ve_multixpu.jl:57-98          */
/* ------------------------------------------------------------------------ */
static void p3_compute_P(or_state* s) {
    arr dV = s->f[ORF_DIVV], Pt = s->f[ORF_PT], Vx = s->f[ORF_VX], Vy = s->f[ORF_VY], Vz = s->f[ORF_VZ],
        Gd = s->f[ORF_GDTAU];
    const double dx = s->dx, dy = s->dy, dz = s->dz, r = s->p.r;
#pragma omp parallel for
    for (int64_t k = 0; k < Pt.n3; ++k)
        for (int64_t j = 0; j < Pt.n2; ++j)
            for (int64_t i = 0; i < Pt.n1; ++i) {
                double d = (A3(Vx, i + 1, j, k) - A3(Vx, i, j, k)) / dx + (A3(Vy, i, j + 1, k) - A3(Vy, i, j, k)) / dy +
                           (A3(Vz, i, j, k + 1) - A3(Vz, i, j, k)) / dz;
                A3(dV, i, j, k) = d;
                A3(Pt, i, j, k) = A3(Pt, i, j, k) - r * A3(Gd, i, j, k) * d;
            }
}

/* Normal-stress update shared by the 3D forms.  e is the already-scaled
 * strain rate (d/dx or _dx*d). */
static inline double tau_n(int ve, double t, double to, double g, double m, double Gdt, double e) {
    if (ve) {
        double Gr = g / Gdt;
        return (t + to * Gr + 2.0 * g * e) / (1.0 + g / m + Gr);
    }
    return (t + 2.0 * g * e) / (1.0 + g / m);
}

static void p3_compute_tau(or_state* s) {
    const double dx = s->dx, dy = s->dy, dz = s->dz, Gdt = s->p.G * s->p.dt;
    const int ve = s->p.ve;
    arr Gd = s->f[ORF_GDTAU], Mu = s->f[ORF_MUS], Vx = s->f[ORF_VX], Vy = s->f[ORF_VY], Vz = s->f[ORF_VZ];
    arr txx = s->f[ORF_TXX], tyy = s->f[ORF_TYY], tzz = s->f[ORF_TZZ], txy = s->f[ORF_TXY], txz = s->f[ORF_TXZ],
        tyz = s->f[ORF_TYZ];
    arr txxo = s->f[ORF_TXX_O], tyyo = s->f[ORF_TYY_O], tzzo = s->f[ORF_TZZ_O], txyo = s->f[ORF_TXY_O],
        txzo = s->f[ORF_TXZ_O], tyzo = s->f[ORF_TYZ_O];
#pragma omp parallel for
    for (int64_t k = 0; k < txx.n3; ++k)
        for (int64_t j = 0; j < txx.n2; ++j)
            for (int64_t i = 0; i < txx.n1; ++i)
                A3(txx, i, j, k) = tau_n(ve, A3(txx, i, j, k), A3(txxo, i, j, k), A3(Gd, i, j + 1, k + 1),
                                         A3(Mu, i, j + 1, k + 1), Gdt,
                                         (A3(Vx, i + 1, j + 1, k + 1) - A3(Vx, i, j + 1, k + 1)) / dx);
#pragma omp parallel for
    for (int64_t k = 0; k < tyy.n3; ++k)
        for (int64_t j = 0; j < tyy.n2; ++j)
            for (int64_t i = 0; i < tyy.n1; ++i)
                A3(tyy, i, j, k) = tau_n(ve, A3(tyy, i, j, k), A3(tyyo, i, j, k), A3(Gd, i + 1, j, k + 1),
                                         A3(Mu, i + 1, j, k + 1), Gdt,
                                         (A3(Vy, i + 1, j + 1, k + 1) - A3(Vy, i + 1, j, k + 1)) / dy);
#pragma omp parallel for
    for (int64_t k = 0; k < tzz.n3; ++k)
        for (int64_t j = 0; j < tzz.n2; ++j)
            for (int64_t i = 0; i < tzz.n1; ++i)
                A3(tzz, i, j, k) = tau_n(ve, A3(tzz, i, j, k), A3(tzzo, i, j, k), A3(Gd, i + 1, j + 1, k),
                                         A3(Mu, i + 1, j + 1, k), Gdt,
                                         (A3(Vz, i + 1, j + 1, k + 1) - A3(Vz, i + 1, j + 1, k)) / dz);
#pragma omp parallel for
    for (int64_t k = 0; k < txy.n3; ++k)
        for (int64_t j = 0; j < txy.n2; ++j)
            for (int64_t i = 0; i < txy.n1; ++i) {
                double ag = (A3(Gd, i, j, k + 1) + A3(Gd, i + 1, j, k + 1) + A3(Gd, i, j + 1, k + 1) + A3(Gd, i + 1, j + 1, k + 1)) * 0.25;
                double am = (A3(Mu, i, j, k + 1) + A3(Mu, i + 1, j, k + 1) + A3(Mu, i, j + 1, k + 1) + A3(Mu, i + 1, j + 1, k + 1)) * 0.25;
                double e = 0.5 * ((A3(Vx, i + 1, j + 1, k + 1) - A3(Vx, i + 1, j, k + 1)) / dy +
                                  (A3(Vy, i + 1, j + 1, k + 1) - A3(Vy, i, j + 1, k + 1)) / dx);
                A3(txy, i, j, k) = tau_n(ve, A3(txy, i, j, k), A3(txyo, i, j, k), ag, am, Gdt, e);
            }
#pragma omp parallel for
    for (int64_t k = 0; k < txz.n3; ++k)
        for (int64_t j = 0; j < txz.n2; ++j)
            for (int64_t i = 0; i < txz.n1; ++i) {
                double ag = (A3(Gd, i, j + 1, k) + A3(Gd, i + 1, j + 1, k) + A3(Gd, i, j + 1, k + 1) + A3(Gd, i + 1, j + 1, k + 1)) * 0.25;
                double am = (A3(Mu, i, j + 1, k) + A3(Mu, i + 1, j + 1, k) + A3(Mu, i, j + 1, k + 1) + A3(Mu, i + 1, j + 1, k + 1)) * 0.25;
                double e = 0.5 * ((A3(Vx, i + 1, j + 1, k + 1) - A3(Vx, i + 1, j + 1, k)) / dz +
                                  (A3(Vz, i + 1, j + 1, k + 1) - A3(Vz, i, j + 1, k + 1)) / dx);
                A3(txz, i, j, k) = tau_n(ve, A3(txz, i, j, k), A3(txzo, i, j, k), ag, am, Gdt, e);
            }
#pragma omp parallel for
    for (int64_t k = 0; k < tyz.n3; ++k)
        for (int64_t j = 0; j < tyz.n2; ++j)
            for (int64_t i = 0; i < tyz.n1; ++i) {
                double ag = (A3(Gd, i + 1, j, k) + A3(Gd, i + 1, j + 1, k) + A3(Gd, i + 1, j, k + 1) + A3(Gd, i + 1, j + 1, k + 1)) * 0.25;
                double am = (A3(Mu, i + 1, j, k) + A3(Mu, i + 1, j + 1, k) + A3(Mu, i + 1, j, k + 1) + A3(Mu, i + 1, j + 1, k + 1)) * 0.25;
                double e = 0.5 * ((A3(Vy, i + 1, j + 1, k + 1) - A3(Vy, i + 1, j + 1, k)) / dz +
                                  (A3(Vz, i + 1, j + 1, k + 1) - A3(Vz, i + 1, j, k + 1)) / dy);
                A3(tyz, i, j, k) = tau_n(ve, A3(tyz, i, j, k), A3(tyzo, i, j, k), ag, am, Gdt, e);
            }
}

/* Momentum residuals in the 4-kernel (divide) form; k3 selects the component. */
static inline double r3_div(const or_state* s, int comp, int64_t i, int64_t j, int64_t k) {
    const arr Pt = s->f[ORF_PT], txx = s->f[ORF_TXX], tyy = s->f[ORF_TYY], tzz = s->f[ORF_TZZ], txy = s->f[ORF_TXY],
              txz = s->f[ORF_TXZ], tyz = s->f[ORF_TYZ], rg = s->f[ORF_RHOG];
    const double dx = s->dx, dy = s->dy, dz = s->dz;
    const int bf = s->p.rhog_in != 0.0 || s->p.rhog_out != 0.0;
    if (comp == 0)
        return (A3(txx, i + 1, j, k) - A3(txx, i, j, k)) / dx + (A3(txy, i, j + 1, k) - A3(txy, i, j, k)) / dy +
               (A3(txz, i, j, k + 1) - A3(txz, i, j, k)) / dz - (A3(Pt, i + 1, j + 1, k + 1) - A3(Pt, i, j + 1, k + 1)) / dx;
    if (comp == 1)
        return (A3(tyy, i, j + 1, k) - A3(tyy, i, j, k)) / dy + (A3(txy, i + 1, j, k) - A3(txy, i, j, k)) / dx +
               (A3(tyz, i, j, k + 1) - A3(tyz, i, j, k)) / dz - (A3(Pt, i + 1, j + 1, k + 1) - A3(Pt, i + 1, j, k + 1)) / dy;
    double R = (A3(tzz, i, j, k + 1) - A3(tzz, i, j, k)) / dz + (A3(txz, i + 1, j, k) - A3(txz, i, j, k)) / dx +
               (A3(tyz, i, j + 1, k) - A3(tyz, i, j, k)) / dy - (A3(Pt, i + 1, j + 1, k + 1) - A3(Pt, i + 1, j + 1, k)) / dz;
    if (bf) R = R - (A3(rg, i + 1, j + 1, k) + A3(rg, i + 1, j + 1, k + 1)) * 0.5;
    return R;
}
static inline double r3_mul(const or_state* s, int comp, int64_t i, int64_t j, int64_t k) {
    const arr Pt = s->f[ORF_PT], txx = s->f[ORF_TXX], tyy = s->f[ORF_TYY], tzz = s->f[ORF_TZZ], txy = s->f[ORF_TXY],
              txz = s->f[ORF_TXZ], tyz = s->f[ORF_TYZ], rg = s->f[ORF_RHOG];
    const double _dx = s->_dx, _dy = s->_dy, _dz = s->_dz;
    const int bf = s->p.rhog_in != 0.0 || s->p.rhog_out != 0.0;
    if (comp == 0)
        return _dx * (A3(txx, i + 1, j, k) - A3(txx, i, j, k)) + _dy * (A3(txy, i, j + 1, k) - A3(txy, i, j, k)) +
               _dz * (A3(txz, i, j, k + 1) - A3(txz, i, j, k)) - _dx * (A3(Pt, i + 1, j + 1, k + 1) - A3(Pt, i, j + 1, k + 1));
    if (comp == 1)
        return _dy * (A3(tyy, i, j + 1, k) - A3(tyy, i, j, k)) + _dx * (A3(txy, i + 1, j, k) - A3(txy, i, j, k)) +
               _dz * (A3(tyz, i, j, k + 1) - A3(tyz, i, j, k)) - _dy * (A3(Pt, i + 1, j + 1, k + 1) - A3(Pt, i + 1, j, k + 1));
    double R = _dz * (A3(tzz, i, j, k + 1) - A3(tzz, i, j, k)) + _dx * (A3(txz, i + 1, j, k) - A3(txz, i, j, k)) +
               _dy * (A3(tyz, i, j + 1, k) - A3(tyz, i, j, k)) - _dz * (A3(Pt, i + 1, j + 1, k + 1) - A3(Pt, i + 1, j + 1, k));
    if (bf) R = R - 0.5 * (A3(rg, i + 1, j + 1, k) + A3(rg, i + 1, j + 1, k + 1));
    return R;
}

/* compute_dV! (multixpu:79-84) */
static void p3_compute_dV(or_state* s) {
    arr Rh = s->f[ORF_DTAU_RHO];
    arr D[3] = {s->f[ORF_DVX], s->f[ORF_DVY], s->f[ORF_DVZ]};
    for (int c = 0; c < 3; ++c) {
        arr d = D[c];
#pragma omp parallel for
        for (int64_t k = 0; k < d.n3; ++k)
            for (int64_t j = 0; j < d.n2; ++j)
                for (int64_t i = 0; i < d.n1; ++i) {
                    double av;
                    if (c == 0) av = (A3(Rh, i, j + 1, k + 1) + A3(Rh, i + 1, j + 1, k + 1)) * 0.5;
                    else if (c == 1) av = (A3(Rh, i + 1, j, k + 1) + A3(Rh, i + 1, j + 1, k + 1)) * 0.5;
                    else av = (A3(Rh, i + 1, j + 1, k) + A3(Rh, i + 1, j + 1, k + 1)) * 0.5;
                    A3(d, i, j, k) = r3_div(s, c, i, j, k) * av;
                }
    }
}
/* compute_V! (multixpu:86-91) */
static void p3_compute_V(or_state* s) {
    arr V[3] = {s->f[ORF_VX], s->f[ORF_VY], s->f[ORF_VZ]};
    arr D[3] = {s->f[ORF_DVX], s->f[ORF_DVY], s->f[ORF_DVZ]};
    for (int c = 0; c < 3; ++c)
        for (int64_t k = 0; k < D[c].n3; ++k)
            for (int64_t j = 0; j < D[c].n2; ++j)
                for (int64_t i = 0; i < D[c].n1; ++i)
                    A3(V[c], i + 1, j + 1, k + 1) = A3(V[c], i + 1, j + 1, k + 1) + A3(D[c], i, j, k);
}
/* compute_Res! (multixpu:93-98) */
static void p3_compute_Res(or_state* s) {
    arr R[3] = {s->f[ORF_RX], s->f[ORF_RY], s->f[ORF_RZ]};
    for (int c = 0; c < 3; ++c)
        for (int64_t k = 0; k < R[c].n3; ++k)
            for (int64_t j = 0; j < R[c].n2; ++j)
                for (int64_t i = 0; i < R[c].n1; ++i) A3(R[c], i, j, k) = r3_div(s, c, i, j, k);
}

/* ------------------------------------------------------------------------ */
/* 3D, fused perf form: Stokes3D_ve_perf.jl:51-96                             */
/* ------------------------------------------------------------------------ */
static void f3_compute_Pt_tau(or_state* s) {
    const double _dx = s->_dx, _dy = s->_dy, _dz = s->_dz, r = s->p.r, Gdt = s->p.G * s->p.dt;
    const int ve = s->p.ve;
    arr Pt = s->f[ORF_PT], Gd = s->f[ORF_GDTAU], Mu = s->f[ORF_MUS], Vx = s->f[ORF_VX], Vy = s->f[ORF_VY],
        Vz = s->f[ORF_VZ];
    arr txx = s->f[ORF_TXX], tyy = s->f[ORF_TYY], tzz = s->f[ORF_TZZ], txy = s->f[ORF_TXY], txz = s->f[ORF_TXZ],
        tyz = s->f[ORF_TYZ];
    arr txxo = s->f[ORF_TXX_O], tyyo = s->f[ORF_TYY_O], tzzo = s->f[ORF_TZZ_O], txyo = s->f[ORF_TXY_O],
        txzo = s->f[ORF_TXZ_O], tyzo = s->f[ORF_TYZ_O];
    const int64_t nx = Pt.n1, ny = Pt.n2, nz = Pt.n3;
#pragma omp parallel for
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                A3(Pt, i, j, k) = A3(Pt, i, j, k) - r * A3(Gd, i, j, k) *
                    (_dx * (A3(Vx, i + 1, j, k) - A3(Vx, i, j, k)) + _dy * (A3(Vy, i, j + 1, k) - A3(Vy, i, j, k)) +
                     _dz * (A3(Vz, i, j, k + 1) - A3(Vz, i, j, k)));
                if (j < ny - 2 && k < nz - 2)
                    A3(txx, i, j, k) = tau_n(ve, A3(txx, i, j, k), A3(txxo, i, j, k), A3(Gd, i, j + 1, k + 1),
                                             A3(Mu, i, j + 1, k + 1), Gdt,
                                             _dx * (A3(Vx, i + 1, j + 1, k + 1) - A3(Vx, i, j + 1, k + 1)));
                if (i < nx - 2 && k < nz - 2)
                    A3(tyy, i, j, k) = tau_n(ve, A3(tyy, i, j, k), A3(tyyo, i, j, k), A3(Gd, i + 1, j, k + 1),
                                             A3(Mu, i + 1, j, k + 1), Gdt,
                                             _dy * (A3(Vy, i + 1, j + 1, k + 1) - A3(Vy, i + 1, j, k + 1)));
                if (i < nx - 2 && j < ny - 2)
                    A3(tzz, i, j, k) = tau_n(ve, A3(tzz, i, j, k), A3(tzzo, i, j, k), A3(Gd, i + 1, j + 1, k),
                                             A3(Mu, i + 1, j + 1, k), Gdt,
                                             _dz * (A3(Vz, i + 1, j + 1, k + 1) - A3(Vz, i + 1, j + 1, k)));
                if (i < nx - 1 && j < ny - 1 && k < nz - 2) {
                    double ag = (A3(Gd, i, j, k + 1) + A3(Gd, i + 1, j, k + 1) + A3(Gd, i, j + 1, k + 1) + A3(Gd, i + 1, j + 1, k + 1)) * 0.25;
                    double am = (A3(Mu, i, j, k + 1) + A3(Mu, i + 1, j, k + 1) + A3(Mu, i, j + 1, k + 1) + A3(Mu, i + 1, j + 1, k + 1)) * 0.25;
                    double e = 0.5 * (_dy * (A3(Vx, i + 1, j + 1, k + 1) - A3(Vx, i + 1, j, k + 1)) +
                                      _dx * (A3(Vy, i + 1, j + 1, k + 1) - A3(Vy, i, j + 1, k + 1)));
                    A3(txy, i, j, k) = tau_n(ve, A3(txy, i, j, k), A3(txyo, i, j, k), ag, am, Gdt, e);
                }
                if (i < nx - 1 && j < ny - 2 && k < nz - 1) {
                    double ag = (A3(Gd, i, j + 1, k) + A3(Gd, i + 1, j + 1, k) + A3(Gd, i, j + 1, k + 1) + A3(Gd, i + 1, j + 1, k + 1)) * 0.25;
                    double am = (A3(Mu, i, j + 1, k) + A3(Mu, i + 1, j + 1, k) + A3(Mu, i, j + 1, k + 1) + A3(Mu, i + 1, j + 1, k + 1)) * 0.25;
                    double e = 0.5 * (_dz * (A3(Vx, i + 1, j + 1, k + 1) - A3(Vx, i + 1, j + 1, k)) +
                                      _dx * (A3(Vz, i + 1, j + 1, k + 1) - A3(Vz, i, j + 1, k + 1)));
                    A3(txz, i, j, k) = tau_n(ve, A3(txz, i, j, k), A3(txzo, i, j, k), ag, am, Gdt, e);
                }
                if (i < nx - 2 && j < ny - 1 && k < nz - 1) {
                    double ag = (A3(Gd, i + 1, j, k) + A3(Gd, i + 1, j + 1, k) + A3(Gd, i + 1, j, k + 1) + A3(Gd, i + 1, j + 1, k + 1)) * 0.25;
                    double am = (A3(Mu, i + 1, j, k) + A3(Mu, i + 1, j + 1, k) + A3(Mu, i + 1, j, k + 1) + A3(Mu, i + 1, j + 1, k + 1)) * 0.25;
                    double e = 0.5 * (_dz * (A3(Vy, i + 1, j + 1, k + 1) - A3(Vy, i + 1, j + 1, k)) +
                                      _dy * (A3(Vz, i + 1, j + 1, k + 1) - A3(Vz, i + 1, j, k + 1)));
                    A3(tyz, i, j, k) = tau_n(ve, A3(tyz, i, j, k), A3(tyzo, i, j, k), ag, am, Gdt, e);
                }
            }
}
static void f3_compute_V(or_state* s) {
    arr Vx = s->f[ORF_VX], Vy = s->f[ORF_VY], Vz = s->f[ORF_VZ], Rh = s->f[ORF_DTAU_RHO];
    const int64_t nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
#pragma omp parallel for
    for (int64_t k = 0; k < nz - 1; ++k)
        for (int64_t j = 0; j < ny - 1; ++j)
            for (int64_t i = 0; i < nx - 1; ++i) {
                if (i <= nx - 2 && j <= ny - 3 && k <= nz - 3)
                    A3(Vx, i + 1, j + 1, k + 1) = A3(Vx, i + 1, j + 1, k + 1) + r3_mul(s, 0, i, j, k) *
                        ((A3(Rh, i, j + 1, k + 1) + A3(Rh, i + 1, j + 1, k + 1)) * 0.5);
                if (i <= nx - 3 && j <= ny - 2 && k <= nz - 3)
                    A3(Vy, i + 1, j + 1, k + 1) = A3(Vy, i + 1, j + 1, k + 1) + r3_mul(s, 1, i, j, k) *
                        ((A3(Rh, i + 1, j, k + 1) + A3(Rh, i + 1, j + 1, k + 1)) * 0.5);
                if (i <= nx - 3 && j <= ny - 3 && k <= nz - 2)
                    A3(Vz, i + 1, j + 1, k + 1) = A3(Vz, i + 1, j + 1, k + 1) + r3_mul(s, 2, i, j, k) *
                        ((A3(Rh, i + 1, j + 1, k) + A3(Rh, i + 1, j + 1, k + 1)) * 0.5);
            }
}
static void f3_compute_Res(or_state* s) {
    arr dV = s->f[ORF_DIVV], Vx = s->f[ORF_VX], Vy = s->f[ORF_VY], Vz = s->f[ORF_VZ];
    const double _dx = s->_dx, _dy = s->_dy, _dz = s->_dz;
    for (int64_t k = 0; k < dV.n3; ++k)
        for (int64_t j = 0; j < dV.n2; ++j)
            for (int64_t i = 0; i < dV.n1; ++i)
                A3(dV, i, j, k) = _dx * (A3(Vx, i + 1, j, k) - A3(Vx, i, j, k)) + _dy * (A3(Vy, i, j + 1, k) - A3(Vy, i, j, k)) +
                                  _dz * (A3(Vz, i, j, k + 1) - A3(Vz, i, j, k));
    arr R[3] = {s->f[ORF_RX], s->f[ORF_RY], s->f[ORF_RZ]};
    for (int c = 0; c < 3; ++c)
        for (int64_t k = 0; k < R[c].n3; ++k)
            for (int64_t j = 0; j < R[c].n2; ++j)
                for (int64_t i = 0; i < R[c].n1; ++i) A3(R[c], i, j, k) = r3_mul(s, c, i, j, k);
}

/* ------------------------------------------------------------------------ */
/* Iteration and solve                                                        */
/* ------------------------------------------------------------------------ */
static void bcs_V(or_state* s) {
    if (s->p.ndim == 2) {
        /* Stokes2D_ve.jl:159-160 */
        bc_y(s->f[ORF_VX]);
        bc_x(s->f[ORF_VY]);
    } else {
        /* Stokes3D_ve_perf.jl:202-207 */
        bc_x(s->f[ORF_VY]);
        bc_x(s->f[ORF_VZ]);
        bc_y(s->f[ORF_VX]);
        bc_y(s->f[ORF_VZ]);
        bc_z(s->f[ORF_VX]);
        bc_z(s->f[ORF_VY]);
    }
}

void or_iterate(or_state* s, int64_t n) {
    for (int64_t q = 0; q < n; ++q) {
        if (s->p.ndim == 2) {
            if (s->p.perf) {
                f2_compute_Pt_tau(s);
                f2_compute_V(s);
            } else {
                p2_compute_P(s);
                p2_compute_tau(s);
                /* Stokes2D.jl stores Rx,Ry in compute_dV!; keeping them
                 * always is harmless for the VE script (it overwrites them
                 * in compute_Res! before use). */
                p2_compute_dV(s, 1);
                p2_compute_V(s);
            }
        } else {
            if (s->p.perf) {
                f3_compute_Pt_tau(s);
                f3_compute_V(s);
            } else {
                p3_compute_P(s);
                p3_compute_tau(s);
                p3_compute_dV(s);
                p3_compute_V(s);
            }
        }
        bcs_V(s);
    }
}

int or_kernel(or_state* s, int which, int arg) {
    const int d2 = s->p.ndim == 2;
    switch (which) {
        case 0: if (d2) p2_compute_P(s); else p3_compute_P(s); return 0;
        case 1: if (d2) p2_compute_tau(s); else p3_compute_tau(s); return 0;
        case 2: if (d2) p2_compute_dV(s, 1); else p3_compute_dV(s); return 0;
        case 3: if (d2) p2_compute_V(s); else p3_compute_V(s); return 0;
        case 4: if (d2) p2_compute_Res(s); else p3_compute_Res(s); return 0;
        case 5: if (d2) f2_compute_Pt_tau(s); else f3_compute_Pt_tau(s); return 0;
        case 6: if (d2) f2_compute_V(s); else f3_compute_V(s); return 0;
        case 7: if (d2) f2_compute_Res(s); else f3_compute_Res(s); return 0;
        case 8: bcs_V(s); return 0;
        case 10: if (arg < 0 || arg >= ORF_NFIELDS) return -1; bc_x(s->f[arg]); return 0;
        case 11: if (arg < 0 || arg >= ORF_NFIELDS) return -1; bc_y(s->f[arg]); return 0;
        case 12: if (arg < 0 || arg >= ORF_NFIELDS) return -1; bc_z(s->f[arg]); return 0;
        default: return -1;
    }
}

void or_compute_res(or_state* s) {
    if (s->p.ndim == 2) {
        if (s->p.perf) f2_compute_Res(s);
        else p2_compute_Res(s);
    } else {
        if (s->p.perf) f3_compute_Res(s);
        else p3_compute_Res(s);
    }
}

static void minmax(const arr* A, double* mn, double* mx) {
    double a = INFINITY, b = -INFINITY;
    for (int64_t q = 0; q < len(A); ++q) {
        double v = A->a[q];
        if (v < a) a = v;
        if (v > b) b = v;
    }
    *mn = a; *mx = b;
}
static double nrm2(const arr* A) {
    double acc = 0.0;
    for (int64_t q = 0; q < len(A); ++q) acc += A->a[q] * A->a[q];
    return sqrt(acc);
}
static double nrminf(const arr* A) {
    double m = 0.0;
    for (int64_t q = 0; q < len(A); ++q) {
        double v = fabs(A->a[q]);
        if (v > m || v != v) m = v;
    }
    return m;
}

/* Stokes2D_ve.jl:164-171, Stokes3D_ve_perf.jl:211-219 (err = max of the
 * scaled RMS norms).  norm_inf replaces ‖·‖₂/√len by max|·|. */
void or_check(or_state* s, double out[9]) {
    double Vmin, Vmax, Pmin, Pmax;
    minmax(&s->f[ORF_VX], &Vmin, &Vmax);
    minmax(&s->f[ORF_PT], &Pmin, &Pmax);
    const double lx = s->p.lx;
    const int ids[4] = {ORF_RX, ORF_RY, ORF_RZ, ORF_DIVV};
    double err = -INFINITY;
    int nan = 0;
    for (int c = 0; c < 4; ++c) {
        const arr* A = &s->f[ids[c]];
        if (c == 2 && s->p.ndim == 2) { out[c] = 0.0; continue; }
        double den = (c == 3) ? (Vmax - Vmin) : (Pmax - Pmin);
        double v;
        if (s->p.norm_inf) v = nrminf(A) / den * lx;
        else v = nrm2(A) / den * lx / sqrt((double)len(A));
        out[c] = v;
        if (v != v) nan = 1;
        if (v > err) err = v;
    }
    out[4] = nan ? NAN : err;
    out[5] = Vmin; out[6] = Vmax; out[7] = Pmin; out[8] = Pmax;
}

int or_solve(or_state* s, int64_t* ittot_out, double* err_out, int64_t* per_step) {
    const or_params* p = &s->p;
    int64_t ittot = 0;
    double err = 0.0;
    for (int it = 0; it < p->nt; ++it) {
        err = 2.0 * p->eps;
        int64_t iter = 0;
        if (p->ve) or_assign_tau(s);
        while (err > p->eps && iter <= p->iter_max) {
            or_iterate(s, 1);
            iter += 1;
            if (iter % p->nout == 0) {
                /* Stokes2D.jl uses Rx,Ry from compute_dV!; the VE scripts
                 * call compute_Res! -- both equal the residual of the
                 * current Pt/τ.  The 4-kernel scripts' ∇V is the one from
                 * compute_P! (before the V update); the perf scripts
                 * recompute it in compute_Res!. */
                if (p->perf || p->ve) or_compute_res(s);
                double o[9];
                or_check(s, o);
                err = o[4];
                if (err != err) {
                    if (ittot_out) *ittot_out = ittot + iter;
                    if (err_out) *err_out = err;
                    return -1;
                }
            }
        }
        if (per_step) per_step[it] = iter;
        ittot += iter;
    }
    if (ittot_out) *ittot_out = ittot;
    if (err_out) *err_out = err;
    return 0;
}
