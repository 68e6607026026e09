// pts_setup.hip -- setup kernels: fill/copy, initial conditions, smoothing,
// maxloc, free-slip copies and iteration parameters.  Each restates one
// reference kernel (cited) with the same floating-point association; the
// library is built with -ffp-contract=off so no a*b+c is fused.
#include "pts_internal.hpp"

namespace pts {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ void unflat(int64_t q, const int64_t n1, const int64_t n2, int& i, int& j, int& k) {
    i = (int)(q % n1);
    int64_t r = q / n1;
    j = (int)(r % n2);
    k = (int)(r / n2);
}

__global__ void k_fill(double* p, double v, int64_t n1, int64_t n2, int64_t n3, int sx, int sxy) {
    const int64_t tot = n1 * n2 * n3;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        unflat(q, n1, n2, i, j, k);
        p[i + sx * j + (int64_t)sxy * k] = v;
    }
}

__global__ void k_copy(double* dst, const double* src, int64_t n1, int64_t n2, int64_t n3, int sx, int sxy) {
    const int64_t tot = n1 * n2 * n3;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        unflat(q, n1, n2, i, j, k);
        const int64_t o = i + sx * j + (int64_t)sxy * k;
        dst[o] = src[o];
    }
}

// Initial conditions: Stokes2D_ve.jl:131-136 (2D), Stokes3D_ve_perf.jl:167-172
// (3D), Stokes3D_ve_multixpu.jl:174-176 (igg: x_g(ix,dx,A) = (ix-1)*dx - dx/2
// for an (nx+1)-sized array on one rank).
__global__ void k_init(Fields f, int ndim, double dx, double dy, double dz, double lx, double ly, double lz,
                       double mus0, double musi, double rad2, double ebg, double rin, double rout, int igg) {
    const int nx = f.nx, ny = f.ny, nz = ndim == 3 ? f.nz : 1;
    // index space covers (nx+1) x (ny+1) x (nz+1 or 1)
    const int64_t n1 = nx + 1, n2 = ny + 1, n3 = ndim == 3 ? nz + 1 : 1;
    const int64_t tot = n1 * n2 * n3;
    double* Mus = f.p[PTS_MUS];
    double* rg = f.p[PTS_RHOG];
    double* Vx = f.p[PTS_VX];
    double* Vy = f.p[PTS_VY];
    double* Vz = f.p[PTS_VZ];
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        unflat(q, n1, n2, i, j, k);
        const int64_t o = i + f.sx * j + (int64_t)f.sxy * k;
        if (ndim == 2) {
            if (i < nx && j < ny) {
                double xr = ((double)i * dx + 0.5 * dx) - 0.5 * lx;
                double yr = ((double)j * dy + 0.5 * dy) - 0.5 * ly;
                double r2 = xr * xr + yr * yr;
                bool in = r2 < rad2;
                Mus[o] = in ? musi : mus0 * 1.0;
                if (rg) rg[o] = in ? rin : rout;
            }
            if (j < ny) Vx[o] = (-ebg) * ((double)i * dx - 0.5 * lx);
            if (i < nx) Vy[o] = ebg * ((double)j * dy - 0.5 * ly);
        } else {
            if (i < nx && j < ny && k < nz) {
                double xr = ((double)i * dx + 0.5 * dx) - 0.5 * lx;
                double yr = ((double)j * dy + 0.5 * dy) - 0.5 * ly;
                double zr = ((double)k * dz + 0.5 * dz) - 0.5 * lz;
                double r2 = xr * xr + yr * yr + zr * zr;
                bool in = r2 < rad2;
                Mus[o] = in ? musi : mus0 * 1.0;
                if (rg) rg[o] = in ? rin : rout;
            }
            if (j < ny && k < nz) {
                double x = igg ? ((((double)i * dx + (-0.5 * dx)) + 0.5 * dx) - 0.5 * lx) : ((double)i * dx - 0.5 * lx);
                Vx[o] = (-ebg) * x;
            }
            if (i < nx && k < nz) Vy[o] = 0.0;
            if (i < nx && j < ny) {
                double z = igg ? ((((double)k * dz + (-0.5 * dz)) + 0.5 * dz) - 0.5 * lz) : ((double)k * dz - 0.5 * lz);
                Vz[o] = ebg * z;
            }
        }
    }
}

// smooth!: Stokes2D_ve.jl:18-21, Stokes3D_ve_perf.jl:24-27.
__global__ void k_smooth(double* dst, const double* src, int ndim, int64_t n1, int64_t n2, int64_t n3, int sx,
                         int sxy, double c) {
    const int64_t m1 = n1 - 2, m2 = n2 - 2, m3 = ndim == 3 ? n3 - 2 : 1;
    const int64_t tot = m1 * m2 * m3;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        unflat(q, m1, m2, i, j, k);
        if (ndim == 2) {
            const int64_t o = (i + 1) + sx * (int64_t)(j + 1);
            double m = src[o];
            double d2x = (src[o + 1] - m) - (m - src[o - 1]);
            double d2y = (src[o + sx] - m) - (m - src[o - sx]);
            dst[o] = m + c * (d2x + d2y);
        } else {
            const int64_t o = (i + 1) + sx * (int64_t)(j + 1) + (int64_t)sxy * (k + 1);
            double m = src[o];
            double d2x = (src[o + 1] - m) - (m - src[o - 1]);
            double d2y = (src[o + sx] - m) - (m - src[o - sx]);
            double d2z = (src[o + sxy] - m) - (m - src[o - sxy]);
            dst[o] = m + c * (d2x + d2y + d2z);
        }
    }
}

// compute_maxloc!: Stokes2D_ve.jl:23-26 (5-point cross), Stokes3D_ve_perf.jl:29-32 (7-point).
__global__ void k_maxloc(double* dst, const double* src, int ndim, int64_t n1, int64_t n2, int64_t n3, int sx,
                         int sxy) {
    const int64_t m1 = n1 - 2, m2 = n2 - 2, m3 = ndim == 3 ? n3 - 2 : 1;
    const int64_t tot = m1 * m2 * m3;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        unflat(q, m1, m2, i, j, k);
        if (ndim == 2) {
            const int64_t o = (i + 1) + sx * (int64_t)(j + 1);
            dst[o] = fmax(fmax(fmax(fmax(src[o - 1], src[o + 1]), src[o]), src[o - sx]), src[o + sx]);
        } else {
            const int64_t o = (i + 1) + sx * (int64_t)(j + 1) + (int64_t)sxy * (k + 1);
            double m = fmax(src[o - 1], src[o + 1]);
            m = fmax(m, src[o]);
            m = fmax(m, src[o - sx]);
            m = fmax(m, src[o + sx]);
            m = fmax(m, src[o - sxy]);
            m = fmax(m, src[o + sxy]);
            dst[o] = m;
        }
    }
}

// Musτ == maxloc(μ) + free-slip copies, cell by cell (2D, bitwise).
__global__ void k_check_mustau2(Fields f, int* bad) {
    const int nx = f.nx, ny = f.ny, sx = f.sx;
    const int64_t tot = (int64_t)nx * ny;
    const double* M = f.p[PTS_MUS];
    const double* Mt = f.p[PTS_MUSTAU];
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % nx), j = (int)(q / nx);
        const int ic = min(max(i, 1), nx - 2), jc = min(max(j, 1), ny - 2);
        const int om = ic + sx * jc;
        const double m = fmax(fmax(fmax(fmax(M[om - 1], M[om + 1]), M[om]), M[om - sx]), M[om + sx]);
        const double t = Mt[i + sx * j];
        if (!(m == t)) atomicOr(bad, 1);
    }
}

// 3D: Musτ == maxloc(μ) + bc_x!, bc_y!, bc_z! (Stokes3D_ve_perf.jl:29-32,
// 178-182), cell by cell and bitwise: after the three copies every cell holds
// the cross maximum at its position clamped into the interior.
__global__ void k_check_mustau3(Fields f, int* bad) {
    const int nx = f.nx, ny = f.ny, nz = f.nz, sx = f.sx, sxy = f.sxy;
    const int64_t tot = (int64_t)nx * ny * nz;
    const double* M = f.p[PTS_MUS];
    const double* Mt = f.p[PTS_MUSTAU];
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((int64_t)nx * ny));
        const int ic = min(max(i, 1), nx - 2), jc = min(max(j, 1), ny - 2), kc = min(max(k, 1), nz - 2);
        const int64_t o = ic + sx * (int64_t)jc + (int64_t)sxy * kc;
        double m = fmax(M[o - 1], M[o + 1]);
        m = fmax(m, M[o]);
        m = fmax(m, M[o - sx]);
        m = fmax(m, M[o + sx]);
        m = fmax(m, M[o - sxy]);
        m = fmax(m, M[o + sxy]);
        const double t = Mt[i + sx * (int64_t)j + (int64_t)sxy * k];
        if (!(m == t)) atomicOr(bad, 1);
    }
}

// bc_x!/bc_y!/bc_z!: copy the first interior plane outward along `dim`.
__global__ void k_bc(double* a, int64_t n1, int64_t n2, int64_t n3, int sx, int sxy, int dim) {
    int64_t e1 = n1, e2 = n2, e3 = n3;
    if (dim == 0) e1 = 1;
    else if (dim == 1) e2 = 1;
    else e3 = 1;
    const int64_t tot = e1 * e2 * e3;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        unflat(q, e1, e2, i, j, k);
        const int64_t base = i + sx * (int64_t)j + (int64_t)sxy * k;
        const int64_t st = dim == 0 ? 1 : (dim == 1 ? sx : sxy);
        const int64_t n = dim == 0 ? n1 : (dim == 1 ? n2 : n3);
        a[base] = a[base + st];
        a[base + (n - 1) * st] = a[base + (n - 2) * st];
    }
}

// compute_iter_params!: Stokes2D.jl:27-31 (viscous), Stokes2D_ve.jl:28-33 and
// Stokes3D_ve_perf.jl:34-39 (VE, Mu_eff = 1/(1/Musτ + 1/(G*dt))).
__global__ void k_iter_params(Fields f, int ndim, int ve, double Vp, double G, double dt, double Re, double r,
                              double ml) {
    const int64_t n1 = f.nx, n2 = f.ny, n3 = ndim == 3 ? f.nz : 1;
    const int64_t tot = n1 * n2 * n3;
    const double* M = f.p[PTS_MUSTAU];
    double* R = f.p[PTS_DTAU_RHO];
    double* Gd = f.p[PTS_GDTAU];
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        unflat(q, n1, n2, i, j, k);
        const int64_t o = i + f.sx * (int64_t)j + (int64_t)f.sxy * k;
        double mu_eff = ve ? 1.0 / (1.0 / M[o] + 1.0 / (G * dt)) : M[o];
        double rho = Vp * ml / Re / mu_eff;
        R[o] = rho;
        Gd[o] = Vp * Vp / rho / (r + 2.0);
    }
}

// Gaussian random field by random Fourier modes: f(x) = sigma*sqrt(2/N) *
// sum_m cos(k_m . x + phi_m) at the cell centres; Mus = 10^clip(f, lo, hi)
// outside the inclusion (Rad2 < rad2), musi inside.  The modes (kx, ky, kz,
// phi per mode) are staged in LDS by each block.
constexpr int kGrfChunk = 512;
__global__ __launch_bounds__(256) void k_grf(Fields f, int ndim, const double* __restrict__ modes, int nmodes,
                                             double amp, double lo, double hi, double dx, double dy, double dz,
                                             double lx, double ly, double lz, double musi, double rad2) {
    __shared__ double sm[kGrfChunk * 4];
    const int nx = f.nx, ny = f.ny, nz = ndim == 3 ? f.nz : 1;
    const int64_t tot = (int64_t)nx * ny * nz;
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int i = 0, j = 0, k = 0;
    if (q < tot) unflat(q, nx, ny, i, j, k);
    const double x = (double)i * dx + 0.5 * dx, y = (double)j * dy + 0.5 * dy, z = (double)k * dz + 0.5 * dz;
    double acc = 0.0;
    for (int m0 = 0; m0 < nmodes; m0 += kGrfChunk) {
        const int nm = min(kGrfChunk, nmodes - m0);
        __syncthreads();
        for (int t = threadIdx.x; t < nm * 4; t += blockDim.x) sm[t] = modes[(int64_t)m0 * 4 + t];
        __syncthreads();
        for (int m = 0; m < nm; ++m) {
            double ph = sm[4 * m] * x + sm[4 * m + 1] * y + sm[4 * m + 3];
            if (ndim == 3) ph += sm[4 * m + 2] * z;
            acc += cos(ph);
        }
    }
    if (q >= tot) return;
    const double xr = x - 0.5 * lx, yr = y - 0.5 * ly, zr = ndim == 3 ? z - 0.5 * lz : 0.0;
    const double r2 = xr * xr + yr * yr + zr * zr;
    const double lg = fmin(fmax(amp * acc, lo), hi);
    f.p[PTS_MUS][i + f.sx * (int64_t)j + (int64_t)f.sxy * k] = r2 < rad2 ? musi : pow(10.0, lg);
}

int grid_for(int64_t tot) {
    int64_t g = (tot + kBlock - 1) / kBlock;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (int)g;
}

}  // namespace

hipError_t launch_fill(hipStream_t s, double* p, double v, const Fields& f, const int64_t sh[3]) {
    const int64_t tot = sh[0] * sh[1] * sh[2];
    if (tot == 0) return hipSuccess;
    k_fill<<<grid_for(tot), kBlock, 0, s>>>(p, v, sh[0], sh[1], sh[2], f.sx, f.sxy);
    return hipGetLastError();
}

hipError_t launch_copy(hipStream_t s, double* dst, const double* src, const Fields& f, const int64_t sh[3]) {
    const int64_t tot = sh[0] * sh[1] * sh[2];
    if (tot == 0) return hipSuccess;
    k_copy<<<grid_for(tot), kBlock, 0, s>>>(dst, src, sh[0], sh[1], sh[2], f.sx, f.sxy);
    return hipGetLastError();
}

hipError_t launch_init_inclusion(hipStream_t s, const Fields& f, int ndim, double lx, double ly, double lz,
                                 double mus0, double musi, double rad2, double ebg, double rin, double rout,
                                 int igg) {
    const double dx = lx / f.nx, dy = ly / f.ny, dz = ndim == 3 ? lz / f.nz : 1.0;
    const int64_t tot = (int64_t)(f.nx + 1) * (f.ny + 1) * (ndim == 3 ? f.nz + 1 : 1);
    k_init<<<grid_for(tot), kBlock, 0, s>>>(f, ndim, dx, dy, dz, lx, ly, lz, mus0, musi, rad2, ebg, rin, rout, igg);
    return hipGetLastError();
}

hipError_t launch_smooth(hipStream_t s, const Fields& f, int ndim, double* dst, const double* src,
                         const int64_t sh[3], double fact) {
    const double c = (ndim == 2 ? 1.0 / 4.1 : 1.0 / 6.1) / fact;
    const int64_t tot = (sh[0] - 2) * (sh[1] - 2) * (ndim == 3 ? sh[2] - 2 : 1);
    if (tot <= 0) return hipSuccess;
    k_smooth<<<grid_for(tot), kBlock, 0, s>>>(dst, src, ndim, sh[0], sh[1], sh[2], f.sx, f.sxy, c);
    return hipGetLastError();
}

hipError_t launch_maxloc(hipStream_t s, const Fields& f, int ndim, double* dst, const double* src,
                         const int64_t sh[3]) {
    const int64_t tot = (sh[0] - 2) * (sh[1] - 2) * (ndim == 3 ? sh[2] - 2 : 1);
    if (tot <= 0) return hipSuccess;
    k_maxloc<<<grid_for(tot), kBlock, 0, s>>>(dst, src, ndim, sh[0], sh[1], sh[2], f.sx, f.sxy);
    return hipGetLastError();
}

hipError_t launch_bc(hipStream_t s, const Fields& f, double* a, const int64_t sh[3], int dim) {
    const int64_t n = dim == 0 ? sh[0] : (dim == 1 ? sh[1] : sh[2]);
    if (n < 2) return hipSuccess;
    int64_t tot = sh[0] * sh[1] * sh[2] / n;
    k_bc<<<grid_for(tot), kBlock, 0, s>>>(a, sh[0], sh[1], sh[2], f.sx, f.sxy, dim);
    return hipGetLastError();
}

hipError_t launch_grf(hipStream_t s, const Fields& f, int ndim, const double* modes, int nmodes, double sigma,
                      double lo, double hi, double lx, double ly, double lz, double musi, double rad2) {
    const int64_t tot = (int64_t)f.nx * f.ny * (ndim == 3 ? f.nz : 1);
    const double amp = sigma * sqrt(2.0 / nmodes);
    const double dx = lx / f.nx, dy = ly / f.ny, dz = ndim == 3 ? lz / f.nz : 1.0;
    k_grf<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(f, ndim, modes, nmodes, amp, lo, hi, dx, dy, dz, lx, ly, lz,
                                                        musi, rad2);
    return hipGetLastError();
}

hipError_t launch_iter_params(hipStream_t s, const Fields& f, int ndim, bool ve, double Vp, double G, double dt,
                              double Re, double r, double ml) {
    const int64_t tot = (int64_t)f.nx * f.ny * (ndim == 3 ? f.nz : 1);
    k_iter_params<<<grid_for(tot), kBlock, 0, s>>>(f, ndim, ve ? 1 : 0, Vp, G, dt, Re, r, ml);
    return hipGetLastError();
}

hipError_t launch_check_mustau2(hipStream_t s, const Fields& f, int* bad) {
    const int64_t tot = (int64_t)f.nx * f.ny;
    k_check_mustau2<<<grid_for(tot), kBlock, 0, s>>>(f, bad);
    return hipGetLastError();
}

hipError_t launch_check_mustau3(hipStream_t s, const Fields& f, int* bad) {
    const int64_t tot = (int64_t)f.nx * f.ny * f.nz;
    k_check_mustau3<<<grid_for(tot), kBlock, 0, s>>>(f, bad);
    return hipGetLastError();
}

}  // namespace pts
