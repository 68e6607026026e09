// pts_step3d.hip -- 3D PT iteration kernels.
//
// Per-kernel reference API (Stokes3D_ve_multixpu.jl:57-98 divide form,
// Stokes3D_ve_perf.jl:69-96 multiply form) and the in-place two-kernel
// iteration (compute_Pt_τ! + compute_V! with the six free-slip copies folded
// into the V update).  Index: 0-based (i,j,k), offset i + sx*j + sxy*k.
#include "pts_internal.hpp"

namespace pts {

namespace {

constexpr int BX = 64, BY = 4;

inline dim3 grid3(int n1, int n2, int n3) { return dim3((n1 + BX - 1) / BX, (n2 + BY - 1) / BY, n3); }

#define IJK                                              \
    const int i = blockIdx.x * BX + threadIdx.x;         \
    const int j = blockIdx.y * BY + threadIdx.y;         \
    const int k = blockIdx.z;                            \
    const int nx = f.nx, ny = f.ny, nz = f.nz;           \
    const int sx = f.sx, sxy = f.sxy;                    \
    const int o = i + sx * j + sxy * k;                  \
    (void)nx; (void)ny; (void)nz; (void)o; (void)sx; (void)sxy;

// ∇V at cell o (either flavour)
template <bool MUL>
__device__ __forceinline__ double divv3(const Fields& f, const StepScalars& c, int o) {
    const double* Vx = f.p[PTS_VX];
    const double* Vy = f.p[PTS_VY];
    const double* Vz = f.p[PTS_VZ];
    return dfd<MUL>(Vx[o + 1] - Vx[o], c.dx, c._dx) + dfd<MUL>(Vy[o + f.sx] - Vy[o], c.dy, c._dy) +
           dfd<MUL>(Vz[o + f.sxy] - Vz[o], c.dz, c._dz);
}

// Momentum residuals (Stokes3D_ve_multixpu.jl:94-96, Stokes3D_ve_perf.jl:92-94)
template <bool MUL>
__device__ __forceinline__ double res3_x(const Fields& f, const StepScalars& c, int o) {
    const int sx = f.sx, sxy = f.sxy;
    const double* txx = f.p[PTS_TXX];
    const double* txy = f.p[PTS_TXY];
    const double* txz = f.p[PTS_TXZ];
    const double* Pt = f.p[PTS_PT];
    return dfd<MUL>(txx[o + 1] - txx[o], c.dx, c._dx) + dfd<MUL>(txy[o + sx] - txy[o], c.dy, c._dy) +
           dfd<MUL>(txz[o + sxy] - txz[o], c.dz, c._dz) -
           dfd<MUL>(Pt[o + 1 + sx + sxy] - Pt[o + sx + sxy], c.dx, c._dx);
}
template <bool MUL>
__device__ __forceinline__ double res3_y(const Fields& f, const StepScalars& c, int o) {
    const int sx = f.sx, sxy = f.sxy;
    const double* tyy = f.p[PTS_TYY];
    const double* txy = f.p[PTS_TXY];
    const double* tyz = f.p[PTS_TYZ];
    const double* Pt = f.p[PTS_PT];
    return dfd<MUL>(tyy[o + sx] - tyy[o], c.dy, c._dy) + dfd<MUL>(txy[o + 1] - txy[o], c.dx, c._dx) +
           dfd<MUL>(tyz[o + sxy] - tyz[o], c.dz, c._dz) -
           dfd<MUL>(Pt[o + 1 + sx + sxy] - Pt[o + 1 + sxy], c.dy, c._dy);
}
template <bool MUL>
__device__ __forceinline__ double res3_z(const Fields& f, const StepScalars& c, int o, bool bf) {
    const int sx = f.sx, sxy = f.sxy;
    const double* tzz = f.p[PTS_TZZ];
    const double* txz = f.p[PTS_TXZ];
    const double* tyz = f.p[PTS_TYZ];
    const double* Pt = f.p[PTS_PT];
    double R = dfd<MUL>(tzz[o + sxy] - tzz[o], c.dz, c._dz) + dfd<MUL>(txz[o + 1] - txz[o], c.dx, c._dx) +
               dfd<MUL>(tyz[o + sx] - tyz[o], c.dy, c._dy) -
               dfd<MUL>(Pt[o + 1 + sx + sxy] - Pt[o + 1 + sx], c.dz, c._dz);
    if (bf) {
        const double* rg = f.p[PTS_RHOG];
        R = R - (rg[o + 1 + sx] + rg[o + 1 + sx + sxy]) * 0.5;
    }
    return R;
}

// compute_P! (multixpu:57-61)
__global__ __launch_bounds__(BX* BY) void k3_P(Fields f, StepScalars c) {
    IJK;
    if (i < nx && j < ny && k < nz) {
        double d = divv3<false>(f, c, o);
        f.p[PTS_DIVV][o] = d;
        f.p[PTS_PT][o] = f.p[PTS_PT][o] - c.r * f.p[PTS_GDTAU][o] * d;
    }
}

// τ updates at (i,j,k) for every component (multixpu:70-75, perf:71-76).
template <bool VE, bool MUL>
__device__ __forceinline__ void tau3_all(const Fields& f, const StepScalars& c, int i, int j, int k, int o) {
    const int nx = f.nx, ny = f.ny, nz = f.nz, sx = f.sx, sxy = f.sxy;
    const double* Vx = f.p[PTS_VX];
    const double* Vy = f.p[PTS_VY];
    const double* Vz = f.p[PTS_VZ];
    const double* Gd = f.p[PTS_GDTAU];
    const double* Mu = f.p[PTS_MUS];
    if (i < nx && j < ny - 2 && k < nz - 2) {
        const int q = o + sx + sxy;
        double e = dfd<MUL>(Vx[q + 1] - Vx[q], c.dx, c._dx);
        double to = VE ? f.p[PTS_TXX_O][o] : 0.0;
        f.p[PTS_TXX][o] = tau_upd<VE>(f.p[PTS_TXX][o], to, Gd[q], Mu[q], c.Gdt, e);
    }
    if (i < nx - 2 && j < ny && k < nz - 2) {
        const int q = o + 1 + sxy;
        double e = dfd<MUL>(Vy[q + sx] - Vy[q], c.dy, c._dy);
        double to = VE ? f.p[PTS_TYY_O][o] : 0.0;
        f.p[PTS_TYY][o] = tau_upd<VE>(f.p[PTS_TYY][o], to, Gd[q], Mu[q], c.Gdt, e);
    }
    if (i < nx - 2 && j < ny - 2 && k < nz) {
        const int q = o + 1 + sx;
        double e = dfd<MUL>(Vz[q + sxy] - Vz[q], c.dz, c._dz);
        double to = VE ? f.p[PTS_TZZ_O][o] : 0.0;
        f.p[PTS_TZZ][o] = tau_upd<VE>(f.p[PTS_TZZ][o], to, Gd[q], Mu[q], c.Gdt, e);
    }
    if (i < nx - 1 && j < ny - 1 && k < nz - 2) {
        const int b = o + sxy;
        double ag = (Gd[b] + Gd[b + 1] + Gd[b + sx] + Gd[b + sx + 1]) * 0.25;
        double am = (Mu[b] + Mu[b + 1] + Mu[b + sx] + Mu[b + sx + 1]) * 0.25;
        double e = 0.5 * (dfd<MUL>(Vx[b + sx + 1] - Vx[b + 1], c.dy, c._dy) +
                          dfd<MUL>(Vy[b + sx + 1] - Vy[b + sx], c.dx, c._dx));
        double to = VE ? f.p[PTS_TXY_O][o] : 0.0;
        f.p[PTS_TXY][o] = tau_upd<VE>(f.p[PTS_TXY][o], to, ag, am, c.Gdt, e);
    }
    if (i < nx - 1 && j < ny - 2 && k < nz - 1) {
        const int b = o + sx;
        double ag = (Gd[b] + Gd[b + 1] + Gd[b + sxy] + Gd[b + sxy + 1]) * 0.25;
        double am = (Mu[b] + Mu[b + 1] + Mu[b + sxy] + Mu[b + sxy + 1]) * 0.25;
        double e = 0.5 * (dfd<MUL>(Vx[b + sxy + 1] - Vx[b + 1], c.dz, c._dz) +
                          dfd<MUL>(Vz[b + sxy + 1] - Vz[b + sxy], c.dx, c._dx));
        double to = VE ? f.p[PTS_TXZ_O][o] : 0.0;
        f.p[PTS_TXZ][o] = tau_upd<VE>(f.p[PTS_TXZ][o], to, ag, am, c.Gdt, e);
    }
    if (i < nx - 2 && j < ny - 1 && k < nz - 1) {
        const int b = o + 1;
        double ag = (Gd[b] + Gd[b + sx] + Gd[b + sxy] + Gd[b + sx + sxy]) * 0.25;
        double am = (Mu[b] + Mu[b + sx] + Mu[b + sxy] + Mu[b + sx + sxy]) * 0.25;
        double e = 0.5 * (dfd<MUL>(Vy[b + sx + sxy] - Vy[b + sx], c.dz, c._dz) +
                          dfd<MUL>(Vz[b + sx + sxy] - Vz[b + sxy], c.dy, c._dy));
        double to = VE ? f.p[PTS_TYZ_O][o] : 0.0;
        f.p[PTS_TYZ][o] = tau_upd<VE>(f.p[PTS_TYZ][o], to, ag, am, c.Gdt, e);
    }
}

// compute_τ! (multixpu:63-77, divide form)
template <bool VE>
__global__ __launch_bounds__(BX* BY) void k3_tau(Fields f, StepScalars c) {
    IJK;
    tau3_all<VE, false>(f, c, i, j, k, o);
}

// compute_dV! (multixpu:79-84): dV = R * av(dτ_Rho)
__global__ __launch_bounds__(BX* BY) void k3_dV(Fields f, StepScalars c, int bf) {
    IJK;
    const double* Rh = f.p[PTS_DTAU_RHO];
    if (i < nx - 1 && j < ny - 2 && k < nz - 2)
        f.p[PTS_DVX][o] = res3_x<false>(f, c, o) * ((Rh[o + sx + sxy] + Rh[o + 1 + sx + sxy]) * 0.5);
    if (i < nx - 2 && j < ny - 1 && k < nz - 2)
        f.p[PTS_DVY][o] = res3_y<false>(f, c, o) * ((Rh[o + 1 + sxy] + Rh[o + 1 + sx + sxy]) * 0.5);
    if (i < nx - 2 && j < ny - 2 && k < nz - 1)
        f.p[PTS_DVZ][o] = res3_z<false>(f, c, o, bf != 0) * ((Rh[o + 1 + sx] + Rh[o + 1 + sx + sxy]) * 0.5);
}

// compute_V! (multixpu:86-91): V[inn] += dV
__global__ __launch_bounds__(BX* BY) void k3_V(Fields f) {
    IJK;
    const int q = o + 1 + sx + sxy;
    if (i < nx - 1 && j < ny - 2 && k < nz - 2) f.p[PTS_VX][q] = f.p[PTS_VX][q] + f.p[PTS_DVX][o];
    if (i < nx - 2 && j < ny - 1 && k < nz - 2) f.p[PTS_VY][q] = f.p[PTS_VY][q] + f.p[PTS_DVY][o];
    if (i < nx - 2 && j < ny - 2 && k < nz - 1) f.p[PTS_VZ][q] = f.p[PTS_VZ][q] + f.p[PTS_DVZ][o];
}

// compute_Res! (multixpu:93-98 divide: R only; perf:90-96 multiply: ∇V and R)
template <bool MUL>
__global__ __launch_bounds__(BX* BY) void k3_Res(Fields f, StepScalars c, int bf) {
    IJK;
    if (MUL && i < nx && j < ny && k < nz) f.p[PTS_DIVV][o] = divv3<true>(f, c, o);
    if (i < nx - 1 && j < ny - 2 && k < nz - 2) f.p[PTS_RX][o] = res3_x<MUL>(f, c, o);
    if (i < nx - 2 && j < ny - 1 && k < nz - 2) f.p[PTS_RY][o] = res3_y<MUL>(f, c, o);
    if (i < nx - 2 && j < ny - 2 && k < nz - 1) f.p[PTS_RZ][o] = res3_z<MUL>(f, c, o, bf != 0);
}

// compute_Pt_τ! (perf:69-78; with MUL=false the same update in divide form,
// which equals compute_P! followed by compute_τ! of the 4-kernel script)
template <bool VE, bool MUL>
__global__ __launch_bounds__(BX* BY) void k3_Pt_tau(Fields f, StepScalars c) {
    IJK;
    if (i < nx && j < ny && k < nz)
        f.p[PTS_PT][o] = f.p[PTS_PT][o] - c.r * f.p[PTS_GDTAU][o] * divv3<MUL>(f, c, o);
    tau3_all<VE, MUL>(f, c, i, j, k, o);
}

// compute_V! of the perf script (perf:83-88).  Thread (I,J,K) indexes the
// velocity arrays directly.  With FOLD the free-slip copies of the six BC
// kernels are written by the thread that computes their source value;
// threads on Dirichlet faces copy their own (never updated) value.
template <bool MUL, bool FOLD>
__global__ __launch_bounds__(BX* BY) void k3_V_perf(Fields f, StepScalars c, int bf) {
    const int I = blockIdx.x * BX + threadIdx.x;
    const int J = blockIdx.y * BY + threadIdx.y;
    const int K = blockIdx.z;
    const int nx = f.nx, ny = f.ny, nz = f.nz, sx = f.sx, sxy = f.sxy;
    const int q = I + sx * J + sxy * K;
    const double* Rh = f.p[PTS_DTAU_RHO];
    // Vx: (nx+1) x ny x nz; interior I in [1,nx-1], J in [1,ny-2], K in [1,nz-2]
    if (I <= nx && J >= 1 && J <= ny - 2 && K >= 1 && K <= nz - 2) {
        double* V = f.p[PTS_VX];
        if (I >= 1 && I <= nx - 1) {
            const int o = q - 1 - sx - sxy;
            double v = V[q] + res3_x<MUL>(f, c, o) * ((Rh[o + sx + sxy] + Rh[o + 1 + sx + sxy]) * 0.5);
            if (FOLD) put_bc(V, q, v, J, ny, sx, K, nz, sxy);
            else V[q] = v;
        } else if (FOLD) {
            put_bc(V, q, V[q], J, ny, sx, K, nz, sxy);
        }
    }
    // Vy: nx x (ny+1) x nz; interior I in [1,nx-2], J in [1,ny-1], K in [1,nz-2]
    if (J <= ny && I >= 1 && I <= nx - 2 && K >= 1 && K <= nz - 2) {
        double* V = f.p[PTS_VY];
        if (J >= 1 && J <= ny - 1) {
            const int o = q - 1 - sx - sxy;
            double v = V[q] + res3_y<MUL>(f, c, o) * ((Rh[o + 1 + sxy] + Rh[o + 1 + sx + sxy]) * 0.5);
            if (FOLD) put_bc(V, q, v, I, nx, 1, K, nz, sxy);
            else V[q] = v;
        } else if (FOLD) {
            put_bc(V, q, V[q], I, nx, 1, K, nz, sxy);
        }
    }
    // Vz: nx x ny x (nz+1); interior I in [1,nx-2], J in [1,ny-2], K in [1,nz-1]
    if (K <= nz && I >= 1 && I <= nx - 2 && J >= 1 && J <= ny - 2) {
        double* V = f.p[PTS_VZ];
        if (K >= 1 && K <= nz - 1) {
            const int o = q - 1 - sx - sxy;
            double v = V[q] + res3_z<MUL>(f, c, o, bf != 0) * ((Rh[o + 1 + sx] + Rh[o + 1 + sx + sxy]) * 0.5);
            if (FOLD) put_bc(V, q, v, I, nx, 1, J, ny, sx);
            else V[q] = v;
        } else if (FOLD) {
            put_bc(V, q, V[q], I, nx, 1, J, ny, sx);
        }
    }
}

}  // namespace

hipError_t launch3_compute_P(hipStream_t s, const Fields& f, const StepScalars& c) {
    k3_P<<<grid3(f.nx, f.ny, f.nz), dim3(BX, BY), 0, s>>>(f, c);
    return hipGetLastError();
}
hipError_t launch3_compute_tau(hipStream_t s, const Fields& f, const StepScalars& c, bool ve) {
    if (ve) k3_tau<true><<<grid3(f.nx, f.ny, f.nz), dim3(BX, BY), 0, s>>>(f, c);
    else k3_tau<false><<<grid3(f.nx, f.ny, f.nz), dim3(BX, BY), 0, s>>>(f, c);
    return hipGetLastError();
}
hipError_t launch3_compute_dV(hipStream_t s, const Fields& f, const StepScalars& c, bool bf) {
    k3_dV<<<grid3(f.nx, f.ny, f.nz), dim3(BX, BY), 0, s>>>(f, c, bf ? 1 : 0);
    return hipGetLastError();
}
hipError_t launch3_compute_V(hipStream_t s, const Fields& f) {
    k3_V<<<grid3(f.nx, f.ny, f.nz), dim3(BX, BY), 0, s>>>(f);
    return hipGetLastError();
}
hipError_t launch3_compute_Res(hipStream_t s, const Fields& f, const StepScalars& c, bool bf, bool mul) {
    if (mul) k3_Res<true><<<grid3(f.nx, f.ny, f.nz), dim3(BX, BY), 0, s>>>(f, c, bf ? 1 : 0);
    else k3_Res<false><<<grid3(f.nx, f.ny, f.nz), dim3(BX, BY), 0, s>>>(f, c, bf ? 1 : 0);
    return hipGetLastError();
}
hipError_t launch3_compute_Pt_tau(hipStream_t s, const Fields& f, const StepScalars& c, bool ve, bool mul) {
    const dim3 g = grid3(f.nx, f.ny, f.nz), b(BX, BY);
    if (ve) { if (mul) k3_Pt_tau<true, true><<<g, b, 0, s>>>(f, c); else k3_Pt_tau<true, false><<<g, b, 0, s>>>(f, c); }
    else { if (mul) k3_Pt_tau<false, true><<<g, b, 0, s>>>(f, c); else k3_Pt_tau<false, false><<<g, b, 0, s>>>(f, c); }
    return hipGetLastError();
}
hipError_t launch3_compute_V_perf(hipStream_t s, const Fields& f, const StepScalars& c, bool bf, bool mul,
                                  bool fold_bc) {
    const dim3 g = grid3(f.nx + 1, f.ny + 1, f.nz + 1), b(BX, BY);
    const int ibf = bf ? 1 : 0;
    if (mul) { if (fold_bc) k3_V_perf<true, true><<<g, b, 0, s>>>(f, c, ibf); else k3_V_perf<true, false><<<g, b, 0, s>>>(f, c, ibf); }
    else { if (fold_bc) k3_V_perf<false, true><<<g, b, 0, s>>>(f, c, ibf); else k3_V_perf<false, false><<<g, b, 0, s>>>(f, c, ibf); }
    return hipGetLastError();
}

}  // namespace pts
