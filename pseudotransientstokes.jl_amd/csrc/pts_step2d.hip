// pts_step2d.hip -- 2D PT iteration kernels.
//
// Per-kernel reference API (in place, one launch per reference kernel) and
// the single-pass iteration k2_iter used by pts_pt_iterate/pts_pt_solve.
// Index convention: 0-based (i,j); every field at offset i + sx*j.
#include <algorithm>

#include "pts_internal.hpp"

namespace pts {

namespace {

constexpr int BX = 64, BY = 4;

inline dim3 grid2(int n1, int n2) { return dim3((n1 + BX - 1) / BX, (n2 + BY - 1) / BY, 1); }

#define IJ                                          \
    const int i = blockIdx.x * BX + threadIdx.x;     \
    const int j = blockIdx.y * BY + threadIdx.y;     \
    const int nx = f.nx, ny = f.ny, sx = f.sx;       \
    const int o = i + sx * j;                        \
    (void)nx; (void)ny; (void)o;

// compute_P! (Stokes2D.jl:33-37, Stokes2D_ve.jl:42-46)
__global__ __launch_bounds__(BX* BY) void k2_P(Fields f, StepScalars c) {
    IJ;
    if (i < nx && j < ny) {
        const double* Vx = f.p[PTS_VX];
        const double* Vy = f.p[PTS_VY];
        double d = (Vx[o + 1] - Vx[o]) / c.dx + (Vy[o + sx] - Vy[o]) / c.dy;
        f.p[PTS_DIVV][o] = d;
        f.p[PTS_PT][o] = f.p[PTS_PT][o] - c.r * f.p[PTS_GDTAU][o] * d;
    }
}

// compute_τ! (Stokes2D.jl:39-44 viscous, Stokes2D_ve.jl:48-55 VE)
template <bool VE>
__global__ __launch_bounds__(BX* BY) void k2_tau(Fields f, StepScalars c) {
    IJ;
    const double* Vx = f.p[PTS_VX];
    const double* Vy = f.p[PTS_VY];
    const double* Gd = f.p[PTS_GDTAU];
    const double* Mu = f.p[PTS_MUS];
    if (i < nx && j < ny) {
        double g = Gd[o], m = Mu[o];
        double ex = Vx[o + 1] - Vx[o], ey = Vy[o + sx] - Vy[o];
        double* txx = f.p[PTS_TXX];
        double* tyy = f.p[PTS_TYY];
        if constexpr (VE) {
            txx[o] = tau_upd<true>(txx[o], f.p[PTS_TXX_O][o], g, m, c.Gdt, ex / c.dx);
            tyy[o] = tau_upd<true>(tyy[o], f.p[PTS_TYY_O][o], g, m, c.Gdt, ey / c.dy);
        } else {
            txx[o] = tau_visc2d_div(txx[o], g, m, ex, c.dx);
            tyy[o] = tau_visc2d_div(tyy[o], g, m, ey, c.dy);
        }
    }
    if (i < nx - 1 && j < ny - 1) {
        double ag = (Gd[o] + Gd[o + 1] + Gd[o + sx] + Gd[o + sx + 1]) * 0.25;
        double am = (Mu[o] + Mu[o + 1] + Mu[o + sx] + Mu[o + sx + 1]) * 0.25;
        double e = 0.5 * ((Vx[o + sx + 1] - Vx[o + 1]) / c.dy + (Vy[o + sx + 1] - Vy[o + sx]) / c.dx);
        double* txy = f.p[PTS_TXY];
        double to = VE ? f.p[PTS_TXY_O][o] : 0.0;
        txy[o] = tau_upd<VE>(txy[o], to, ag, am, c.Gdt, e);
    }
}

// Momentum residuals at (i,j) in either flavour (Stokes2D_ve.jl:70-71,
// Stokes2D_ve_perf.jl:63-64), body force per Stokes2D_param.jl:59.
template <bool MUL>
__device__ __forceinline__ double res2_x(const Fields& f, const StepScalars& c, int o) {
    const int sx = f.sx;
    const double* txx = f.p[PTS_TXX];
    const double* txy = f.p[PTS_TXY];
    const double* Pt = f.p[PTS_PT];
    return dfd<MUL>(txx[o + sx + 1] - txx[o + sx], c.dx, c._dx) + dfd<MUL>(txy[o + sx] - txy[o], c.dy, c._dy) -
           dfd<MUL>(Pt[o + sx + 1] - Pt[o + sx], c.dx, c._dx);
}
template <bool MUL>
__device__ __forceinline__ double res2_y(const Fields& f, const StepScalars& c, int o, bool bf) {
    const int sx = f.sx;
    const double* tyy = f.p[PTS_TYY];
    const double* txy = f.p[PTS_TXY];
    const double* Pt = f.p[PTS_PT];
    double R = dfd<MUL>(tyy[o + sx + 1] - tyy[o + 1], c.dy, c._dy) + dfd<MUL>(txy[o + 1] - txy[o], c.dx, c._dx) -
               dfd<MUL>(Pt[o + sx + 1] - Pt[o + 1], c.dy, c._dy);
    if (bf) {
        const double* rg = f.p[PTS_RHOG];
        R = R - (rg[o + 1] + rg[o + sx + 1]) * 0.5;
    }
    return R;
}

// compute_dV! (Stokes2D.jl:46-52 stores Rx, Ry as well; Stokes2D_ve.jl:57-61)
__global__ __launch_bounds__(BX* BY) void k2_dV(Fields f, StepScalars c, int bf) {
    IJ;
    const double* Rh = f.p[PTS_DTAU_RHO];
    if (i < nx - 1 && j < ny - 2) {
        double R = res2_x<false>(f, c, o);
        f.p[PTS_RX][o] = R;
        f.p[PTS_DVX][o] = R * ((Rh[o + sx] + Rh[o + sx + 1]) * 0.5);
    }
    if (i < nx - 2 && j < ny - 1) {
        double R = res2_y<false>(f, c, o, bf != 0);
        f.p[PTS_RY][o] = R;
        f.p[PTS_DVY][o] = R * ((Rh[o + 1] + Rh[o + sx + 1]) * 0.5);
    }
}

// compute_V! (Stokes2D.jl:54-58)
__global__ __launch_bounds__(BX* BY) void k2_V(Fields f) {
    IJ;
    if (i < nx - 1 && j < ny - 2) f.p[PTS_VX][o + sx + 1] = f.p[PTS_VX][o + sx + 1] + f.p[PTS_DVX][o];
    if (i < nx - 2 && j < ny - 1) f.p[PTS_VY][o + sx + 1] = f.p[PTS_VY][o + sx + 1] + f.p[PTS_DVY][o];
}

// compute_Res! (Stokes2D_ve.jl:69-73 divide form: Rx, Ry;
// Stokes2D_ve_perf.jl:61-66 multiply form: ∇V, Rx, Ry)
template <bool MUL>
__global__ __launch_bounds__(BX* BY) void k2_Res(Fields f, StepScalars c, int bf) {
    IJ;
    if (MUL && i < nx && j < ny) {
        const double* Vx = f.p[PTS_VX];
        const double* Vy = f.p[PTS_VY];
        f.p[PTS_DIVV][o] = c._dx * (Vx[o + 1] - Vx[o]) + c._dy * (Vy[o + sx] - Vy[o]);
    }
    if (i < nx - 1 && j < ny - 2) f.p[PTS_RX][o] = res2_x<MUL>(f, c, o);
    if (i < nx - 2 && j < ny - 1) f.p[PTS_RY][o] = res2_y<MUL>(f, c, o, bf != 0);
}

// compute_Pt_τ! (Stokes2D_ve_perf.jl:47-53)
template <bool VE>
__global__ __launch_bounds__(BX* BY) void k2_Pt_tau(Fields f, StepScalars c) {
    IJ;
    const double* Vx = f.p[PTS_VX];
    const double* Vy = f.p[PTS_VY];
    const double* Gd = f.p[PTS_GDTAU];
    const double* Mu = f.p[PTS_MUS];
    if (i < nx && j < ny) {
        double g = Gd[o], m = Mu[o];
        double ex = c._dx * (Vx[o + 1] - Vx[o]), ey = c._dy * (Vy[o + sx] - Vy[o]);
        f.p[PTS_PT][o] = f.p[PTS_PT][o] - c.r * g * (ex + ey);
        double* txx = f.p[PTS_TXX];
        double* tyy = f.p[PTS_TYY];
        double txo = VE ? f.p[PTS_TXX_O][o] : 0.0, tyo = VE ? f.p[PTS_TYY_O][o] : 0.0;
        txx[o] = tau_upd<VE>(txx[o], txo, g, m, c.Gdt, ex);
        tyy[o] = tau_upd<VE>(tyy[o], tyo, g, m, c.Gdt, ey);
    }
    if (i < nx - 1 && j < ny - 1) {
        double ag = 0.25 * (Gd[o] + Gd[o + 1] + Gd[o + sx] + Gd[o + sx + 1]);
        double am = 0.25 * (Mu[o] + Mu[o + 1] + Mu[o + sx] + Mu[o + sx + 1]);
        double e = 0.5 * (c._dy * (Vx[o + sx + 1] - Vx[o + 1]) + c._dx * (Vy[o + sx + 1] - Vy[o + sx]));
        double* txy = f.p[PTS_TXY];
        double to = VE ? f.p[PTS_TXY_O][o] : 0.0;
        txy[o] = tau_upd<VE>(txy[o], to, ag, am, c.Gdt, e);
    }
}

// compute_V! of the perf script (Stokes2D_ve_perf.jl:55-59)
__global__ __launch_bounds__(BX* BY) void k2_V_perf(Fields f, StepScalars c, int bf) {
    IJ;
    const double* Rh = f.p[PTS_DTAU_RHO];
    if (i < nx - 1 && j < ny - 2) {
        double R = res2_x<true>(f, c, o);
        f.p[PTS_VX][o + sx + 1] = f.p[PTS_VX][o + sx + 1] + R * (0.5 * (Rh[o + sx] + Rh[o + sx + 1]));
    }
    if (i < nx - 2 && j < ny - 1) {
        double R = res2_y<true>(f, c, o, bf != 0);
        f.p[PTS_VY][o + sx + 1] = f.p[PTS_VY][o + sx + 1] + R * (0.5 * (Rh[o + 1] + Rh[o + sx + 1]));
    }
}

// ---------------------------------------------------------------------------
// Single-pass iteration.  One 64-lane wave owns a strip of 62 output columns
// (lane t handles column c = 62*strip - 1 + t; lanes 0 and 63 are halo) and a
// chunk of rows, marching in y.  At row step J a lane computes
//   * the new cell values Pt, τxx, τyy at (c, J)            (compute_Pt_τ!)
//   * the new vertex τxy at (c-1, J-1)
//   * the new Vx at face (c, J-1) and Vy at face (c, J)     (compute_V!)
// with free-slip copies folded in (bc_y!(Vx), bc_x!(Vy)).  Neighbour values
// across x come from DPP lane shifts; values from the previous row are kept
// in registers.  The state is read from `in` and written to `out`, so every
// array is streamed once per iteration.  Dirichlet (never-updated) entries of
// Vx/Vy are not written: both ping-pong buffers hold them.
// ---------------------------------------------------------------------------
// Inputs of one row step (cell row J, vertex row J-1).
struct Row2 {
    double Pt, txx, tyy, txo, tyo, Gd, Mu, Rh, VyN, rg, Vx, txy, txyo;
};

// Musτ of cell (cx, J) from μ: compute_maxloc! (the 5-point cross maximum,
// Stokes2D_ve.jl:23-26, in the order of k_maxloc) at the cell clamped into
// the interior, which is what bc_x!, bc_y! leave on the boundary cells
// (Stokes2D_ve.jl:142-145).
__device__ __forceinline__ double mustau_of(const Fields& in, int cx, int J) {
    const int cc = min(max(cx, 1), in.nx - 2), jj = min(max(J, 1), in.ny - 2);
    const int om = cc + in.sx * jj;
    const double* M = in.p[PTS_MUS];
    return fmax(fmax(fmax(fmax(M[om - 1], M[om + 1]), M[om]), M[om - in.sx]), M[om + in.sx]);
}

template <bool VE, bool BF, bool DP, int MX>
__device__ __forceinline__ void load_row2(Row2& w, const Fields& in, const StepScalars& c, int o, int sx, bool cin,
                                          bool fin, bool vtx, int cx, int J) {
    w = Row2{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (cin) {
        w.Pt = in.p[PTS_PT][o];
        w.txx = in.p[PTS_TXX][o];
        w.tyy = in.p[PTS_TYY][o];
        if constexpr (VE) {
            w.txo = in.p[PTS_TXX_O][o];
            w.tyo = in.p[PTS_TYY_O][o];
        }
        if constexpr (DP) {
            iter_params_of(MX ? mustau_of(in, cx, J) : in.p[PTS_MUSTAU][o], c, w.Rh, w.Gd);
        } else {
            w.Gd = in.p[PTS_GDTAU][o];
            w.Rh = in.p[PTS_DTAU_RHO][o];
        }
        w.Mu = in.p[PTS_MUS][o];
        w.VyN = in.p[PTS_VY][o + sx];
        if constexpr (BF) w.rg = in.p[PTS_RHOG][o];
    }
    if (fin) w.Vx = in.p[PTS_VX][o];
    if (vtx) {
        w.txy = in.p[PTS_TXY][o - 1 - sx];
        if constexpr (VE) w.txyo = in.p[PTS_TXY_O][o - 1 - sx];
    }
}

// Branch-free variant: every lane loads from an in-bounds (clamped) address,
// so the number of loads in flight is static and the compiler can wait for
// exactly the row it needs.  Lanes outside the grid then hold copies of edge
// values; no stored output depends on them (each output's operands come from
// lanes and rows inside its own validity range).
// MX == 2: lanes with `mxload` (boundary columns and the strip's halo lanes)
// fetch Musτ itself with the row, into Gd (it equals the cross maximum: the
// MX kernels only run after a device check of that).
template <bool VE, bool BF, bool DP, int MX>
__device__ __forceinline__ void load_row2_clamped(Row2& w, const Fields& in, const StepScalars& c, int cx, int J,
                                                  int nx, int sx, bool mxload = false) {
    const int oc = min(max(cx, 0), nx - 1) + sx * J;
    const int of = min(max(cx, 0), nx) + sx * J;
    const int ov = min(max(cx - 1, 0), nx - 2) + sx * max(J - 1, 0);
    w.Pt = in.p[PTS_PT][oc];
    w.txx = in.p[PTS_TXX][oc];
    w.tyy = in.p[PTS_TYY][oc];
    w.txo = w.tyo = w.txyo = w.rg = 0.0;
    if constexpr (VE) {
        w.txo = in.p[PTS_TXX_O][oc];
        w.tyo = in.p[PTS_TYY_O][oc];
        w.txyo = in.p[PTS_TXY_O][ov];
    }
    if constexpr (DP && MX == 2) {
        w.Gd = w.Rh = 0.0;  // derived from μ in the row step
        if (mxload) w.Gd = in.p[PTS_MUSTAU][oc];
    } else if constexpr (DP) {
        iter_params_of(MX ? mustau_of(in, cx, J) : in.p[PTS_MUSTAU][oc], c, w.Rh, w.Gd);
    } else {
        w.Gd = in.p[PTS_GDTAU][oc];
        w.Rh = in.p[PTS_DTAU_RHO][oc];
    }
    w.Mu = in.p[PTS_MUS][oc];
    w.VyN = in.p[PTS_VY][oc + sx];
    if constexpr (BF) w.rg = in.p[PTS_RHOG][oc];
    w.Vx = in.p[PTS_VX][of];
    w.txy = in.p[PTS_TXY][ov];
}

template <bool NT>
__device__ __forceinline__ void st(double* p, double v) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// ---------------------------------------------------------------------------
// Single-pass iteration.  One 64-lane wave owns a strip of 62 output columns
// (lane t handles column c = 62*strip - 1 + t; lanes 0 and 63 are halo) and a
// chunk of rows, marching in y.  At row step J a lane computes
//   * the new cell values Pt, τxx, τyy at (c, J)            (compute_Pt_τ!)
//   * the new vertex τxy at (c-1, J-1)
//   * the new Vx at face (c, J-1) and Vy at face (c, J)     (compute_V!)
// with free-slip copies folded in (bc_y!(Vx), bc_x!(Vy)).  Neighbour values
// across x come from DPP lane shifts; values from the previous row are kept
// in registers.  The state is read from `in` and written to `out`, so every
// array is streamed once per iteration.  Dirichlet (never-updated) entries of
// Vx/Vy are not written: both ping-pong buffers hold them.  With PF the loads
// of row J+1 are issued before row J is computed; with NT the stores bypass
// the caches (the outputs are not read again in this launch).
// ---------------------------------------------------------------------------
//
// NW > 1: a block of NW waves owns a strip of 64·NW − 2 columns (only its
// first and last lane are halo), so a row of each array is read as one
// contiguous run of 64·NW values; values crossing the wave boundaries inside
// the block go through LDS (two barriers per row step).
//
// MX: Musτ derived from μ (Gdτ, dτ_Rho from Musτ need DP).  MX = 1 loads the
// five μ of each cell's cross; MX = 2 (with PF) takes them from the row
// pipeline: x-neighbours by DPP (wave edges: exchanged one row ahead through
// LDS), the row below from registers, the row above from the prefetched row;
// rows whose cross is not in the pipeline (first and last row of a chunk,
// grid row 0), the boundary columns and the strip's two halo lanes fall
// back to the loads.
// Grid-wide barrier of a cooperative launch (every block resident): bar[0]
// counts the arrivals, bar[1] is the generation (it only grows, so a launch
// starts from whatever value the previous one left).  The writes before the
// barrier are released and the reads after it acquired at device scope (the
// XCDs' L2 caches are not coherent with each other).  The wait is bounded:
// a block that sees no new generation after about 2^22 polls sets *err and
// goes on, so a broken barrier ends the kernel with wrong results and an
// error the host reports, never a hang.
__device__ __forceinline__ unsigned barrier_generation(const unsigned* bar) {
    return __hip_atomic_load(bar + 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ void grid_barrier(unsigned* bar, unsigned nblocks, unsigned& gen, int* err) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned g = gen;
        const unsigned a = __hip_atomic_fetch_add(bar, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (a == nblocks - 1) {
            __hip_atomic_store(bar, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(bar + 1, g + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            for (int spins = 0; __hip_atomic_load(bar + 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g;) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > (1 << 22)) {
                    __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        gen = g + 1u;
    }
    __syncthreads();
}

// MULTI: one cooperative launch runs `niter` iterations, the ping-pong
// buffers swapping with the parity of the iteration, a grid barrier between
// iterations (small grids, where a launch costs more than the iteration).
template <bool VE, bool MUL, bool BF, bool PF, bool NT, int NW, bool DP, int MX, bool MULTI = false>
__global__ __launch_bounds__(64 * NW, (!PF && NW == 8) ? 6 : 1) void k2_iter(const Fields f0, const Fields f1, const StepScalars c, int L,
                                                   int nstrips, int ntiles, int niter, unsigned* bar, int* berr) {
    constexpr int W = 64 * NW - 2;  // output columns per strip
    constexpr int NXL = 7;          // values passed to the right neighbour wave
    __shared__ double sL[NW > 1 ? NW : 1][NXL], sR[NW > 1 ? NW : 1][2], sT[NW > 1 ? NW : 1];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, gl = threadIdx.x;
    const int nx = f0.nx, ny = f0.ny, sx = f0.sx;
    const double dx = c.dx, dy = c.dy, _dx = c._dx, _dy = c._dy, r = c.r, Gdt = c.Gdt;
    const int nb = gridDim.x;
    unsigned gen = 0;
    if constexpr (MULTI) {
        if (threadIdx.x == 0) gen = barrier_generation(bar);
    }

    for (int it = 0; it < (MULTI ? niter : 1); ++it) {
    const Fields& in = (MULTI && (it & 1)) ? f1 : f0;
    const Fields& out = (MULTI && (it & 1)) ? f0 : f1;
    double* __restrict__ Pt1 = out.p[PTS_PT];
    double* __restrict__ txx1 = out.p[PTS_TXX];
    double* __restrict__ tyy1 = out.p[PTS_TYY];
    double* __restrict__ txy1 = out.p[PTS_TXY];
    double* __restrict__ Vx1 = out.p[PTS_VX];
    double* __restrict__ Vy1 = out.p[PTS_VY];

    for (int b0 = blockIdx.x; b0 < ntiles; b0 += nb) {
        int b = b0;
        // XCD-aware order: the hardware deals consecutive blocks to the 8 XCDs
        // round-robin; remap so each XCD gets a contiguous run of tiles and the
        // halo columns shared by neighbouring strips stay in one L2.
        if (nb >= ntiles) b = xcd_tile(b, ntiles);
        const int strip = b % nstrips, chunk = b / nstrips;
        const int cx = strip * W - 1 + gl;
        const int jb = chunk * L;
        const int je = min(jb + L, ny);
        const bool own = gl >= 1 && gl <= W;
        const bool cin = cx >= 0 && cx < nx;       // cell column
        const bool fin = cx >= 0 && cx <= nx;      // Vx face
        const bool vin = cx >= 1 && cx <= nx - 1;  // vertex column cx-1 / interior Vx face
        const bool vyin = cx >= 1 && cx <= nx - 2; // interior Vy column
        const bool mxload = cx == 0 || cx == nx - 1 || gl == 0 || gl == 64 * NW - 1;

        // values carried from row J-1
        double VxP = 0.0, GdP = 0.0, MuP = 0.0, GdPl = 0.0, MuPl = 0.0, RhP = 0.0, rgP = 0.0;
        double Pt1P = 0.0, txx1P = 0.0, tyy1P = 0.0, txy1PP = 0.0;
        double MtP = 0.0, muLn = 0.0, muRn = 0.0;  // MX == 2: Musτ of row J-1; edge μ of row J
        const int j0 = max(jb - 1, 0), j1 = min(je, ny - 1);
        double VyC = in.p[PTS_VY][min(max(cx, 0), nx - 1) + sx * j0];  // Vy0[c, J]

        auto step = [&](const int J, Row2 cur, const double nxtMu) {
            const int o = cx + sx * J;
            const bool vtx = vin && J >= 1;
            if constexpr (DP && MX == 2) {
                double muL = from_left(cur.Mu), muR = from_right(cur.Mu);
                if (NW > 1 && lane == 0 && wv > 0) muL = muLn;
                if (NW > 1 && lane == 63 && wv < NW - 1) muR = muRn;
                double Mt;
                if (J == 0 || J == j0 || (J == j1 && J < ny - 1)) {
                    Mt = mustau_of(in, cx, J);
                } else if (J == ny - 1) {
                    Mt = MtP;  // Musτ(:, ny-1) = Musτ(:, ny-2)
                } else {
                    Mt = fmax(fmax(fmax(fmax(muL, muR), cur.Mu), MuP), nxtMu);
                    // clamped boundary columns, and the strip's two halo lanes, which
                    // have no x-neighbour in the block (their Gdτ feeds the
                    // vertices of the first and last own lanes): Musτ as loaded
                    if (mxload) Mt = cur.Gd;
                }
                MtP = Mt;
                iter_params_of(Mt, c, cur.Rh, cur.Gd);
            }
            // lane exchanges (whole wave active)
            double VxR = from_right(cur.Vx);
            double GdL = from_left(cur.Gd);
            double MuL = from_left(cur.Mu);
            double VyCL = from_left(VyC);
            double txx1PL = from_left(txx1P);
            double Pt1PL = from_left(Pt1P);
            double RhPL = from_left(RhP);
            if constexpr (NW > 1) {  // across the wave boundaries of the block
                if (lane == 63) {
                    sL[wv][0] = cur.Gd;
                    sL[wv][1] = cur.Mu;
                    sL[wv][2] = VyC;
                    sL[wv][3] = txx1P;
                    sL[wv][4] = Pt1P;
                    sL[wv][5] = RhP;
                    sL[wv][6] = nxtMu;
                }
                if (lane == 0) {
                    sR[wv][0] = cur.Vx;
                    sR[wv][1] = nxtMu;
                }
                __syncthreads();
                if (lane == 0 && wv > 0) {
                    GdL = sL[wv - 1][0];
                    MuL = sL[wv - 1][1];
                    VyCL = sL[wv - 1][2];
                    txx1PL = sL[wv - 1][3];
                    Pt1PL = sL[wv - 1][4];
                    RhPL = sL[wv - 1][5];
                    muLn = sL[wv - 1][6];
                }
                if (lane == 63 && wv < NW - 1) {
                    VxR = sR[wv + 1][0];
                    muRn = sR[wv + 1][1];
                }
            }

            // cell (c, J)
            const double dVx = VxR - cur.Vx, dVy = cur.VyN - VyC;
            const double ex = dfd<MUL>(dVx, dx, _dx), ey = dfd<MUL>(dVy, dy, _dy);
            const double Pt1n = cur.Pt - r * cur.Gd * (ex + ey);
            double txx1n, tyy1n;
            if constexpr (!VE && !MUL) {
                txx1n = tau_visc2d_div(cur.txx, cur.Gd, cur.Mu, dVx, dx);
                tyy1n = tau_visc2d_div(cur.tyy, cur.Gd, cur.Mu, dVy, dy);
            } else {
                txx1n = tau_upd<VE>(cur.txx, cur.txo, cur.Gd, cur.Mu, Gdt, ex);
                tyy1n = tau_upd<VE>(cur.tyy, cur.tyo, cur.Gd, cur.Mu, Gdt, ey);
            }
            // vertex (c-1, J-1): 4-cell means over [c-1..c] x [J-1..J]
            const double ag = (GdPl + GdP + GdL + cur.Gd) * 0.25;
            const double am = (MuPl + MuP + MuL + cur.Mu) * 0.25;
            const double e = 0.5 * (dfd<MUL>(cur.Vx - VxP, dy, _dy) + dfd<MUL>(VyC - VyCL, dx, _dx));
            const double txy1n = tau_upd<VE>(cur.txy, cur.txyo, ag, am, Gdt, e);
            double txy1nR = from_right(txy1n);
            if constexpr (NW > 1) {
                if (lane == 0) sT[wv] = txy1n;
                __syncthreads();
                if (lane == 63 && wv < NW - 1) txy1nR = sT[wv + 1];
            }

            // stores of the new cell/vertex values owned by this tile
            if (own && cin && J >= jb && J < je) {
                st<NT>(Pt1 + o, Pt1n);
                st<NT>(txx1 + o, txx1n);
                st<NT>(tyy1 + o, tyy1n);
            }
            const int Jm = J - 1;
            if (own && vtx && Jm >= jb && Jm < je) st<NT>(txy1 + o - 1 - sx, txy1n);

            // Vx at face (c, J-1)
            if (own && Jm >= jb && Jm < je && Jm >= 1 && Jm <= ny - 2) {
                if (vin) {
                    double R = dfd<MUL>(txx1P - txx1PL, dx, _dx) + dfd<MUL>(txy1n - txy1PP, dy, _dy) -
                               dfd<MUL>(Pt1P - Pt1PL, dx, _dx);
                    double v = VxP + R * (0.5 * (RhPL + RhP));
                    st<NT>(Vx1 + o - sx, v);
                    if (Jm == 1) Vx1[o - 2 * sx] = v;           // bc_y!: Vx[:,1] = Vx[:,2]
                    if (Jm == ny - 2) Vx1[o] = v;               //        Vx[:,end] = Vx[:,end-1]
                } else if (fin) {                               // Dirichlet faces c = 0, nx
                    if (Jm == 1) Vx1[o - 2 * sx] = VxP;
                    if (Jm == ny - 2) Vx1[o] = VxP;
                }
            }
            // Vy at face (c, J)
            if (own && J >= jb && J < je && J >= 1) {
                if (vyin) {
                    double R = dfd<MUL>(tyy1n - tyy1P, dy, _dy) + dfd<MUL>(txy1nR - txy1n, dx, _dx) -
                               dfd<MUL>(Pt1n - Pt1P, dy, _dy);
                    if constexpr (BF) R = R - (rgP + cur.rg) * 0.5;
                    double v = VyC + R * (0.5 * (RhP + cur.Rh));
                    st<NT>(Vy1 + o, v);
                    if (cx == 1) Vy1[o - 1] = v;                // bc_x!: Vy[1,:] = Vy[2,:]
                    if (cx == nx - 2) Vy1[o + 1] = v;           //        Vy[end,:] = Vy[end-1,:]
                }
            }
            // Dirichlet rows of Vy (face rows 0 and ny): their x-copies
            if (own && (cx == 1 || cx == nx - 2)) {
                if (J == 0 && jb == 0) {
                    if (cx == 1) Vy1[o - 1] = VyC;
                    if (cx == nx - 2) Vy1[o + 1] = VyC;
                }
                if (J == ny - 1 && je == ny) {
                    if (cx == 1) Vy1[o + sx - 1] = cur.VyN;
                    if (cx == nx - 2) Vy1[o + sx + 1] = cur.VyN;
                }
            }
            // carry
            VxP = cur.Vx;
            VyC = cur.VyN;
            GdPl = GdL;
            MuPl = MuL;
            GdP = cur.Gd;
            MuP = cur.Mu;
            RhP = cur.Rh;
            rgP = cur.rg;
            Pt1P = Pt1n;
            txx1P = txx1n;
            tyy1P = tyy1n;
            txy1PP = txy1n;
        };
        Row2 ra, rb;
        if constexpr (PF) {
            load_row2_clamped<VE, BF, DP, MX>(ra, in, c, cx, j0, nx, sx, mxload);
            for (int J = j0; J <= j1; J += 2) {
                if (J + 1 <= j1) load_row2_clamped<VE, BF, DP, MX>(rb, in, c, cx, J + 1, nx, sx, mxload);
                step(J, ra, J + 1 <= j1 ? rb.Mu : ra.Mu);
                if (J + 1 > j1) break;
                if (J + 2 <= j1) load_row2_clamped<VE, BF, DP, MX>(ra, in, c, cx, J + 2, nx, sx, mxload);
                step(J + 1, rb, J + 2 <= j1 ? ra.Mu : rb.Mu);
            }
        } else {
            for (int J = j0; J <= j1; ++J) {
                load_row2<VE, BF, DP, MX>(ra, in, c, cx + sx * J, sx, cin, fin, vin && J >= 1, cx, J);
                step(J, ra, ra.Mu);
            }
        }
    }
    if constexpr (MULTI) {
        if (it + 1 < niter) grid_barrier(bar, (unsigned)nb, gen, berr);
    }
    }
}

// ---------------------------------------------------------------------------
// Small 2D grids: the whole state in LDS for a whole check interval.  The
// positions (i, J), 0 <= i <= nx, 0 <= J <= ny, are dealt in tiles of R rows
// and C columns (the last tile of a row or column of tiles: what is left,
// at least R or C); block b = by·nbx + bx holds its tile of every field plus
// H halo rows and columns on each side, and runs `niter` iterations on them:
// the updates of compute_P!/compute_τ! (cells and vertices, in place) and of
// compute_V! (faces, in place, with the free-slip copies of bc_y!(Vx),
// bc_x!(Vy)), each a phase between barriers.  Every H iterations a block
// hands its positions within H of its edges to the blocks that hold them as
// halo (H may exceed R or C) through a small global buffer (two parities)
// and a per-block counter (release and acquire at device scope; the wait is
// bounded by spin_limit polls and sets *err when it gives up; wskew, a test
// hook, raises the generation waited for); in between it also updates the
// halo positions that are still valid (H - 1 - step beyond its own on each
// side), so only the halo moves through memory: the state is read once at
// the start and written once at the end (to f1 when niter is odd, else to
// f0; in the divide form the velocities before the last iteration go to the
// other one, as the ping-pong leaves them for the err).  A cooperative
// launch keeps every block resident.  Bit-identical to the per-kernel
// sequence.
// ---------------------------------------------------------------------------
// x / h for a grid spacing h > 0, with rh = 1/h rounded to nearest:
// Markstein's correction of the product (P. W. Markstein, "Computation of
// elementary functions on the IBM RISC System/6000 processor", IBM J. Res.
// Dev. 34(1), 1990): q = x·rh, r = h·q − x (exact with the fused
// multiply-add), q − r·rh rounded once is x/h rounded to nearest, signed
// zeros included.  Three instructions instead of the general division's
// ten; checked against IEEE division on 4.2·10^9 numerators for 16 grid
// spacings, among them 3.2·10^9 whose quotients lie next to a midpoint
// between doubles (no difference).
__device__ __forceinline__ double div_h(double x, double h, double rh) {
    const double q = x * rh;
    const double r = __builtin_fma(h, q, -x);
    return __builtin_fma(-r, rh, q);
}
// The scaled difference of the two script flavours: _dx·(a − b) (perf
// scripts, ih the caller's reciprocal) or (a − b)/dx (4-kernel scripts).
template <bool MUL>
__device__ __forceinline__ double dfd_h(double diff, double h, double ih, double rh) {
    if constexpr (MUL) return ih * diff;
    else return div_h(diff, h, rh);
}
// res2_x / res2_y with dfd_h (the same operations and order)
template <bool MUL>
__device__ __forceinline__ double res2_x_h(const Fields& f, const StepScalars& c, int o, double rdx, double rdy) {
    const int sx = f.sx;
    const double* txx = f.p[PTS_TXX];
    const double* txy = f.p[PTS_TXY];
    const double* Pt = f.p[PTS_PT];
    return dfd_h<MUL>(txx[o + sx + 1] - txx[o + sx], c.dx, c._dx, rdx) +
           dfd_h<MUL>(txy[o + sx] - txy[o], c.dy, c._dy, rdy) -
           dfd_h<MUL>(Pt[o + sx + 1] - Pt[o + sx], c.dx, c._dx, rdx);
}
template <bool MUL>
__device__ __forceinline__ double res2_y_h(const Fields& f, const StepScalars& c, int o, bool bf, double rdx,
                                           double rdy) {
    const int sx = f.sx;
    const double* tyy = f.p[PTS_TYY];
    const double* txy = f.p[PTS_TXY];
    const double* Pt = f.p[PTS_PT];
    double R = dfd_h<MUL>(tyy[o + sx + 1] - tyy[o + 1], c.dy, c._dy, rdy) +
               dfd_h<MUL>(txy[o + 1] - txy[o], c.dx, c._dx, rdx) -
               dfd_h<MUL>(Pt[o + sx + 1] - Pt[o + 1], c.dy, c._dy, rdy);
    if (bf) {
        const double* rg = f.p[PTS_RHOG];
        R = R - (rg[o + 1] + rg[o + sx + 1]) * 0.5;
    }
    return R;
}

constexpr int kLdsState = 6;  // Pt, τxx, τyy, τxy, Vx, Vy
__device__ __forceinline__ int lds_state_field(int q) {
    return q == 0 ? PTS_PT : q == 1 ? PTS_TXX : q == 2 ? PTS_TYY : q == 3 ? PTS_TXY : q == 4 ? PTS_VX : PTS_VY;
}

// LDS slot q holds rows L0 = max(r0 - H, 0) .. and columns K0 = max(c0 - H,
// 0) .. of one field, lds2_extent of each (the most any tile holds);
// element (i, J) is at (i - K0) + sl·(J - L0) of the slot, sl the column
// extent (a local index: every pointer stays inside the LDS array).  Slots
// 0-5: the state (lds_state_field order), 6 μ, 7 Gdτ, 8 dτ_Rho, then ρg
// (BF) and τxx_o, τyy_o, τxy_o (VE).
template <bool VE, bool MUL, bool BF, bool DP, int H>
__global__ __launch_bounds__(1024) void k2_lds(const Fields f0, const Fields f1, const StepScalars c, int R, int C,
                                               int nbx, int niter, double* xb, unsigned long long* flags,
                                               unsigned long long gen0, int* err, int spin_limit,
                                               unsigned long long wskew) {
    extern __shared__ double smem[];
    const int nx = f0.nx, ny = f0.ny, sx = f0.sx;
    const int nb = gridDim.x, b = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const int nby = nb / nbx, by = b / nbx, bx = b % nbx;
    // own rows and columns (Vy has faces up to row ny, Vx up to column nx);
    // the last tile of a row or column of tiles owns what is left
    const int r0 = by * R, r1 = by == nby - 1 ? ny + 1 : r0 + R;
    const int c0 = bx * C, c1 = bx == nbx - 1 ? nx + 1 : c0 + C;
    const int RL = ny + 1 - (nby - 1) * R, CL = nx + 1 - (nbx - 1) * C;  // the last tiles' extents
    const int L0 = max(r0 - H, 0), K0 = max(c0 - H, 0);  // LDS row and column 0
    const int sl = lds2_extent(nx + 1, C, CL, H);        // LDS row stride
    const size_t slot = (size_t)lds2_extent(ny + 1, R, RL, H) * sl;
    double* const Pt = smem;
    double* const txx = smem + slot;
    double* const tyy = smem + 2 * slot;
    double* const txy = smem + 3 * slot;
    double* const Vx = smem + 4 * slot;
    double* const Vy = smem + 5 * slot;
    double* const Mu = smem + 6 * slot;
    double* const Gd = smem + 7 * slot;
    double* const Rh = smem + 8 * slot;
    double* const rg = smem + 9 * slot;                  // BF
    double* const txxo = smem + (9 + BF) * slot;         // VE
    double* const tyyo = smem + (10 + BF) * slot;
    double* const txyo = smem + (11 + BF) * slot;
    // the LDS tile as Fields, for the residual functions of the per-kernel path
    Fields lf = f0;
    lf.p[PTS_PT] = Pt;
    lf.p[PTS_TXX] = txx;
    lf.p[PTS_TYY] = tyy;
    lf.p[PTS_TXY] = txy;
    lf.p[PTS_RHOG] = rg;
    lf.sx = sl;
    const int ja = L0, jz = min(r1 - 1 + H, ny);  // rows held
    const int ia = K0, iz = min(c1 - 1 + H, nx);  // columns held
    const int nw = iz - ia + 1, ncol = c1 - c0;
    const double rdx = 1.0 / c.dx, rdy = 1.0 / c.dy;  // divide form: see div_h

    // load the tile (positions outside the grid are never read)
    for (int e = tid; e < (jz - ja + 1) * nw; e += nt) {
        const int J = ja + e / nw, i = ia + e % nw, o = i + sx * J, l = (i - K0) + sl * (J - L0);
#pragma unroll
        for (int q = 0; q < kLdsState; ++q) smem[q * slot + l] = f0.p[lds_state_field(q)][o];
        Mu[l] = f0.p[PTS_MUS][o];
        if constexpr (DP) {
            double rh, gd;
            iter_params_of(f0.p[PTS_MUSTAU][o], c, rh, gd);
            Rh[l] = rh;
            Gd[l] = gd;
        } else {
            Gd[l] = f0.p[PTS_GDTAU][o];
            Rh[l] = f0.p[PTS_DTAU_RHO][o];
        }
        if constexpr (BF) rg[l] = f0.p[PTS_RHOG][o];
        if constexpr (VE) {
            txxo[l] = f0.p[PTS_TXX_O][o];
            tyyo[l] = f0.p[PTS_TYY_O][o];
            txyo[l] = f0.p[PTS_TXY_O][o];
        }
    }
    __syncthreads();

    // published positions: per block RL rows of kLdsState rows of CL values
    const size_t xblk = (size_t)RL * kLdsState * CL;
    for (int it = 0; it < niter; ++it) {
        // step st of an exchange period of H iterations: the rows and columns
        // still valid shrink by one per step on each side (g beyond the own
        // ones); columns mirror rows (Vx is to columns what Vy is to rows)
        const int st = it % H, g = H - 1 - st;
        const int ca = max(r0 - H + st, 0), cz = min(r1 - 1 + g, ny - 1);  // cell rows of the P, τ phase
        const int vz = min(r1 - 1 + g, ny - 2);                            // vertex rows ca .. vz
        const int xz = min(r1 - 1 + g, ny - 2);                            // Vx face rows ya .. xz
        const int ya = max(r0 - g, 1), yz = min(r1 - 1 + g, ny - 1);       // Vy face rows
        const int ba = max(r0 - g, 0), bz = min(r1 - 1 + g, ny);           // rows of the free-slip copies
        const int ka = max(c0 - H + st, 0), kz = min(c1 - 1 + g, nx - 1);  // cell columns
        const int vkz = min(c1 - 1 + g, nx - 2);                           // vertex columns ka .. vkz
        const int fa = max(c0 - g, 1), fz = min(c1 - 1 + g, nx - 1);       // Vx face columns
        const int fyz = min(c1 - 1 + g, nx - 2);                           // Vy face columns fa .. fyz
        const int qa = max(c0 - g, 0), qz = min(c1 - 1 + g, nx);           // columns of the free-slip copies
        if (it > 0 && st == 0) {
            // the halo positions of the previous iteration from the blocks
            // that own them (H may exceed R or C: then from more than one
            // block on a side); the lanes of wave 0 wait on one owner each
            const int tylo = min(ja / R, nby - 1), tyhi = min(jz / R, nby - 1);
            const int txlo = min(ia / C, nbx - 1), txhi = min(iz / C, nbx - 1), ntx = txhi - txlo + 1;
            if (tid < 64) {
                const int k = tid, nbr = (tylo + k / ntx) * nbx + txlo + k % ntx;
                if (k < (tyhi - tylo + 1) * ntx && nbr != b) {
                    for (int spins = 0; __hip_atomic_load(flags + nbr, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) <
                                        gen0 + (unsigned long long)(it / H) + wskew;) {
                        __builtin_amdgcn_s_sleep(1);
                        if (++spins > spin_limit) {
                            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            break;
                        }
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
            __syncthreads();
            // the halo: the held rows below and above the tile (all held
            // columns), then the held columns left and right of it (the
            // tile's rows); two positions per thread and round, loads first
            const double* xp = xb + (size_t)((it / H - 1) & 1) * nb * xblk;
            const int nbot = (r0 - ja) * nw, ntop = (jz - r1 + 1) * nw, nside = r1 - r0;
            const int nleft = nside * (c0 - ia), nh = nbot + ntop + nleft + nside * (iz - c1 + 1);
            auto halo_pos = [&](int e, int& J, int& i) {
                if (e < nbot) {
                    J = ja + e / nw, i = ia + e % nw;
                } else if ((e -= nbot) < ntop) {
                    J = r1 + e / nw, i = ia + e % nw;
                } else if ((e -= ntop) < nleft) {
                    J = r0 + e / (c0 - ia), i = ia + e % (c0 - ia);
                } else {
                    e -= nleft;
                    J = r0 + e / (iz - c1 + 1), i = c1 + e % (iz - c1 + 1);
                }
            };
            for (int e = tid; e < nh; e += 2 * nt) {
                double v[2][kLdsState];
                int l[2];
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    l[u] = -1;
                    if (e + u * nt >= nh) continue;
                    int J, i;
                    halo_pos(e + u * nt, J, i);
                    const int oy = min(J / R, nby - 1), ox = min(i / C, nbx - 1);
                    const double* src = xp + (size_t)(oy * nbx + ox) * xblk +
                                        (size_t)(J - oy * R) * kLdsState * CL + (i - ox * C);
#pragma unroll
                    for (int q = 0; q < kLdsState; ++q) v[u][q] = src[(size_t)q * CL];
                    l[u] = (i - K0) + sl * (J - L0);
                }
#pragma unroll
                for (int u = 0; u < 2; ++u)
                    if (l[u] >= 0) {
#pragma unroll
                        for (int q = 0; q < kLdsState; ++q) smem[q * slot + l[u]] = v[u][q];
                    }
            }
            __syncthreads();
        }
        if (!MUL && it + 1 == niter) {
            // the divide-form err takes ∇V from the velocities before the
            // last update, from the ping-pong partner: leave them there
            const Fields& prev = (niter & 1) ? f0 : f1;
            for (int e = tid; e < (r1 - r0) * ncol; e += nt) {
                const int J = r0 + e / ncol, i = c0 + e % ncol, o = i + sx * J, l = (i - K0) + sl * (J - L0);
                prev.p[PTS_VX][o] = Vx[l];
                prev.p[PTS_VY][o] = Vy[l];
            }
        }
        // compute_P!, compute_τ! (Stokes2D.jl:33-44, Stokes2D_ve.jl:42-55,
        // Stokes2D_ve_perf.jl:47-53): cells and vertices, in place; a thread
        // updates cell (i, J) and vertex (i, J) together (the vertex rows
        // and columns lie within the cell rows and columns), two independent
        // chains in one pass
        const int nkc = max(kz - ka + 1, 0), ncell = max(cz - ca + 1, 0) * nkc;
        for (int e = tid; e < ncell; e += nt) {
            const int J = ca + e / nkc, i = ka + e % nkc, o = (i - K0) + sl * (J - L0);
            {
                const double g = Gd[o], m = Mu[o];
                if constexpr (MUL) {
                    const double ex = c._dx * (Vx[o + 1] - Vx[o]), ey = c._dy * (Vy[o + sl] - Vy[o]);
                    Pt[o] = Pt[o] - c.r * g * (ex + ey);
                    const double txo = VE ? txxo[o] : 0.0, tyo = VE ? tyyo[o] : 0.0;
                    txx[o] = tau_upd<VE>(txx[o], txo, g, m, c.Gdt, ex);
                    tyy[o] = tau_upd<VE>(tyy[o], tyo, g, m, c.Gdt, ey);
                } else {
                    const double d = div_h(Vx[o + 1] - Vx[o], c.dx, rdx) + div_h(Vy[o + sl] - Vy[o], c.dy, rdy);
                    Pt[o] = Pt[o] - c.r * g * d;
                    const double ex = Vx[o + 1] - Vx[o], ey = Vy[o + sl] - Vy[o];
                    if constexpr (VE) {
                        txx[o] = tau_upd<true>(txx[o], txxo[o], g, m, c.Gdt, div_h(ex, c.dx, rdx));
                        tyy[o] = tau_upd<true>(tyy[o], tyyo[o], g, m, c.Gdt, div_h(ey, c.dy, rdy));
                    } else {  // tau_visc2d_div: (τ + ((2·Gdτ)·d)/dx)/(Gdτ/μ + 1)
                        txx[o] = (txx[o] + div_h(2.0 * g * ex, c.dx, rdx)) / (g / m + 1.0);
                        tyy[o] = (tyy[o] + div_h(2.0 * g * ey, c.dy, rdy)) / (g / m + 1.0);
                    }
                }
            }
            if (J <= vz && i <= vkz) {
                const double ag = (Gd[o] + Gd[o + 1] + Gd[o + sl] + Gd[o + sl + 1]) * 0.25;
                const double am = (Mu[o] + Mu[o + 1] + Mu[o + sl] + Mu[o + sl + 1]) * 0.25;
                const double ev = 0.5 * (dfd_h<MUL>(Vx[o + sl + 1] - Vx[o + 1], c.dy, c._dy, rdy) +
                                         dfd_h<MUL>(Vy[o + sl + 1] - Vy[o + sl], c.dx, c._dx, rdx));
                const double to = VE ? txyo[o] : 0.0;
                txy[o] = tau_upd<VE>(txy[o], to, ag, am, c.Gdt, ev);
            }
        }
        __syncthreads();
        // compute_dV!, compute_V! (Stokes2D.jl:46-58, Stokes2D_ve_perf.jl:55-59),
        // in place: the residual at index (i-1, J-1) updates face (i, J); the
        // free-slip copies of bc_y!(Vx), bc_x!(Vy) (Stokes2D_ve.jl:75-85) are
        // stored with it (a block updates the row or column it copies: R, C
        // >= 2, and the updated rows or columns never start at ny-1 or nx-1),
        // and the copies of the Dirichlet corners (columns 0, nx of Vx; rows
        // 0, ny of Vy) follow.  A thread updates Vx and Vy of face (i, J)
        // together: the Vx rows ya..xz lie within the Vy rows ya..yz, the Vy
        // columns fa..fyz within the Vx columns fa..fz.
        const int nkf = max(fz - fa + 1, 0), nface = max(yz - ya + 1, 0) * nkf;
        for (int e = tid; e < nface + 8; e += nt) {
            if (e < nface) {
                const int J = ya + e / nkf, i = fa + e % nkf, o = (i - 1 - K0) + sl * (J - 1 - L0);
                if (J <= xz) {
                    const double Rr = res2_x_h<MUL>(lf, c, o, rdx, rdy);
                    const double v = Vx[o + sl + 1] + Rr * ((Rh[o + sl] + Rh[o + sl + 1]) * 0.5);
                    Vx[o + sl + 1] = v;
                    if (J == 1) Vx[o + 1] = v;                   // Vx[:, 0] = Vx[:, 1]
                    if (J == ny - 2) Vx[o + 2 * sl + 1] = v;     // Vx[:, ny-1] = Vx[:, ny-2]
                }
                if (i <= fyz) {
                    const double Rr = res2_y_h<MUL>(lf, c, o, BF, rdx, rdy);
                    const double v = Vy[o + sl + 1] + Rr * ((Rh[o + 1] + Rh[o + sl + 1]) * 0.5);
                    Vy[o + sl + 1] = v;
                    if (i == 1) Vy[o + sl] = v;                  // Vy[0, :] = Vy[1, :]
                    if (i == nx - 2) Vy[o + sl + 2] = v;         // Vy[nx-1, :] = Vy[nx-2, :]
                }
            } else {
                const int k = e - nface;
                if (k < 4) {  // Vx at the Dirichlet columns 0 and nx, rows 0 and ny-1
                    const int i = (k & 1) ? nx : 0, J = k < 2 ? 0 : ny - 1, Js = k < 2 ? 1 : ny - 2;
                    if (J >= ba && J <= bz && i >= qa && i <= qz)
                        Vx[(i - K0) + sl * (J - L0)] = Vx[(i - K0) + sl * (Js - L0)];
                } else {      // Vy at the Dirichlet rows 0 and ny, columns 0 and nx-1
                    const int J = (k & 1) ? ny : 0, i = k < 6 ? 0 : nx - 1, is = k < 6 ? 1 : nx - 2;
                    if (J >= ba && J <= bz && i >= qa && i <= qz)
                        Vy[(i - K0) + sl * (J - L0)] = Vy[(is - K0) + sl * (J - L0)];
                }
            }
        }
        __syncthreads();
        if (st == H - 1 && it + 1 < niter) {
            // publish the own positions that other blocks hold as halo (those
            // within H of an edge of the tile), at their place in the tile
            double* xp = xb + (size_t)((it / H) & 1) * nb * xblk + (size_t)b * xblk;
            const int nrow = r1 - r0;
            for (int e = tid; e < nrow * ncol; e += nt) {
                const int h = e / ncol, w = e % ncol;
                if (h >= H && h < nrow - H && w >= H && w < ncol - H) continue;
                const int l = (c0 + w - K0) + sl * (r0 + h - L0);
                double* dst = xp + (size_t)h * kLdsState * CL + w;
#pragma unroll
                for (int q = 0; q < kLdsState; ++q) dst[(size_t)q * CL] = smem[q * slot + l];
            }
            __syncthreads();
            if (tid == 0) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                __hip_atomic_store(flags + b, gen0 + (unsigned long long)(it / H) + 1ull, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
    // the own positions of the state to the buffer the iterations end in
    const Fields& out = (niter & 1) ? f1 : f0;
    for (int e = tid; e < (r1 - r0) * ncol; e += nt) {
        const int J = r0 + e / ncol, i = c0 + e % ncol, o = i + sx * J, l = (i - K0) + sl * (J - L0);
#pragma unroll
        for (int q = 0; q < kLdsState; ++q) out.p[lds_state_field(q)][o] = smem[q * slot + l];
    }
}

}  // namespace

hipError_t launch2_compute_P(hipStream_t s, const Fields& f, const StepScalars& c) {
    k2_P<<<grid2(f.nx + 1, f.ny + 1), dim3(BX, BY), 0, s>>>(f, c);
    return hipGetLastError();
}
hipError_t launch2_compute_tau(hipStream_t s, const Fields& f, const StepScalars& c, bool ve) {
    if (ve) k2_tau<true><<<grid2(f.nx + 1, f.ny + 1), dim3(BX, BY), 0, s>>>(f, c);
    else k2_tau<false><<<grid2(f.nx + 1, f.ny + 1), dim3(BX, BY), 0, s>>>(f, c);
    return hipGetLastError();
}
hipError_t launch2_compute_dV(hipStream_t s, const Fields& f, const StepScalars& c, bool bf) {
    k2_dV<<<grid2(f.nx + 1, f.ny + 1), dim3(BX, BY), 0, s>>>(f, c, bf ? 1 : 0);
    return hipGetLastError();
}
hipError_t launch2_compute_V(hipStream_t s, const Fields& f) {
    k2_V<<<grid2(f.nx + 1, f.ny + 1), dim3(BX, BY), 0, s>>>(f);
    return hipGetLastError();
}
hipError_t launch2_compute_Res(hipStream_t s, const Fields& f, const StepScalars& c, bool bf, bool mul) {
    if (mul) k2_Res<true><<<grid2(f.nx + 1, f.ny + 1), dim3(BX, BY), 0, s>>>(f, c, bf ? 1 : 0);
    else k2_Res<false><<<grid2(f.nx + 1, f.ny + 1), dim3(BX, BY), 0, s>>>(f, c, bf ? 1 : 0);
    return hipGetLastError();
}
hipError_t launch2_compute_Pt_tau(hipStream_t s, const Fields& f, const StepScalars& c, bool ve) {
    if (ve) k2_Pt_tau<true><<<grid2(f.nx + 1, f.ny + 1), dim3(BX, BY), 0, s>>>(f, c);
    else k2_Pt_tau<false><<<grid2(f.nx + 1, f.ny + 1), dim3(BX, BY), 0, s>>>(f, c);
    return hipGetLastError();
}
hipError_t launch2_compute_V_perf(hipStream_t s, const Fields& f, const StepScalars& c, bool bf) {
    k2_V_perf<<<grid2(f.nx + 1, f.ny + 1), dim3(BX, BY), 0, s>>>(f, c, bf ? 1 : 0);
    return hipGetLastError();
}

// Resident blocks of one k2_iter instance on the whole device.
template <bool VE, bool MUL, bool BF, bool PF, bool NT, int NW, bool DP, int MX, bool MULTI = false>
static int concurrency2() {
    static int cached = 0;
    if (!cached) {
        int dev = 0, cus = 0, per = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k2_iter<VE, MUL, BF, PF, NT, NW, DP, MX, MULTI>, 64 * NW, 0);
        cached = std::max(1, cus) * std::max(1, per);
    }
    return cached;
}

template <bool VE, bool MUL, bool BF, bool PF, bool NT, int NW = 1, bool DP = false, int MX = 0>
static void go2v(hipStream_t s, const Fields& in, const Fields& out, const StepScalars& c, int L, bool persist) {
    constexpr int W = 64 * NW - 2;
    const int ns = (in.nx + 1 + W - 1) / W;
    const int conc = concurrency2<VE, MUL, BF, PF, NT, NW, DP, MX>();
    if (L < 1) {  // about one round of resident waves (L < 0: within [4, -L] rows)
        const int nch = std::max(1, conc / ns);
        const int cap = -L;
        L = (in.ny + nch - 1) / nch;
        if (cap > 0) L = std::max(4, std::min(L, cap));
    }
    const int ntiles = ns * ((in.ny + L - 1) / L);
    const int grid = persist ? std::min(ntiles, conc) : ntiles;
    k2_iter<VE, MUL, BF, PF, NT, NW, DP, MX><<<grid, 64 * NW, 0, s>>>(in, out, c, L, ns, ntiles, 1, nullptr, nullptr);
}

// One cooperative launch of `niter` iterations of the default single-pass
// kernel (8-wave strips, Gdτ, dτ_Rho and Musτ derived), `rows` rows per tile;
// hipErrorNotSupported when its tiles do not fit one round of resident blocks.
template <bool VE, bool MUL, bool BF>
static hipError_t go2multi(hipStream_t s, const Fields& in, const Fields& out, const StepScalars& c, int rows,
                           int niter, unsigned* bar, int* err) {
    constexpr int NW = 8, W = 64 * NW - 2;
    auto kern = k2_iter<VE, MUL, BF, true, false, NW, true, 2, true>;
    const int ns = (in.nx + 1 + W - 1) / W;
    const int conc = concurrency2<VE, MUL, BF, true, false, NW, true, 2, true>();
    int L = std::max(1, rows);
    const int ntiles = ns * ((in.ny + L - 1) / L);
    if (ntiles > conc) return hipErrorNotSupported;
    Fields a = in, b = out;
    StepScalars cc = c;
    int nss = ns, nt = ntiles, ni = niter;
    void* args[] = {&a, &b, &cc, &L, &nss, &nt, &ni, &bar, &err};
    return hipLaunchCooperativeKernel((const void*)kern, dim3(ntiles), dim3(64 * NW), args, 0, s);
}

// dp: derive Gdτ, dτ_Rho from Musτ (variants 1, 7 and 10; the others read
// them); mx: derive Musτ from μ as well (variant 7)
template <bool VE, bool MUL, bool BF>
static void go2(hipStream_t s, const Fields& in, const Fields& out, const StepScalars& c, int L, int variant,
                bool persist, bool dp, bool mx) {
    if (dp && mx && variant == 7) return go2v<VE, MUL, BF, true, false, 8, true, 2>(s, in, out, c, L, persist);
    if (dp && mx && variant == 11) return go2v<VE, MUL, BF, true, false, 8, true, 1>(s, in, out, c, L, persist);
    if (dp && variant == 7) return go2v<VE, MUL, BF, true, false, 8, true>(s, in, out, c, L, persist);
    if (dp && variant == 10) return go2v<VE, MUL, BF, false, false, 8, true>(s, in, out, c, L, persist);
    if (dp && variant == 1) return go2v<VE, MUL, BF, true, false, 1, true>(s, in, out, c, L, persist);
    switch (variant) {
        case 0: go2v<VE, MUL, BF, false, false>(s, in, out, c, L, persist); break;
        case 2: go2v<VE, MUL, BF, true, true>(s, in, out, c, L, persist); break;
        case 3: go2v<VE, MUL, BF, false, true>(s, in, out, c, L, persist); break;
        case 4: go2v<VE, MUL, BF, true, false, 4>(s, in, out, c, L, persist); break;
        case 5: go2v<VE, MUL, BF, true, false, 2>(s, in, out, c, L, persist); break;
        case 6: go2v<VE, MUL, BF, true, true, 4>(s, in, out, c, L, persist); break;
        case 7: go2v<VE, MUL, BF, true, false, 8>(s, in, out, c, L, persist); break;
        case 8: go2v<VE, MUL, BF, true, true, 8>(s, in, out, c, L, persist); break;
        case 9: go2v<VE, MUL, BF, true, false, 16>(s, in, out, c, L, persist); break;
        default: go2v<VE, MUL, BF, true, false>(s, in, out, c, L, persist); break;
    }
}

// Fields held in LDS by k2_lds, and its LDS bytes for nl x nc positions.
int lds2_fields(bool ve, bool bf) { return 9 + (bf ? 1 : 0) + (ve ? 3 : 0); }
size_t lds2_bytes(int nl, int nc, bool ve, bool bf) {
    return (size_t)lds2_fields(ve, bf) * nl * nc * sizeof(double);
}

template <bool VE, bool MUL, bool BF, bool DP>
static hipError_t go2lds(hipStream_t s, const Fields& f0, const Fields& f1, const StepScalars& c, int R, int nby,
                         int C, int nbx, int H, int niter, double* xb, unsigned long long* flags,
                         unsigned long long gen0, int* err, int threads, int spin_limit, unsigned long long wskew) {
    auto kern = H == 4   ? k2_lds<VE, MUL, BF, DP, 4>
                : H == 3 ? k2_lds<VE, MUL, BF, DP, 3>
                : H == 2 ? k2_lds<VE, MUL, BF, DP, 2>
                         : k2_lds<VE, MUL, BF, DP, 1>;
    const int RL = f0.ny + 1 - (nby - 1) * R, CL = f0.nx + 1 - (nbx - 1) * C;
    if (nby < 1 || nbx < 1 || R < 2 || C < 2 || RL < R || CL < C) return hipErrorInvalidValue;
    const size_t bytes = lds2_bytes(lds2_extent(f0.ny + 1, R, RL, H), lds2_extent(f0.nx + 1, C, CL, H), VE, BF);
    hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
    Fields a = f0, bb = f1;
    StepScalars cc = c;
    int rr = R, cl = C, nx_ = nbx, ni = niter, sp = spin_limit;
    unsigned long long g0 = gen0, ws = wskew;
    void* args[] = {&a, &bb, &cc, &rr, &cl, &nx_, &ni, &xb, &flags, &g0, &err, &sp, &ws};
    return hipLaunchCooperativeKernel((const void*)kern, dim3(nby * nbx), dim3(threads), args, (unsigned)bytes, s);
}

hipError_t launch2_iter_lds(hipStream_t s, const Fields& f0, const Fields& f1, const StepScalars& c, bool ve,
                            bool mul, bool bf, bool dp, int R, int nby, int C, int nbx, int H, int niter, double* xb,
                            unsigned long long* flags, unsigned long long gen0, int* err, int threads, int spin_limit,
                            unsigned long long wskew) {
#define PTS_LDS_GO(V, M, B, D) \
    return go2lds<V, M, B, D>(s, f0, f1, c, R, nby, C, nbx, H, niter, xb, flags, gen0, err, threads, spin_limit, wskew)
    if (ve) {
        if (mul) { if (bf) { if (dp) PTS_LDS_GO(true, true, true, true); PTS_LDS_GO(true, true, true, false); }
                   if (dp) PTS_LDS_GO(true, true, false, true); PTS_LDS_GO(true, true, false, false); }
        if (bf) { if (dp) PTS_LDS_GO(true, false, true, true); PTS_LDS_GO(true, false, true, false); }
        if (dp) PTS_LDS_GO(true, false, false, true); PTS_LDS_GO(true, false, false, false);
    }
    if (mul) { if (bf) { if (dp) PTS_LDS_GO(false, true, true, true); PTS_LDS_GO(false, true, true, false); }
               if (dp) PTS_LDS_GO(false, true, false, true); PTS_LDS_GO(false, true, false, false); }
    if (bf) { if (dp) PTS_LDS_GO(false, false, true, true); PTS_LDS_GO(false, false, true, false); }
    if (dp) PTS_LDS_GO(false, false, false, true); PTS_LDS_GO(false, false, false, false);
#undef PTS_LDS_GO
}

hipError_t launch2_iter_multi(hipStream_t s, const Fields& in, const Fields& out, const StepScalars& c, bool ve,
                              bool mul, bool bf, int rows, int niter, unsigned* bar, int* err) {
    if (ve) {
        if (mul) return bf ? go2multi<true, true, true>(s, in, out, c, rows, niter, bar, err)
                           : go2multi<true, true, false>(s, in, out, c, rows, niter, bar, err);
        return bf ? go2multi<true, false, true>(s, in, out, c, rows, niter, bar, err)
                  : go2multi<true, false, false>(s, in, out, c, rows, niter, bar, err);
    }
    if (mul) return bf ? go2multi<false, true, true>(s, in, out, c, rows, niter, bar, err)
                       : go2multi<false, true, false>(s, in, out, c, rows, niter, bar, err);
    return bf ? go2multi<false, false, true>(s, in, out, c, rows, niter, bar, err)
              : go2multi<false, false, false>(s, in, out, c, rows, niter, bar, err);
}

hipError_t launch2_iter(hipStream_t s, const Fields& in, const Fields& out, const StepScalars& c, bool ve,
                        bool mul, bool bf, int L, int variant, bool persist, bool dp, bool mx) {
    if (ve) {
        if (mul) { if (bf) go2<true, true, true>(s, in, out, c, L, variant, persist, dp, mx); else go2<true, true, false>(s, in, out, c, L, variant, persist, dp, mx); }
        else { if (bf) go2<true, false, true>(s, in, out, c, L, variant, persist, dp, mx); else go2<true, false, false>(s, in, out, c, L, variant, persist, dp, mx); }
    } else {
        if (mul) { if (bf) go2<false, true, true>(s, in, out, c, L, variant, persist, dp, mx); else go2<false, true, false>(s, in, out, c, L, variant, persist, dp, mx); }
        else { if (bf) go2<false, false, true>(s, in, out, c, L, variant, persist, dp, mx); else go2<false, false, false>(s, in, out, c, L, variant, persist, dp, mx); }
    }
    return hipGetLastError();
}

}  // namespace pts
