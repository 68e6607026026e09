// pts_internal.hpp -- shared definitions of libpts (gfx950 HIP).
//
// Device layout: every field of a context lives in its own allocation with
// the SAME strides: element (i,j,k) of any field is at i + sx*j + sxy*k,
// sx = nx+1 rounded up to 32 elements (256 B rows), sxy = sx*(ny+1).  A field
// only uses its reference extent inside that box (e.g. τxy uses
// (nx-1)x(ny-1)); one index therefore addresses every array of a stencil, so
// a kernel computes one offset per cell instead of one per array.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/pts.h"

namespace pts {

constexpr int kNF = PTS_NFIELDS;

// All device pointers of a context plus the shared strides, passed by value.
struct Fields {
    double* p[kNF];
    int sx;    // row stride (elements)
    int sxy;   // plane stride (elements)
    int nx, ny, nz;
};

// FP32 fields of the mixed-precision correction solve (same strides).
struct Fields32 {
    float* p[kNF];
    int sx, sxy;
    int nx, ny, nz;
};

// Reference shape of a field for the grid (shape[2] = 1 in 2D).
void field_shape(int ndim, int nx, int ny, int nz, int field, int64_t out[3]);

struct Ctx;

// ---- shuffles across the 64-wide wavefront ---------------------------------
// from_left(v)[lane] = v[lane-1]; from_right(v)[lane] = v[lane+1].  Lanes at
// the wave ends receive 0.  DPP wave_shr/wave_shl (GFX9 family) move one
// dword per instruction without touching LDS.
// XCD-aware tile order.  The hardware deals consecutive blocks of a launch to
// the 8 XCDs round-robin (block b runs on XCD b % 8); this bijection of
// [0, n) gives each XCD one contiguous run of tiles, so neighbouring tiles
// (which share halo rows and cache lines) meet in the same L2.
__device__ __forceinline__ int xcd_tile(int b, int n) {
    const int x = b & 7, i = b >> 3, q = n >> 3, rem = n & 7;
    return x * q + min(x, rem) + i;
}

// bound_ctrl: lanes without a source lane write 0 (no old value to set up).
__device__ __forceinline__ double from_left(double v) {
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_mov_dpp(lo, 0x138, 0xf, 0xf, true);
    hi = __builtin_amdgcn_mov_dpp(hi, 0x138, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double from_right(double v) {
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_mov_dpp(lo, 0x130, 0xf, 0xf, true);
    hi = __builtin_amdgcn_mov_dpp(hi, 0x130, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}

// Scaled difference in the two flavours of the reference: (a-b)/dx in the
// 4-kernel scripts, _dx*(a-b) in the *_perf scripts.
template <bool MUL>
__device__ __forceinline__ double dfd(double diff, double h, double ih) {
    if constexpr (MUL) return ih * diff;
    else return diff / h;
}

// Normal-stress update τ ← (τ + τ_o·Gr + 2·Gdτ·ε̇)/(1 + Gdτ/μ + Gr), Gr = Gdτ/(G·dt)
// (Stokes2D_ve.jl:51-52, Stokes3D_ve_perf.jl:71-73); viscous drops τ_o and Gr.
// e is the already-scaled strain rate.
template <bool VE>
__device__ __forceinline__ double tau_upd(double t, double to, double g, double m, double Gdt, double e) {
    if constexpr (VE) {
        double Gr = g / Gdt;
        return (t + to * Gr + 2.0 * g * e) / (1.0 + g / m + Gr);
    } else {
        return (t + 2.0 * g * e) / (1.0 + g / m);
    }
}

// Stokes2D.jl:40-41 writes the viscous normal stress as
// (τ + 2.0*Gdτ*d/dx)/(Gdτ/μ + 1.0), i.e. ((2·Gdτ)·d)/dx.
__device__ __forceinline__ double tau_visc2d_div(double t, double g, double m, double d, double h) {
    return (t + 2.0 * g * d / h) / (g / m + 1.0);
}

// Write v at (a,b) on the tangential plane of a velocity component and at the
// free-slip copies it feeds: index a (stride sa, extent na) and b (stride sb,
// extent nb); a==1 also writes a=0, a==na-2 also writes a=na-1 (both when
// na == 3: the one interior plane feeds both boundary planes), same for b
// (bc_y!/bc_z! etc., Stokes3D_ve_perf.jl:202-207; edge values end up equal to
// the diagonal interior neighbour whatever the order of the copies).
// E3 = false drops the extent-3 code (for kernels instantiated for grids
// with every extent > 3).
template <bool E3 = true, typename T>
__device__ __forceinline__ void put_bc(T* V, int q, T v, int a, int na, int sa, int b, int nb, int sb) {
    V[q] = v;
    const int da = (a == 1) ? -1 : ((a == na - 2) ? 1 : 0);
    const int db = (b == 1) ? -1 : ((b == nb - 2) ? 1 : 0);
    if (da) V[q + da * sa] = v;
    if (db) V[q + db * sb] = v;
    if (da && db) V[q + da * sa + db * sb] = v;
    if (E3 && (na == 3 || nb == 3)) {  // grid-uniform: the far-side copies of extent 3
        const bool a3 = na == 3 && a == 1, b3 = nb == 3 && b == 1;
        if (a3) {
            V[q + sa] = v;
            if (db) V[q + sa + db * sb] = v;
        }
        if (b3) {
            V[q + sb] = v;
            if (da) V[q + da * sa + sb] = v;
            if (a3) V[q + sa + sb] = v;
        }
    }
}

// Scalars of one PT iteration.  dp kernels (`dp` launch flag) read Musτ and
// derive Gdτ, dτ_Rho from it.
struct StepScalars {
    double r, Gdt;
    double dx, dy, dz, _dx, _dy, _dz;
    // constants of compute_iter_params! for kernels that derive Gdτ and
    // dτ_Rho from Musτ instead of reading them (dp_ve: visco-elastic μ_eff)
    int dp_ve;
    double ip_igdt, ip_k1, ip_k2, ip_r2;
    // 1/(G·dt) when G·dt is a power of two (then Gdτ·iGdt == Gdτ/(G·dt)), else 0
    double iGdt;
};

// dτ_Rho and Gdτ of one cell from Musτ, operation for operation as
// compute_iter_params! (Stokes2D_ve.jl:28-33, Stokes2D.jl:27-31) evaluates
// them, with the uniform sub-expressions 1/(G·dt), Vpdτ·max_l/Re, Vpdτ·Vpdτ
// and r+2 precomputed on the host: bit-identical to the stored arrays.
__device__ __forceinline__ void iter_params_of(double M, const StepScalars& c, double& rho, double& gd) {
    const double mu_eff = c.dp_ve ? 1.0 / (1.0 / M + c.ip_igdt) : M;
    rho = c.ip_k1 / mu_eff;
    gd = c.ip_k2 / rho / c.ip_r2;
}

struct StepScalars32 {
    float r, _dx, _dy, _dz;
    // derived pseudo-time steps (3D sweep with dp): dτ_Rho = k1 / Musτ,
    // Gdτ = kg · Musτ (compute_iter_params!'s Vpdτ·max_l/Re / Musτ and
    // Vpdτ² / dτ_Rho / (r+2), folded)
    float k1, kg;
};

// ---- kernel launchers (defined in the .hip files) ---------------------------
hipError_t launch_fill(hipStream_t s, double* p, double v, const Fields& f, const int64_t shape[3]);
hipError_t launch_copy(hipStream_t s, double* dst, const double* src, const Fields& f, const int64_t shape[3]);
hipError_t launch_init_inclusion(hipStream_t s, const Fields& f, int ndim, double lx, double ly, double lz,
                                 double mus0, double musi, double rad2, double ebg, double rin, double rout,
                                 int igg);
hipError_t launch_smooth(hipStream_t s, const Fields& f, int ndim, double* dst, const double* src,
                         const int64_t shape[3], double fact);
hipError_t launch_maxloc(hipStream_t s, const Fields& f, int ndim, double* dst, const double* src,
                         const int64_t shape[3]);
hipError_t launch_bc(hipStream_t s, const Fields& f, double* a, const int64_t shape[3], int dim);
hipError_t launch_grf(hipStream_t s, const Fields& f, int ndim, const double* modes, int nmodes, double sigma,
                      double lo, double hi, double lx, double ly, double lz, double musi, double rad2);
hipError_t launch_iter_params(hipStream_t s, const Fields& f, int ndim, bool ve, double Vp, double G, double dt,
                              double Re, double r, double ml);

// 2D per-kernel API (in place, reference semantics)
hipError_t launch2_compute_P(hipStream_t s, const Fields& f, const StepScalars& c);
hipError_t launch2_compute_tau(hipStream_t s, const Fields& f, const StepScalars& c, bool ve);
hipError_t launch2_compute_dV(hipStream_t s, const Fields& f, const StepScalars& c, bool bf);
hipError_t launch2_compute_V(hipStream_t s, const Fields& f);
hipError_t launch2_compute_Res(hipStream_t s, const Fields& f, const StepScalars& c, bool bf, bool mul);
hipError_t launch2_compute_Pt_tau(hipStream_t s, const Fields& f, const StepScalars& c, bool ve);
hipError_t launch2_compute_V_perf(hipStream_t s, const Fields& f, const StepScalars& c, bool bf);
// 2D single-pass iteration: reads `in`, writes `out` (ping-pong).
// rows_per_block < 1 sizes the row chunks so that the tiles fill about one
// round of resident waves; persist caps the launch at that many waves, each
// looping over tiles.  variant: 0 plain, 1 prefetch, 2 prefetch + non-temporal
// stores, 3 non-temporal stores.
hipError_t launch2_iter(hipStream_t s, const Fields& in, const Fields& out, const StepScalars& c, bool ve,
                        bool mul, bool bf, int rows_per_block, int variant, bool persist, bool dp, bool mx);
// `niter` iterations of the default single-pass kernel (derived Gdτ, dτ_Rho and
// Musτ) in one cooperative launch with a grid barrier between iterations
// (bar: 2 unsigned, count and generation, zeroed once; *err set if a barrier
// wait gave up).  in/out swap with the parity of the iteration, as
// pts_pt_iterate's ping-pong does.  hipErrorNotSupported when the tiles of
// `rows` rows do not fit one round of resident blocks.
// Small 2D grids with the state in LDS (k2_lds): niter iterations in one
// cooperative launch of nby x nbx blocks, tiles of R >= 2 rows and C >= 2
// columns of the positions 0..nx x 0..ny (the last tile of a row or column
// of tiles: what is left, at least R or C), exchanging H positions (1 to 4)
// every H iterations; xb holds 2 * nby * nbx * RL * 6 * CL doubles (RL =
// ny + 1 - (nby - 1) * R, CL = nx + 1 - (nbx - 1) * C: the last tiles'
// extents), flags one counter per block (gen0: their value when the launch
// starts; they end at gen0 + (niter - 1) / H); threads per block: a multiple
// of 64 up to 1024; a wait gives up (setting *err) after
// spin_limit polls, and waits for its generation + wskew (0; the tests raise
// it so that every wait gives up).  lds2_extent: the rows (columns) of LDS a block holds,
// for tiles of t positions (the last tl) over m with h halo positions on
// each side, clipped to the grid; lds2_bytes: the LDS of nl x nc positions.
__host__ __device__ inline int lds2_extent(int m, int t, int tl, int h) {
    const int e = (t > tl ? t : tl) + 2 * h;
    return e < m ? e : m;
}
int lds2_fields(bool ve, bool bf);
size_t lds2_bytes(int nl, int nc, bool ve, bool bf);
hipError_t launch2_iter_lds(hipStream_t s, const Fields& f0, const Fields& f1, const StepScalars& c, bool ve,
                            bool mul, bool bf, bool dp, int R, int nby, int C, int nbx, int H, int niter, double* xb,
                            unsigned long long* flags, unsigned long long gen0, int* err, int threads, int spin_limit,
                            unsigned long long wskew);
hipError_t launch2_iter_multi(hipStream_t s, const Fields& in, const Fields& out, const StepScalars& c, bool ve,
                              bool mul, bool bf, int rows, int niter, unsigned* bar, int* err);
// 1 in *bad unless Musτ equals compute_maxloc! of μ followed by bc_x!, bc_y!
// (2D; the precondition of the mx iteration kernels).  *bad must be zero.
hipError_t launch_check_mustau2(hipStream_t s, const Fields& f, int* bad);
// 3D: 1 in *bad unless Musτ equals compute_maxloc! of μ followed by bc_x!,
// bc_y!, bc_z! (the precondition of the mx form of k3_iter2); needs n >= 3.
hipError_t launch_check_mustau3(hipStream_t s, const Fields& f, int* bad);

// 3D
hipError_t launch3_compute_P(hipStream_t s, const Fields& f, const StepScalars& c);
hipError_t launch3_compute_tau(hipStream_t s, const Fields& f, const StepScalars& c, bool ve);
hipError_t launch3_compute_dV(hipStream_t s, const Fields& f, const StepScalars& c, bool bf);
hipError_t launch3_compute_V(hipStream_t s, const Fields& f);
hipError_t launch3_compute_Res(hipStream_t s, const Fields& f, const StepScalars& c, bool bf, bool mul);
hipError_t launch3_compute_Pt_tau(hipStream_t s, const Fields& f, const StepScalars& c, bool ve, bool mul);
hipError_t launch3_compute_V_perf(hipStream_t s, const Fields& f, const StepScalars& c, bool bf, bool mul,
                                  bool fold_bc);
// 3D single-pass iteration (ping-pong)
// variant 0: k3_iter; 1: k3_iter2; 2: k3_iter2 with the next plane's loads
// issued before the current plane's arithmetic.  k3_iter2 needs the grouped
// 3D allocations (slot_bytes apart, with guard zones: see ensure_group and
// field_guard_elems in pts_api.hip); otherwise k3_iter runs.  mx: form Musτ
// from μ (k3_iter2 without prefetch); map_flat: k3_iter2 takes the tiles in
// the flat (chunk, tile) order instead of each XCD sweeping its own tiles;
// w64: k3_iter2 with derived parameters addresses each slot of a grouped
// field through its own pointer also where 32-bit slot offsets would fit (it
// must where they do not, e.g. at 1023^3; at 511^3 it is about 1 % faster).
hipError_t launch3_iter(hipStream_t s, const Fields& in, const Fields& out, const StepScalars& c, bool ve,
                        bool mul, bool bf, int planes_per_block, int tile_rows, bool dp, int variant,
                        int64_t slot_bytes, bool mx, bool map_flat, bool w64, bool stream_loads);

// Mixed-precision defect correction (pts_dc.hip)
// lz: planes per block of the 3D sweep (<0: fewest block rounds within [4, -lz])
// dp: the 3D sweep reads Musτ and derives Gdτ, dτ_Rho from it (c.k1, c.kg)
// map_xcd: each XCD sweeps its own eighth of the tiles chunk by chunk
hipError_t launch_dc_iter(hipStream_t s, const Fields32& in, const Fields32& out, const StepScalars32& c, int ndim,
                          int lz, bool dp, bool map_xcd);
// z planes per block: Lz > 0 as given, else the length within [4, -Lz] with the
// fewest rounds of resident blocks × plane steps
int planes_per_block(int Lz, int nz, int64_t ntile, int64_t conc);
hipError_t launch_dc_project(hipStream_t s, const Fields& f, const StepScalars& c, int ndim);
hipError_t launch_dc_sources(hipStream_t s, const Fields& f, const Fields32& g, const StepScalars& c, int ndim,
                             bool bf);
hipError_t launch_dc_accumulate(hipStream_t s, const Fields& f, const Fields32& g, int ndim);
hipError_t launch_to_f32(hipStream_t s, float* dst, const double* src, int64_t n);
// out[0..3] = Σ R², out[4..7] = max|R| of the correction problem (Rx, Ry, Rz, ∇V)
hipError_t launch_dc_norms(hipStream_t s, const Fields32& g, const StepScalars32& c, int ndim, double* partials,
                           double* out);

// Adjoint of the converged 2D viscous iteration.  `lam`, `tmp`, `out` carry
// the adjoint fields under the forward field ids (PT, TXX, TYY, TXY, VX, VY).
// igdt = 1/(G·dt) for the visco-elastic τ update, 0 for viscous; g carries
// the sources under the field ids VX, VY, TXX, TYY, TXY (null pointers: none).
// fused: the single-pass sweep (same bits as the two passes; tmp unused)
hipError_t launch2_adj_step(hipStream_t s, const Fields& f, const Fields& lam, const Fields& tmp, const Fields& out,
                            const StepScalars& c, double igdt, const Fields& g, bool fused);
hipError_t launch2_adj_pass1(hipStream_t s, const Fields& f, const Fields& lam, const Fields& tmp,
                             const StepScalars& c);
hipError_t launch2_adj_grad(hipStream_t s, const Fields& f, const Fields& l1, double igdt, double* grad, bool acc);
hipError_t launch2_adj_couple(hipStream_t s, const Fields& f, const Fields& l1, double igdt, const Fields& g);
hipError_t launch2_adj_change(hipStream_t s, const Fields& f, const Fields& a, const Fields& b, double* part,
                              double* out);

// Reductions.  `red` holds kRedSlots partial slots per block.
constexpr int kRedBlocks = 512;
constexpr int kRedVals = 12;  // Vmin Vmax Pmin Pmax ss[4] ma[4]
hipError_t launch_norms_stored(hipStream_t s, const Fields& f, int ndim, double* partials, double* out);
// Reductions of launch_norms_stored over dense device arrays (reference
// shapes): `nblocks` partial slots of kRedVals doubles in `partials`, the
// result (kRedVals doubles) in `out`.
struct DenseNorms {
    const double* vx;
    const double* pt;
    const double* r[4];  // Rx, Ry, Rz, ∇V (null: none)
    int64_t nvx, npt, nr[4];
};
hipError_t launch_norms_dense(hipStream_t s, const DenseNorms& d, double* partials, int nblocks, double* out);
// Message reported by pts_last_error(NULL) (calls without a context).
void set_global_error(const std::string& msg);
// On-the-fly residual norms of the current state: R from Pt/τ, ∇V from the
// V fields in `fv` (pass the pre-update V for the 4-kernel semantics).
hipError_t launch_norms_live(hipStream_t s, const Fields& f, const Fields& fv, const StepScalars& c, int ndim,
                             bool bf, bool mul, double* partials, double* out);

}  // namespace pts
