// pts_dc.hip -- mixed-precision defect correction for the viscous PT solve
// (no counterpart in the reference, which is FP64 only).
//
// Outer loop in FP64 (driven from pts_api.hip):
//   1. project τ onto the constitutive law: τ = 2 μ ε̇(V)            (k_dc_project)
//   2. FP64 residual of the full problem, R_P = ∇·V and
//      R_V = ∇·τ - ∇P - ρg, reduced to the stopping norm and cast to FP32
//      source arrays                                                 (k_dc_sources)
//   3. FP32 PT solve of the correction δ = (δP, δτ, δV) of
//         δP ← δP - r Gdτ (∇·δV + R_P)
//         δτ ← (δτ + 2 Gdτ ε̇(δV)) / (1 + Gdτ/μ)
//         δV ← δV + (∇·δτ - ∇δP + R_V) · avg(dτ_Rho)
//      with homogeneous Dirichlet values and the same free-slip copies, to a
//      relative drop of its residual                                 (k2_dc, k3_dc)
//   4. P += δP, V += δV in FP64                                       (k_dc_accumulate)
// Since step 1 makes the stress residual vanish, u + δ obeys the same PT
// update as u itself; only the correction sweeps run in FP32 (half the bytes).
// The FP32 fields reuse the field ids of the FP64 ones; the correction
// sources live in the slots of ∇V (R_P) and Rx, Ry, Rz (R_V).
#include <algorithm>

#include "pts_internal.hpp"

namespace pts {

namespace {

__device__ __forceinline__ float fleft(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float fright(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, false));
}

// The FP32 sweeps keep the arithmetic of the residual that stops them (no
// FMA contraction, correctly rounded division): with FMA or the 1-ulp
// hardware reciprocal the correction solve of the 3D random-viscosity 127^3
// case needs 1.6x the sweeps to reach its relative residual drop.
__device__ __forceinline__ float tau_visc_f(float t, float g, float m, float e) {
    return (t + 2.0f * g * e) / (1.0f + g / m);
}

// ---------------------------------------------------------------------------
// 2D FP32 correction sweep: the single-pass structure of k2_iter (62-column
// strips per wave, marching in y), viscous, multiply form, with sources.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k2_dc(const Fields32 in, const Fields32 out, const StepScalars32 cs, int L,
                                            int nstrips) {
    const int b = xcd_tile(blockIdx.x, gridDim.x);
    const int strip = b % nstrips, chunk = b / nstrips;
    const int lane = threadIdx.x;
    const int cx = strip * 62 - 1 + lane;
    const int nx = in.nx, ny = in.ny, sx = in.sx;
    const int jb = chunk * L;
    const int je = min(jb + L, ny);
    const bool own = lane >= 1 && lane <= 62;
    const bool cin = cx >= 0 && cx < nx;
    const bool fin = cx >= 0 && cx <= nx;
    const bool vin = cx >= 1 && cx <= nx - 1;
    const bool vyin = cx >= 1 && cx <= nx - 2;
    const float _dx = cs._dx, _dy = cs._dy, r = cs.r;

    float VxP = 0.f, GdP = 0.f, MuP = 0.f, GdPl = 0.f, MuPl = 0.f, RhP = 0.f;
    float Pt1P = 0.f, txx1P = 0.f, tyy1P = 0.f, txy1PP = 0.f;
    const int j0 = max(jb - 1, 0), j1 = min(je, ny - 1);
    float VyC = cin ? in.p[PTS_VY][cx + sx * j0] : 0.f;
    for (int J = j0; J <= j1; ++J) {
        const int o = cx + sx * J;
        float Pt0 = 0.f, tx0 = 0.f, ty0 = 0.f, Gd = 0.f, Mu = 0.f, Rh = 0.f, VyN = 0.f, sP = 0.f, sVy = 0.f;
        float Vx0 = 0.f, tz0 = 0.f, sVx = 0.f;
        if (cin) {
            Pt0 = in.p[PTS_PT][o];
            tx0 = in.p[PTS_TXX][o];
            ty0 = in.p[PTS_TYY][o];
            Gd = in.p[PTS_GDTAU][o];
            Mu = in.p[PTS_MUS][o];
            Rh = in.p[PTS_DTAU_RHO][o];
            VyN = in.p[PTS_VY][o + sx];
            sP = in.p[PTS_DIVV][o];
            sVy = in.p[PTS_RY][o];
        }
        if (fin) {
            Vx0 = in.p[PTS_VX][o];
            if (J >= 1) sVx = in.p[PTS_RX][o - sx];
        }
        const bool vtx = vin && J >= 1;
        if (vtx) tz0 = in.p[PTS_TXY][o - 1 - sx];
        const float VxR = fright(Vx0);
        const float GdL = fleft(Gd), MuL = fleft(Mu), VyCL = fleft(VyC);
        const float txx1PL = fleft(txx1P), Pt1PL = fleft(Pt1P), RhPL = fleft(RhP);
        const float ex = _dx * (VxR - Vx0), ey = _dy * (VyN - VyC);
        const float Pt1n = Pt0 - r * Gd * (ex + ey + sP);
        const float txx1n = tau_visc_f(tx0, Gd, Mu, ex);
        const float tyy1n = tau_visc_f(ty0, Gd, Mu, ey);
        const float ag = (GdPl + GdP + GdL + Gd) * 0.25f;
        const float am = (MuPl + MuP + MuL + Mu) * 0.25f;
        const float e = 0.5f * (_dy * (Vx0 - VxP) + _dx * (VyC - VyCL));
        const float txy1n = tau_visc_f(tz0, ag, am, e);
        const float txy1nR = fright(txy1n);
        if (own && cin && J >= jb && J < je) {
            out.p[PTS_PT][o] = Pt1n;
            out.p[PTS_TXX][o] = txx1n;
            out.p[PTS_TYY][o] = tyy1n;
        }
        const int Jm = J - 1;
        if (own && vtx && Jm >= jb && Jm < je) out.p[PTS_TXY][o - 1 - sx] = txy1n;
        if (own && vin && Jm >= jb && Jm < je && Jm >= 1 && Jm <= ny - 2) {
            float R = _dx * (txx1P - txx1PL) + _dy * (txy1n - txy1PP) - _dx * (Pt1P - Pt1PL) + sVx;
            float v = VxP + R * (0.5f * (RhPL + RhP));
            float* V = out.p[PTS_VX];
            V[o - sx] = v;
            if (Jm == 1) V[o - 2 * sx] = v;
            if (Jm == ny - 2) V[o] = v;
        }
        if (own && vyin && J >= jb && J < je && J >= 1) {
            float R = _dy * (tyy1n - tyy1P) + _dx * (txy1nR - txy1n) - _dy * (Pt1n - Pt1P) + sVy;
            float v = VyC + R * (0.5f * (RhP + Rh));
            float* V = out.p[PTS_VY];
            V[o] = v;
            if (cx == 1) V[o - 1] = v;
            if (cx == nx - 2) V[o + 1] = v;
        }
        VxP = Vx0;
        VyC = VyN;
        GdPl = GdL;
        MuPl = MuL;
        GdP = Gd;
        MuP = Mu;
        RhP = Rh;
        Pt1P = Pt1n;
        txx1P = txx1n;
        tyy1P = tyy1n;
        txy1PP = txy1n;
    }
}

// ---------------------------------------------------------------------------
// 3D FP32 correction sweep: the tile structure of k3_iter (64 x 12 threads,
// marching in z), viscous, multiply form, with sources.
// ---------------------------------------------------------------------------
constexpr int BY3 = 12;
// E3: some extent is 3 (the free-slip copies of put_bc).  DP: Gdτ and dτ_Rho
// from Musτ (one array read instead of two; they only set the pseudo-time
// steps, so the correction's fixed point does not depend on their rounding).
// nchunk > 0: each XCD sweeps its own eighth of the tiles chunk by chunk
// (as k3_iter2: all XCDs on the same planes, different rows; 511^3:
// 4.31 -> 4.01 ms per sweep against the flat (chunk, tile) order dealt to
// the XCDs in contiguous runs, nchunk = 0).
template <bool E3, bool DP>
__global__ __launch_bounds__(64 * BY3, 6) void k3_dc(const Fields32 in, const Fields32 out, const StepScalars32 cs,
                                                  int Lz, int ntx, int nty, int nchunk) {
    __shared__ float sh[4][BY3][64];
    const int ntile = ntx * nty;
    int tile, chunk;
    if (nchunk > 0) {
        const int xcd = blockIdx.x & 7, seq = blockIdx.x >> 3;
        const int lo = (int)((int64_t)ntile * xcd / 8), cnt = (int)((int64_t)ntile * (xcd + 1) / 8) - lo;
        if (seq >= cnt * nchunk) return;
        tile = lo + seq % cnt;
        chunk = seq / cnt;
    } else {
        const int b = xcd_tile(blockIdx.x, gridDim.x);
        tile = b % ntile;
        chunk = b / ntile;
    }
    const int tx = tile % ntx, ty = tile / ntx;
    const int lane = threadIdx.x, w = threadIdx.y;
    const int c = tx * 62 - 1 + lane;
    const int d = ty * (BY3 - 2) - 1 + w;
    const int nx = in.nx, ny = in.ny, nz = in.nz, sx = in.sx, sxy = in.sxy;
    const int kb = chunk * Lz;
    const int ke = min(kb + Lz, nz);
    const bool own = lane >= 1 && lane <= 62 && w >= 1 && w <= BY3 - 2;
    const int wu = w > 0 ? w - 1 : 0, wd = w < BY3 - 1 ? w + 1 : BY3 - 1;
    // Row roles: the top halo row only supplies Pt and τyy of the row above
    // the tile (its Vy update), the bottom halo row τxy and τyz of the row
    // below; each loads only the arrays those need.
    const bool top = w == 0, bot = w == BY3 - 1, mid = !top && !bot;
    const bool cell = c >= 0 && c < nx && d >= 0 && d < ny;
    const bool cellU = c >= 0 && c < nx && d >= 1 && d <= ny;
    const bool vxv = c >= 0 && c <= nx && d >= 0 && d < ny;
    const bool vxU = c >= 0 && c <= nx && d >= 1 && d <= ny;
    const bool vyv = c >= 0 && c < nx && d >= 0 && d <= ny;
    const bool txx_cd = c >= 0 && c < nx && d >= 1 && d <= ny - 2;
    const bool tyy_cd = c >= 1 && c <= nx - 2 && d >= 0 && d <= ny - 1;
    const bool tzz_cd = c >= 1 && c <= nx - 2 && d >= 1 && d <= ny - 2;
    const bool txy_cd = c >= 1 && c <= nx - 1 && d >= 1 && d <= ny - 1;
    const bool txz_cd = c >= 1 && c <= nx - 1 && d >= 1 && d <= ny - 2;
    const bool tyz_cd = c >= 1 && c <= nx - 2 && d >= 1 && d <= ny - 1;
    const bool vx_in = c >= 1 && c <= nx - 1 && d >= 1 && d <= ny - 2;
    const bool vy_in = c >= 1 && c <= nx - 2 && d >= 1 && d <= ny - 1;
    const bool vz_in = c >= 1 && c <= nx - 2 && d >= 1 && d <= ny - 2;
    const float _dx = cs._dx, _dy = cs._dy, _dz = cs._dz, r = cs.r;

    float Vx0P = 0.f, Vy0P = 0.f, GdP = 0.f, MuP = 0.f, GdLP = 0.f, MuLP = 0.f, GdUP = 0.f, MuUP = 0.f;
    float RhP = 0.f, RhUP = 0.f;
    float Pt1P = 0.f, txx1P = 0.f, tyy1P = 0.f, tzz1P = 0.f, txy1P = 0.f, txz1P = 0.f, tyz1P = 0.f;
    float Pt1PU = 0.f, tyy1PU = 0.f, txy1PD = 0.f;
    const int k0 = max(kb - 1, 0), k1 = min(ke, nz - 1);
    const int o0 = c + sx * d + sxy * k0;
    float VzC = cell ? in.p[PTS_VZ][o0] : 0.f;
    float VzUC = cellU && !top ? in.p[PTS_VZ][o0 - sx] : 0.f;
    for (int k = k0; k <= k1; ++k) {
        const int o = c + sx * d + sxy * k;
        const bool kin2 = k >= 1 && k <= nz - 2;
        const bool kin1 = k >= 1 && k <= nz - 1;
        float Pt0 = 0.f, Gd = 0.f, Mu = 0.f, Rh = 0.f, Vy0n = 0.f, VzN = 0.f, sP = 0.f, sVz = 0.f;
        float tx0 = 0.f, ty0 = 0.f, tz0 = 0.f, txy0 = 0.f, txz0 = 0.f, tyz0 = 0.f;
        float Vx0c = 0.f, Vy0c = 0.f, GdU = 0.f, MuU = 0.f, RhU = 0.f, VxU = 0.f, VzUN = 0.f, sVx = 0.f, sVy = 0.f;
        if (cell) {
            if constexpr (DP) {
                const float M = in.p[PTS_MUSTAU][o];
                Gd = cs.kg * M;
                if (mid) Rh = cs.k1 / M;
            } else {
                Gd = in.p[PTS_GDTAU][o];
                if (mid) Rh = in.p[PTS_DTAU_RHO][o];
            }
            Mu = in.p[PTS_MUS][o];
            VzN = in.p[PTS_VZ][o + sxy];
            if (!bot) {
                Pt0 = in.p[PTS_PT][o];
                Vy0n = in.p[PTS_VY][o + sx];
                sP = in.p[PTS_DIVV][o];
            }
            if (mid) {
                sVz = in.p[PTS_RZ][o];
                if (k >= 1) sVy = in.p[PTS_RY][o - sxy];
            }
        }
        if (cellU && !top) {
            if constexpr (DP) {
                const float M = in.p[PTS_MUSTAU][o - sx];
                GdU = cs.kg * M;
                if (mid) RhU = cs.k1 / M;
            } else {
                GdU = in.p[PTS_GDTAU][o - sx];
                if (mid) RhU = in.p[PTS_DTAU_RHO][o - sx];
            }
            MuU = in.p[PTS_MUS][o - sx];
            VzUN = in.p[PTS_VZ][o - sx + sxy];
        }
        if (vxv) {
            Vx0c = in.p[PTS_VX][o];
            if (mid && k >= 1) sVx = in.p[PTS_RX][o - sxy];
        }
        if (vxU && !top) VxU = in.p[PTS_VX][o - sx];
        if (vyv) {
            Vy0c = in.p[PTS_VY][o];
            if (mid && !cell && k >= 1) sVy = in.p[PTS_RY][o - sxy];
        }
        const bool hxx = txx_cd && kin2, hyy = tyy_cd && kin2, hzz = tzz_cd;
        const bool hxy = txy_cd && kin2, hxz = txz_cd && kin1, hyz = tyz_cd && kin1;
        const int oe = o - 1 - sx - sxy;
        if (hxx && mid) tx0 = in.p[PTS_TXX][o - sx - sxy];
        if (hyy && !bot) ty0 = in.p[PTS_TYY][o - 1 - sxy];
        if (hzz && mid) tz0 = in.p[PTS_TZZ][o - 1 - sx];
        if (hxy && !top) txy0 = in.p[PTS_TXY][oe];
        if (hxz && mid) txz0 = in.p[PTS_TXZ][oe];
        if (hyz && !top) tyz0 = in.p[PTS_TYZ][oe];

        const float VxR = fright(Vx0c);
        const float GdL = fleft(Gd), MuL = fleft(Mu), GdUL = fleft(GdU), MuUL = fleft(MuU);
        const float VyL = fleft(Vy0c), VzL = fleft(VzC);
        const float Pt1PL = fleft(Pt1P), txx1PL = fleft(txx1P), RhPL = fleft(RhP);
        const float txy1PR = fright(txy1P);

        const float ex = _dx * (VxR - Vx0c), ey = _dy * (Vy0n - Vy0c), ez = _dz * (VzN - VzC);
        const float Pt1n = Pt0 - r * Gd * (ex + ey + ez + sP);
        const float txx1n = tau_visc_f(tx0, Gd, Mu, ex);
        const float tyy1n = tau_visc_f(ty0, Gd, Mu, ey);
        const float tzz1n = tau_visc_f(tz0, Gd, Mu, ez);
        const float txy1n = tau_visc_f(txy0, (GdUL + GdU + GdL + Gd) * 0.25f, (MuUL + MuU + MuL + Mu) * 0.25f,
                                       0.5f * (_dy * (Vx0c - VxU) + _dx * (Vy0c - VyL)));
        const float txz1n = tau_visc_f(txz0, (GdLP + GdP + GdL + Gd) * 0.25f, (MuLP + MuP + MuL + Mu) * 0.25f,
                                       0.5f * (_dz * (Vx0c - Vx0P) + _dx * (VzC - VzL)));
        const float tyz1n = tau_visc_f(tyz0, (GdUP + GdP + GdU + Gd) * 0.25f, (MuUP + MuP + MuU + Mu) * 0.25f,
                                       0.5f * (_dz * (Vy0c - Vy0P) + _dy * (VzC - VzUC)));
        const float txz1R = fright(txz1n);
        sh[0][w][lane] = Pt1n;
        sh[1][w][lane] = tyy1n;
        sh[2][w][lane] = txy1n;
        sh[3][w][lane] = tyz1n;
        __syncthreads();
        const float Pt1U = sh[0][wu][lane], tyy1U = sh[1][wu][lane];
        const float txy1D = sh[2][wd][lane], tyz1D = sh[3][wd][lane];
        __syncthreads();
        const bool kown = k >= kb && k < ke;
        if (own && kown) {
            if (cell) out.p[PTS_PT][o] = Pt1n;
            if (hxx) out.p[PTS_TXX][o - sx - sxy] = txx1n;
            if (hyy) out.p[PTS_TYY][o - 1 - sxy] = tyy1n;
            if (hzz) out.p[PTS_TZZ][o - 1 - sx] = tzz1n;
            if (hxy) out.p[PTS_TXY][oe] = txy1n;
            if (hxz) out.p[PTS_TXZ][oe] = txz1n;
            if (hyz) out.p[PTS_TYZ][oe] = tyz1n;
            if (vz_in && k >= 1) {
                float R = _dz * (tzz1n - tzz1P) + _dx * (txz1R - txz1n) + _dy * (tyz1D - tyz1n) - _dz * (Pt1n - Pt1P) + sVz;
                put_bc<E3>(out.p[PTS_VZ], o, VzC + R * ((RhP + Rh) * 0.5f), c, nx, 1, d, ny, sx);
            }
        }
        const int km = k - 1;
        if (own && km >= kb && km < ke && km >= 1 && km <= nz - 2) {
            const int q = o - sxy;
            if (vx_in) {
                float R = _dx * (txx1P - txx1PL) + _dy * (txy1PD - txy1P) + _dz * (txz1n - txz1P) - _dx * (Pt1P - Pt1PL) + sVx;
                put_bc<E3>(out.p[PTS_VX], q, Vx0P + R * ((RhPL + RhP) * 0.5f), d, ny, sx, km, nz, sxy);
            }
            if (vy_in) {
                float R = _dy * (tyy1P - tyy1PU) + _dx * (txy1PR - txy1P) + _dz * (tyz1n - tyz1P) - _dy * (Pt1P - Pt1PU) + sVy;
                put_bc<E3>(out.p[PTS_VY], q, Vy0P + R * ((RhUP + RhP) * 0.5f), c, nx, 1, km, nz, sxy);
            }
        }
        Vx0P = Vx0c; Vy0P = Vy0c; VzC = VzN; VzUC = VzUN;
        GdLP = GdL; MuLP = MuL; GdUP = GdU; MuUP = MuU; GdP = Gd; MuP = Mu; RhUP = RhU; RhP = Rh;
        Pt1P = Pt1n; txx1P = txx1n; tyy1P = tyy1n; tzz1P = tzz1n; txy1P = txy1n; txz1P = txz1n; tyz1P = tyz1n;
        Pt1PU = Pt1U; tyy1PU = tyy1U; txy1PD = txy1D;
    }
}

// ---------------------------------------------------------------------------
// FP64 helpers of the outer loop
// ---------------------------------------------------------------------------
template <int ND>
__global__ __launch_bounds__(256) void k_dc_project(Fields f, StepScalars c) {
    // τ = 2 μ ε̇(V) on every stress point (multiply form)
    const int nx = f.nx, ny = f.ny, nz = ND == 3 ? f.nz : 1, sx = f.sx, sxy = f.sxy;
    const int64_t n1 = nx + 1, n2 = ny + 1, tot = n1 * n2 * (ND == 3 ? nz + 1 : 1);
    const double* Vx = f.p[PTS_VX];
    const double* Vy = f.p[PTS_VY];
    const double* Vz = f.p[PTS_VZ];
    const double* Mu = f.p[PTS_MUS];
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % n1);
        const int j = (int)((q / n1) % n2);
        const int k = (int)(q / (n1 * n2));
        const int o = i + sx * j + sxy * k;
        if (ND == 2) {
            if (i < nx && j < ny) {
                f.p[PTS_TXX][o] = 2.0 * Mu[o] * (c._dx * (Vx[o + 1] - Vx[o]));
                f.p[PTS_TYY][o] = 2.0 * Mu[o] * (c._dy * (Vy[o + sx] - Vy[o]));
            }
            if (i < nx - 1 && j < ny - 1) {
                double am = 0.25 * (Mu[o] + Mu[o + 1] + Mu[o + sx] + Mu[o + sx + 1]);
                f.p[PTS_TXY][o] = 2.0 * am * (0.5 * (c._dy * (Vx[o + sx + 1] - Vx[o + 1]) + c._dx * (Vy[o + sx + 1] - Vy[o + sx])));
            }
        } else {
            if (i < nx && j < ny - 2 && k < nz - 2) {
                const int a = o + sx + sxy;
                f.p[PTS_TXX][o] = 2.0 * Mu[a] * (c._dx * (Vx[a + 1] - Vx[a]));
            }
            if (i < nx - 2 && j < ny && k < nz - 2) {
                const int a = o + 1 + sxy;
                f.p[PTS_TYY][o] = 2.0 * Mu[a] * (c._dy * (Vy[a + sx] - Vy[a]));
            }
            if (i < nx - 2 && j < ny - 2 && k < nz) {
                const int a = o + 1 + sx;
                f.p[PTS_TZZ][o] = 2.0 * Mu[a] * (c._dz * (Vz[a + sxy] - Vz[a]));
            }
            if (i < nx - 1 && j < ny - 1 && k < nz - 2) {
                const int b = o + sxy;
                double am = (Mu[b] + Mu[b + 1] + Mu[b + sx] + Mu[b + sx + 1]) * 0.25;
                f.p[PTS_TXY][o] = 2.0 * am * (0.5 * (c._dy * (Vx[b + sx + 1] - Vx[b + 1]) + c._dx * (Vy[b + sx + 1] - Vy[b + sx])));
            }
            if (i < nx - 1 && j < ny - 2 && k < nz - 1) {
                const int b = o + sx;
                double am = (Mu[b] + Mu[b + 1] + Mu[b + sxy] + Mu[b + sxy + 1]) * 0.25;
                f.p[PTS_TXZ][o] = 2.0 * am * (0.5 * (c._dz * (Vx[b + sxy + 1] - Vx[b + 1]) + c._dx * (Vz[b + sxy + 1] - Vz[b + sxy])));
            }
            if (i < nx - 2 && j < ny - 1 && k < nz - 1) {
                const int b = o + 1;
                double am = (Mu[b] + Mu[b + sx] + Mu[b + sxy] + Mu[b + sx + sxy]) * 0.25;
                f.p[PTS_TYZ][o] = 2.0 * am * (0.5 * (c._dz * (Vy[b + sx + sxy] - Vy[b + sx]) + c._dy * (Vz[b + sx + sxy] - Vz[b + sxy])));
            }
        }
    }
}

// FP32 sources: R_P at cells, R_V at the interior velocity points (stored at
// the velocity index of the face), zero elsewhere.
template <int ND>
__global__ __launch_bounds__(256) void k_dc_sources(Fields f, Fields32 g, StepScalars c, int bf) {
    const int nx = f.nx, ny = f.ny, nz = ND == 3 ? f.nz : 1, sx = f.sx, sxy = f.sxy;
    const int64_t n1 = nx + 1, n2 = ny + 1, tot = n1 * n2 * (ND == 3 ? nz + 1 : 1);
    const double* Vx = f.p[PTS_VX];
    const double* Vy = f.p[PTS_VY];
    const double* Vz = f.p[PTS_VZ];
    const double* Pt = f.p[PTS_PT];
    const double* rg = f.p[PTS_RHOG];
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % n1);
        const int j = (int)((q / n1) % n2);
        const int k = (int)(q / (n1 * n2));
        const int o = i + sx * j + sxy * k;
        if (ND == 2) {
            if (i < nx && j < ny)
                g.p[PTS_DIVV][o] = (float)(c._dx * (Vx[o + 1] - Vx[o]) + c._dy * (Vy[o + sx] - Vy[o]));
            float sx_ = 0.f, sy_ = 0.f;
            if (i >= 1 && i <= nx - 1 && j >= 1 && j <= ny - 2) {
                const int a = o - 1 - sx;  // residual index (i-1, j-1)
                const double* txx = f.p[PTS_TXX];
                const double* txy = f.p[PTS_TXY];
                sx_ = (float)(c._dx * (txx[a + sx + 1] - txx[a + sx]) + c._dy * (txy[a + sx] - txy[a]) -
                              c._dx * (Pt[a + sx + 1] - Pt[a + sx]));
            }
            if (i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 1) {
                const int a = o - 1 - sx;
                const double* tyy = f.p[PTS_TYY];
                const double* txy = f.p[PTS_TXY];
                double R = c._dy * (tyy[a + sx + 1] - tyy[a + 1]) + c._dx * (txy[a + 1] - txy[a]) -
                           c._dy * (Pt[a + sx + 1] - Pt[a + 1]);
                if (bf) R -= 0.5 * (rg[a + 1] + rg[a + sx + 1]);
                sy_ = (float)R;
            }
            if (i <= nx && j < ny) g.p[PTS_RX][o] = sx_;
            if (i < nx && j <= ny) g.p[PTS_RY][o] = sy_;
        } else {
            if (i < nx && j < ny && k < nz)
                g.p[PTS_DIVV][o] = (float)(c._dx * (Vx[o + 1] - Vx[o]) + c._dy * (Vy[o + sx] - Vy[o]) +
                                           c._dz * (Vz[o + sxy] - Vz[o]));
            const double* txx = f.p[PTS_TXX];
            const double* tyy = f.p[PTS_TYY];
            const double* tzz = f.p[PTS_TZZ];
            const double* txy = f.p[PTS_TXY];
            const double* txz = f.p[PTS_TXZ];
            const double* tyz = f.p[PTS_TYZ];
            const int a = o - 1 - sx - sxy;  // residual index (i-1, j-1, k-1)
            float rx = 0.f, ry = 0.f, rz = 0.f;
            if (i >= 1 && i <= nx - 1 && j >= 1 && j <= ny - 2 && k >= 1 && k <= nz - 2)
                rx = (float)(c._dx * (txx[a + 1] - txx[a]) + c._dy * (txy[a + sx] - txy[a]) + c._dz * (txz[a + sxy] - txz[a]) -
                             c._dx * (Pt[a + 1 + sx + sxy] - Pt[a + sx + sxy]));
            if (i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 1 && k >= 1 && k <= nz - 2)
                ry = (float)(c._dy * (tyy[a + sx] - tyy[a]) + c._dx * (txy[a + 1] - txy[a]) + c._dz * (tyz[a + sxy] - tyz[a]) -
                             c._dy * (Pt[a + 1 + sx + sxy] - Pt[a + 1 + sxy]));
            if (i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2 && k >= 1 && k <= nz - 1) {
                double R = c._dz * (tzz[a + sxy] - tzz[a]) + c._dx * (txz[a + 1] - txz[a]) + c._dy * (tyz[a + sx] - tyz[a]) -
                           c._dz * (Pt[a + 1 + sx + sxy] - Pt[a + 1 + sx]);
                if (bf) R -= 0.5 * (rg[a + 1 + sx] + rg[a + 1 + sx + sxy]);
                rz = (float)R;
            }
            if (i <= nx && j < ny && k < nz) g.p[PTS_RX][o] = rx;
            if (i < nx && j <= ny && k < nz) g.p[PTS_RY][o] = ry;
            if (i < nx && j < ny && k <= nz) g.p[PTS_RZ][o] = rz;
        }
    }
}

// P += δP, V += δV (whole arrays; δ is zero on Dirichlet entries).
template <int ND>
__global__ __launch_bounds__(256) void k_dc_accumulate(Fields f, Fields32 g) {
    const int nx = f.nx, ny = f.ny, nz = ND == 3 ? f.nz : 1, sx = f.sx, sxy = f.sxy;
    const int64_t n1 = nx + 1, n2 = ny + 1, tot = n1 * n2 * (ND == 3 ? nz + 1 : 1);
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % n1);
        const int j = (int)((q / n1) % n2);
        const int k = (int)(q / (n1 * n2));
        const int o = i + sx * j + sxy * k;
        if (i < nx && j < ny && k < nz) f.p[PTS_PT][o] += (double)g.p[PTS_PT][o];
        if (i <= nx && j < ny && k < nz) f.p[PTS_VX][o] += (double)g.p[PTS_VX][o];
        if (i < nx && j <= ny && k < nz) f.p[PTS_VY][o] += (double)g.p[PTS_VY][o];
        if (ND == 3 && i < nx && j < ny && k <= nz) f.p[PTS_VZ][o] += (double)g.p[PTS_VZ][o];
    }
}

// FP32 copies of the constant fields (Mus, Gdτ, dτ_Rho).
__global__ __launch_bounds__(256) void k_to_f32(float* dst, const double* src, int64_t n) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        dst[q] = (float)src[q];
}

// Correction-problem residual norms (FP32 data, FP64 accumulation):
// ∇·δV + R_P and ∇·δτ - ∇δP + R_V; out[0..3] = Σ², out[4..7] = max|.|
// for Rx, Ry, Rz, ∇V.
// NaN-propagating max (fmax drops a NaN operand)
__device__ __forceinline__ double nanmax(double a, double b) { return (a != a) ? a : ((b != b) ? b : fmax(a, b)); }

template <int ND>
__global__ __launch_bounds__(256) void k_dc_norms(Fields32 g, StepScalars32 c, double* partials) {
    __shared__ double sm[8][256];
    double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int nx = g.nx, ny = g.ny, nz = ND == 3 ? g.nz : 1, sx = g.sx, sxy = g.sxy;
    const int64_t n1 = nx + 1, n2 = ny + 1, tot = n1 * n2 * (ND == 3 ? nz + 1 : 1);
    const float* Vx = g.p[PTS_VX];
    const float* Vy = g.p[PTS_VY];
    const float* Vz = g.p[PTS_VZ];
    const float* Pt = g.p[PTS_PT];
    auto acc = [&](int comp, double v) {
        a[comp] += v * v;
        a[4 + comp] = nanmax(a[4 + comp], fabs(v));
    };
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % n1);
        const int j = (int)((q / n1) % n2);
        const int k = (int)(q / (n1 * n2));
        const int o = i + sx * j + sxy * k;
        if (ND == 2) {
            if (i < nx && j < ny)
                acc(3, (double)(c._dx * (Vx[o + 1] - Vx[o]) + c._dy * (Vy[o + sx] - Vy[o]) + g.p[PTS_DIVV][o]));
            const float* txx = g.p[PTS_TXX];
            const float* tyy = g.p[PTS_TYY];
            const float* txy = g.p[PTS_TXY];
            if (i < nx - 1 && j < ny - 2)
                acc(0, (double)(c._dx * (txx[o + sx + 1] - txx[o + sx]) + c._dy * (txy[o + sx] - txy[o]) -
                                c._dx * (Pt[o + sx + 1] - Pt[o + sx]) + g.p[PTS_RX][o + 1 + sx]));
            if (i < nx - 2 && j < ny - 1)
                acc(1, (double)(c._dy * (tyy[o + sx + 1] - tyy[o + 1]) + c._dx * (txy[o + 1] - txy[o]) -
                                c._dy * (Pt[o + sx + 1] - Pt[o + 1]) + g.p[PTS_RY][o + 1 + sx]));
        } else {
            if (i < nx && j < ny && k < nz)
                acc(3, (double)(c._dx * (Vx[o + 1] - Vx[o]) + c._dy * (Vy[o + sx] - Vy[o]) + c._dz * (Vz[o + sxy] - Vz[o]) +
                                g.p[PTS_DIVV][o]));
            const float* txx = g.p[PTS_TXX];
            const float* tyy = g.p[PTS_TYY];
            const float* tzz = g.p[PTS_TZZ];
            const float* txy = g.p[PTS_TXY];
            const float* txz = g.p[PTS_TXZ];
            const float* tyz = g.p[PTS_TYZ];
            const int v = o + 1 + sx + sxy;
            if (i < nx - 1 && j < ny - 2 && k < nz - 2)
                acc(0, (double)(c._dx * (txx[o + 1] - txx[o]) + c._dy * (txy[o + sx] - txy[o]) + c._dz * (txz[o + sxy] - txz[o]) -
                                c._dx * (Pt[o + 1 + sx + sxy] - Pt[o + sx + sxy]) + g.p[PTS_RX][v]));
            if (i < nx - 2 && j < ny - 1 && k < nz - 2)
                acc(1, (double)(c._dy * (tyy[o + sx] - tyy[o]) + c._dx * (txy[o + 1] - txy[o]) + c._dz * (tyz[o + sxy] - tyz[o]) -
                                c._dy * (Pt[o + 1 + sx + sxy] - Pt[o + 1 + sxy]) + g.p[PTS_RY][v]));
            if (i < nx - 2 && j < ny - 2 && k < nz - 1)
                acc(2, (double)(c._dz * (tzz[o + sxy] - tzz[o]) + c._dx * (txz[o + 1] - txz[o]) + c._dy * (tyz[o + sx] - tyz[o]) -
                                c._dz * (Pt[o + 1 + sx + sxy] - Pt[o + 1 + sx]) + g.p[PTS_RZ][v]));
        }
    }
    const int t = threadIdx.x;
    for (int m = 0; m < 8; ++m) sm[m][t] = a[m];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            for (int m = 0; m < 4; ++m) sm[m][t] += sm[m][t + w];
            for (int m = 4; m < 8; ++m) sm[m][t] = nanmax(sm[m][t], sm[m][t + w]);
        }
        __syncthreads();
    }
    if (t < 8) partials[(int64_t)blockIdx.x * 8 + t] = sm[t][0];
}

// 256 threads: each folds every 256th block, then a fixed-order tree
// (deterministic sums; NaN-propagating maxima).
__global__ __launch_bounds__(256) void k_dc_norms_final(const double* partials, int nb, double* out) {
    __shared__ double sm[8][256];
    const int t = threadIdx.x;
    double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = t; b < nb; b += 256) {
        for (int m = 0; m < 4; ++m) a[m] += partials[(int64_t)b * 8 + m];
        for (int m = 4; m < 8; ++m) a[m] = nanmax(a[m], partials[(int64_t)b * 8 + m]);
    }
    for (int m = 0; m < 8; ++m) sm[m][t] = a[m];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            for (int m = 0; m < 4; ++m) sm[m][t] += sm[m][t + w];
            for (int m = 4; m < 8; ++m) sm[m][t] = nanmax(sm[m][t], sm[m][t + w]);
        }
        __syncthreads();
    }
    if (t < 8) out[t] = sm[t][0];
}

inline int gridN(int64_t tot) { return (int)std::min<int64_t>(65536, std::max<int64_t>(1, (tot + 255) / 256)); }

}  // namespace

// Resident blocks of `kern` with `threads` threads on the whole device.
template <typename K>
static int device_concurrency(K kern, int threads, int& cached) {
    if (!cached) {
        int dev = 0, cus = 0, per = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, threads, 0);
        cached = std::max(1, cus) * std::max(1, per);
    }
    return cached;
}

// Rows (2D) or planes (3D) per block: about one round of resident blocks,
// within [4, cap], so small grids still fill the device.
static int march_length(int n, int columns, int conc, int cap) {
    const int nch = std::max(1, conc / std::max(1, columns));
    return std::max(4, std::min((n + nch - 1) / nch, cap));
}

hipError_t launch_dc_iter(hipStream_t s, const Fields32& in, const Fields32& out, const StepScalars32& c, int ndim,
                          int lz, bool dp, bool map_xcd) {
    static int conc2 = 0, conc3 = 0;
    if (ndim == 2) {
        const int ns = (in.nx + 1 + 61) / 62;
        const int L = march_length(in.ny, ns, device_concurrency(k2_dc, 64, conc2), 128);
        const int nb = ns * ((in.ny + L - 1) / L);
        k2_dc<<<nb, 64, 0, s>>>(in, out, c, L, ns);
    } else {
        const int ntx = (in.nx + 1 + 61) / 62, nty = (in.ny + 1 + BY3 - 3) / (BY3 - 2);
        const int64_t conc = device_concurrency(k3_dc<false, false>, 64 * BY3, conc3);
        const dim3 block(64, BY3);
        const bool e3 = in.nx == 3 || in.ny == 3 || in.nz == 3;
        int Lz, nchk;
        dim3 grid;
        if (map_xcd) {  // per XCD: ceil(tiles / 8) tiles per chunk on an eighth of the resident blocks
            const int64_t per = ((int64_t)ntx * nty + 7) / 8;
            Lz = planes_per_block(lz, in.nz, per, std::max<int64_t>(1, conc / 8));
            nchk = (in.nz + Lz - 1) / Lz;
            grid = dim3((unsigned)(8 * per * nchk));
        } else {
            Lz = planes_per_block(lz, in.nz, (int64_t)ntx * nty, conc);
            grid = dim3((unsigned)(ntx * nty * ((in.nz + Lz - 1) / Lz)));
            nchk = 0;
        }
        if (e3 && dp) k3_dc<true, true><<<grid, block, 0, s>>>(in, out, c, Lz, ntx, nty, nchk);
        else if (e3) k3_dc<true, false><<<grid, block, 0, s>>>(in, out, c, Lz, ntx, nty, nchk);
        else if (dp) k3_dc<false, true><<<grid, block, 0, s>>>(in, out, c, Lz, ntx, nty, nchk);
        else k3_dc<false, false><<<grid, block, 0, s>>>(in, out, c, Lz, ntx, nty, nchk);
    }
    return hipGetLastError();
}

hipError_t launch_dc_project(hipStream_t s, const Fields& f, const StepScalars& c, int ndim) {
    const int64_t tot = (int64_t)(f.nx + 1) * (f.ny + 1) * (ndim == 3 ? f.nz + 1 : 1);
    if (ndim == 2) k_dc_project<2><<<gridN(tot), 256, 0, s>>>(f, c);
    else k_dc_project<3><<<gridN(tot), 256, 0, s>>>(f, c);
    return hipGetLastError();
}

hipError_t launch_dc_sources(hipStream_t s, const Fields& f, const Fields32& g, const StepScalars& c, int ndim,
                             bool bf) {
    const int64_t tot = (int64_t)(f.nx + 1) * (f.ny + 1) * (ndim == 3 ? f.nz + 1 : 1);
    if (ndim == 2) k_dc_sources<2><<<gridN(tot), 256, 0, s>>>(f, g, c, bf ? 1 : 0);
    else k_dc_sources<3><<<gridN(tot), 256, 0, s>>>(f, g, c, bf ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_dc_accumulate(hipStream_t s, const Fields& f, const Fields32& g, int ndim) {
    const int64_t tot = (int64_t)(f.nx + 1) * (f.ny + 1) * (ndim == 3 ? f.nz + 1 : 1);
    if (ndim == 2) k_dc_accumulate<2><<<gridN(tot), 256, 0, s>>>(f, g);
    else k_dc_accumulate<3><<<gridN(tot), 256, 0, s>>>(f, g);
    return hipGetLastError();
}

hipError_t launch_to_f32(hipStream_t s, float* dst, const double* src, int64_t n) {
    k_to_f32<<<gridN(n), 256, 0, s>>>(dst, src, n);
    return hipGetLastError();
}

hipError_t launch_dc_norms(hipStream_t s, const Fields32& g, const StepScalars32& c, int ndim, double* partials,
                           double* out) {
    if (ndim == 2) k_dc_norms<2><<<kRedBlocks, 256, 0, s>>>(g, c, partials);
    else k_dc_norms<3><<<kRedBlocks, 256, 0, s>>>(g, c, partials);
    k_dc_norms_final<<<1, 256, 0, s>>>(partials, kRedBlocks, out);
    return hipGetLastError();
}

}  // namespace pts
