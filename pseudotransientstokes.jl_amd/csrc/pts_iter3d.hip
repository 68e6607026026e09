// pts_iter3d.hip -- 3D single-pass PT iteration.
//
// One block of 64 x BY threads owns a tile of 62 x (BY-2) output columns in
// the x-y plane (lanes 0/63 and rows 0/BY-1 are halo) and a chunk of z
// planes, marching in z.  Thread (c, d) at plane step k computes
//   * the new cell values Pt, τxx, τyy, τzz at (c, d, k)          (compute_Pt_τ!)
//   * the new edge stresses τxy at (x-face c, y-face d, cell k),
//     τxz at (x-face c, cell d, z-face k), τyz at (cell c, y-face d, z-face k)
//   * Vz at z-face k, and Vx, Vy at plane k-1                      (compute_V!)
// with the six free-slip copies folded into the V writes
// (Stokes3D_ve_perf.jl:69-88, 202-207).  x-neighbours come from DPP lane
// shifts, y-neighbours of computed values through LDS (two barriers per
// plane), y-neighbours of read-only inputs from the cache, z-neighbours from
// registers.  State is read from `in` and written to `out` (ping-pong); the
// never-updated Dirichlet entries are held by both buffers.
#include <type_traits>

#include "pts_internal.hpp"

namespace pts {

namespace {

template <bool VE, bool MUL, bool BF, int BY, bool DP>
__global__ __launch_bounds__(64 * BY) void k3_iter(const Fields in, const Fields out, const StepScalars cs, int Lz,
                                                   int ntx, int nty) {
    __shared__ double sh[4][BY][64];  // Pt1, τyy1 (to row +1); τxy1, τyz1 (to row -1)

    int b = blockIdx.x;
    const int nb = gridDim.x;
    b = xcd_tile(b, nb);  // contiguous tiles per XCD
    const int ntile = ntx * nty;
    const int tile = b % ntile, chunk = b / ntile;
    const int tx = tile % ntx, ty = tile / ntx;
    const int lane = threadIdx.x, w = threadIdx.y;
    const int c = tx * 62 - 1 + lane;
    const int d = ty * (BY - 2) - 1 + w;
    const int nx = in.nx, ny = in.ny, nz = in.nz, sx = in.sx, sxy = in.sxy;
    const int kb = chunk * Lz;
    const int ke = min(kb + Lz, nz);
    const bool own = lane >= 1 && lane <= 62 && w >= 1 && w <= BY - 2;
    const int wu = w > 0 ? w - 1 : 0, wd = w < BY - 1 ? w + 1 : BY - 1;
    // The halo rows (one wave each) only feed their neighbour row: the top row
    // its new Pt and τyy, the bottom row its new τxy and τyz.  They skip the
    // loads those values do not need (wave-uniform branches).
    const bool top = w == 0, bot = w == BY - 1, mid = !top && !bot;

    const bool cell = c >= 0 && c < nx && d >= 0 && d < ny;
    const bool cellU = c >= 0 && c < nx && d >= 1 && d <= ny;  // cell (c, d-1)
    const bool vxv = c >= 0 && c <= nx && d >= 0 && d < ny;
    const bool vxU = c >= 0 && c <= nx && d >= 1 && d <= ny;   // Vx (c, d-1)
    const bool vyv = c >= 0 && c < nx && d >= 0 && d <= ny;
    const bool txx_cd = c >= 0 && c < nx && d >= 1 && d <= ny - 2;
    const bool tyy_cd = c >= 1 && c <= nx - 2 && d >= 0 && d <= ny - 1;
    const bool tzz_cd = c >= 1 && c <= nx - 2 && d >= 1 && d <= ny - 2;
    const bool txy_cd = c >= 1 && c <= nx - 1 && d >= 1 && d <= ny - 1;
    const bool txz_cd = c >= 1 && c <= nx - 1 && d >= 1 && d <= ny - 2;
    const bool tyz_cd = c >= 1 && c <= nx - 2 && d >= 1 && d <= ny - 1;
    // interior velocity columns and Dirichlet faces
    const bool vx_in = c >= 1 && c <= nx - 1 && d >= 1 && d <= ny - 2;
    const bool vx_dir = (c == 0 || c == nx) && d >= 1 && d <= ny - 2;
    const bool vy_in = c >= 1 && c <= nx - 2 && d >= 1 && d <= ny - 1;
    const bool vy_dir = c >= 1 && c <= nx - 2 && (d == 0 || d == ny);
    const bool vz_in = c >= 1 && c <= nx - 2 && d >= 1 && d <= ny - 2;

    const double* __restrict__ Pt0 = in.p[PTS_PT];
    const double* __restrict__ txx0 = in.p[PTS_TXX];
    const double* __restrict__ tyy0 = in.p[PTS_TYY];
    const double* __restrict__ tzz0 = in.p[PTS_TZZ];
    const double* __restrict__ txy0 = in.p[PTS_TXY];
    const double* __restrict__ txz0 = in.p[PTS_TXZ];
    const double* __restrict__ tyz0 = in.p[PTS_TYZ];
    const double* __restrict__ txxo = in.p[PTS_TXX_O];
    const double* __restrict__ tyyo = in.p[PTS_TYY_O];
    const double* __restrict__ tzzo = in.p[PTS_TZZ_O];
    const double* __restrict__ txyo = in.p[PTS_TXY_O];
    const double* __restrict__ txzo = in.p[PTS_TXZ_O];
    const double* __restrict__ tyzo = in.p[PTS_TYZ_O];
    const double* __restrict__ Vx0 = in.p[PTS_VX];
    const double* __restrict__ Vy0 = in.p[PTS_VY];
    const double* __restrict__ Vz0 = in.p[PTS_VZ];
    const double* __restrict__ Mtp = in.p[PTS_MUSTAU];
    const double* __restrict__ Gdp = in.p[PTS_GDTAU];
    const double* __restrict__ Mup = in.p[PTS_MUS];
    const double* __restrict__ Rhp = in.p[PTS_DTAU_RHO];
    const double* __restrict__ rgp = in.p[PTS_RHOG];
    double* __restrict__ Pt1 = out.p[PTS_PT];
    double* __restrict__ txx1 = out.p[PTS_TXX];
    double* __restrict__ tyy1 = out.p[PTS_TYY];
    double* __restrict__ tzz1 = out.p[PTS_TZZ];
    double* __restrict__ txy1 = out.p[PTS_TXY];
    double* __restrict__ txz1 = out.p[PTS_TXZ];
    double* __restrict__ tyz1 = out.p[PTS_TYZ];
    double* __restrict__ Vx1 = out.p[PTS_VX];
    double* __restrict__ Vy1 = out.p[PTS_VY];
    double* __restrict__ Vz1 = out.p[PTS_VZ];

    const double dx = cs.dx, dy = cs.dy, dz = cs.dz, _dx = cs._dx, _dy = cs._dy, _dz = cs._dz;
    const double r = cs.r, Gdt = cs.Gdt;

    // values carried from plane k-1
    double Vx0P = 0.0, Vy0P = 0.0, GdP = 0.0, MuP = 0.0, GdLP = 0.0, MuLP = 0.0, GdUP = 0.0, MuUP = 0.0;
    double RhP = 0.0, RhUP = 0.0, rgP = 0.0;
    double Pt1P = 0.0, txx1P = 0.0, tyy1P = 0.0, tzz1P = 0.0, txy1P = 0.0, txz1P = 0.0, tyz1P = 0.0;
    double Pt1PU = 0.0, tyy1PU = 0.0, txy1PD = 0.0;
    const int k0 = max(kb - 1, 0), k1 = min(ke, nz - 1);
    const int o0 = c + sx * d + sxy * k0;
    double VzC = cell ? Vz0[o0] : 0.0;   // Vz0 at z-face k
    double VzUC = cellU ? Vz0[o0 - sx] : 0.0;

    for (int k = k0; k <= k1; ++k) {
        const int o = c + sx * d + sxy * k;
        const bool kin2 = k >= 1 && k <= nz - 2;  // plane index valid for τxx/τyy/τxy
        const bool kin1 = k >= 1 && k <= nz - 1;  // z-face index valid for τxz/τyz
        // ---- loads of plane k
        double Pt0v = 0.0, Gd = 0.0, Mu = 0.0, Rh = 0.0, Vy0n = 0.0, VzN = 0.0, rg = 0.0;
        double tx0 = 0.0, ty0 = 0.0, tz0 = 0.0, txo = 0.0, tyo = 0.0, tzo = 0.0;
        double txy0v = 0.0, txz0v = 0.0, tyz0v = 0.0, txyov = 0.0, txzov = 0.0, tyzov = 0.0;
        double Vx0c = 0.0, Vy0c = 0.0, GdU = 0.0, MuU = 0.0, RhU = 0.0, VxU = 0.0, VzUN = 0.0;
        if (cell) {
            if constexpr (DP) {
                iter_params_of(Mtp[o], cs, Rh, Gd);
            } else {
                Gd = Gdp[o];
                if (mid) Rh = Rhp[o];
            }
            Mu = Mup[o];
            VzN = Vz0[o + sxy];
            if (!bot) {
                Pt0v = Pt0[o];
                Vy0n = Vy0[o + sx];
            }
            if (mid) {
                if constexpr (BF) rg = rgp[o];
            }
        }
        if (cellU && !top) {
            if constexpr (DP) {
                iter_params_of(Mtp[o - sx], cs, RhU, GdU);
            } else {
                GdU = Gdp[o - sx];
                if (mid) RhU = Rhp[o - sx];
            }
            MuU = Mup[o - sx];
            VzUN = Vz0[o - sx + sxy];
        }
        if (vxv) Vx0c = Vx0[o];
        if (vxU && !top) VxU = Vx0[o - sx];
        if (vyv) Vy0c = Vy0[o];
        const bool hxx = txx_cd && kin2, hyy = tyy_cd && kin2, hzz = tzz_cd;
        const bool hxy = txy_cd && kin2, hxz = txz_cd && kin1, hyz = tyz_cd && kin1;
        const int oe = o - 1 - sx - sxy;  // edge-stress index (c-1, d-1, k-1)
        if (hxx && mid) {
            tx0 = txx0[o - sx - sxy];
            if constexpr (VE) txo = txxo[o - sx - sxy];
        }
        if (hyy && !bot) {
            ty0 = tyy0[o - 1 - sxy];
            if constexpr (VE) tyo = tyyo[o - 1 - sxy];
        }
        if (hzz && mid) {
            tz0 = tzz0[o - 1 - sx];
            if constexpr (VE) tzo = tzzo[o - 1 - sx];
        }
        if (hxy && !top) {
            txy0v = txy0[oe];
            if constexpr (VE) txyov = txyo[oe];
        }
        if (hxz && mid) {
            txz0v = txz0[oe];
            if constexpr (VE) txzov = txzo[oe];
        }
        if (hyz && !top) {
            tyz0v = tyz0[oe];
            if constexpr (VE) tyzov = tyzo[oe];
        }

        // ---- lane exchanges (every lane active)
        const double VxR = from_right(Vx0c);     // Vx0 (c+1, d, k)
        const double GdL = from_left(Gd);        // (c-1, d, k)
        const double MuL = from_left(Mu);
        const double GdUL = from_left(GdU);      // (c-1, d-1, k)
        const double MuUL = from_left(MuU);
        const double VyL = from_left(Vy0c);      // Vy0 (c-1, y-face d, k)
        const double VzL = from_left(VzC);       // Vz0 (c-1, d, z-face k)
        const double Pt1PL = from_left(Pt1P);    // (c-1, d, k-1)
        const double txx1PL = from_left(txx1P);
        const double RhPL = from_left(RhP);
        const double txy1PR = from_right(txy1P); // τxy (x-face c+1, y-face d, k-1)

        // ---- cell (c, d, k)
        const double ex = dfd<MUL>(VxR - Vx0c, dx, _dx);
        const double ey = dfd<MUL>(Vy0n - Vy0c, dy, _dy);
        const double ez = dfd<MUL>(VzN - VzC, dz, _dz);
        const double Pt1n = Pt0v - r * Gd * (ex + ey + ez);
        const double txx1n = tau_upd<VE>(tx0, txo, Gd, Mu, Gdt, ex);
        const double tyy1n = tau_upd<VE>(ty0, tyo, Gd, Mu, Gdt, ey);
        const double tzz1n = tau_upd<VE>(tz0, tzo, Gd, Mu, Gdt, ez);

        // ---- edges
        double txy1n, txz1n, tyz1n;
        {   // τxy at (x-face c, y-face d, cell k): cells (c-1..c) x (d-1..d) of plane k
            const double ag = (GdUL + GdU + GdL + Gd) * 0.25;
            const double am = (MuUL + MuU + MuL + Mu) * 0.25;
            const double e = 0.5 * (dfd<MUL>(Vx0c - VxU, dy, _dy) + dfd<MUL>(Vy0c - VyL, dx, _dx));
            txy1n = tau_upd<VE>(txy0v, txyov, ag, am, Gdt, e);
        }
        {   // τxz at (x-face c, cell d, z-face k): cells (c-1..c) x (k-1..k) of row d
            const double ag = (GdLP + GdP + GdL + Gd) * 0.25;
            const double am = (MuLP + MuP + MuL + Mu) * 0.25;
            const double e = 0.5 * (dfd<MUL>(Vx0c - Vx0P, dz, _dz) + dfd<MUL>(VzC - VzL, dx, _dx));
            txz1n = tau_upd<VE>(txz0v, txzov, ag, am, Gdt, e);
        }
        {   // τyz at (cell c, y-face d, z-face k): cells (d-1..d) x (k-1..k) of column c
            const double ag = (GdUP + GdP + GdU + Gd) * 0.25;
            const double am = (MuUP + MuP + MuU + Mu) * 0.25;
            const double e = 0.5 * (dfd<MUL>(Vy0c - Vy0P, dz, _dz) + dfd<MUL>(VzC - VzUC, dy, _dy));
            tyz1n = tau_upd<VE>(tyz0v, tyzov, ag, am, Gdt, e);
        }
        const double txz1R = from_right(txz1n);  // τxz (x-face c+1, d, z-face k)

        // ---- y-neighbour exchange of computed values
        sh[0][w][lane] = Pt1n;
        sh[1][w][lane] = tyy1n;
        sh[2][w][lane] = txy1n;
        sh[3][w][lane] = tyz1n;
        __syncthreads();
        const double Pt1U = sh[0][wu][lane];   // (c, d-1, k): used at step k+1
        const double tyy1U = sh[1][wu][lane];
        const double txy1D = sh[2][wd][lane];  // τxy (c, y-face d+1, k): used at step k+1
        const double tyz1D = sh[3][wd][lane];  // τyz (c, y-face d+1, z-face k)
        __syncthreads();

        // ---- stores of the new cell / edge values owned by this tile
        const bool kown = k >= kb && k < ke;
        if (own && kown) {
            if (cell) Pt1[o] = Pt1n;
            if (hxx) txx1[o - sx - sxy] = txx1n;
            if (hyy) tyy1[o - 1 - sxy] = tyy1n;
            if (hzz) tzz1[o - 1 - sx] = tzz1n;
            if (hxy) txy1[oe] = txy1n;
            if (hxz) txz1[oe] = txz1n;
            if (hyz) tyz1[oe] = tyz1n;
        }
        // ---- Vz at z-face k
        if (own && kown) {
            if (vz_in && k >= 1) {
                double R = dfd<MUL>(tzz1n - tzz1P, dz, _dz) + dfd<MUL>(txz1R - txz1n, dx, _dx) +
                           dfd<MUL>(tyz1D - tyz1n, dy, _dy) - dfd<MUL>(Pt1n - Pt1P, dz, _dz);
                if constexpr (BF) R = R - (rgP + rg) * 0.5;
                put_bc(Vz1, o, VzC + R * ((RhP + Rh) * 0.5), c, nx, 1, d, ny, sx);
            } else if (vz_in && k == 0) {  // Dirichlet face z = 0
                put_bc(Vz1, o, VzC, c, nx, 1, d, ny, sx);
            }
            if (vz_in && k == nz - 1) {    // Dirichlet face z = nz
                put_bc(Vz1, o + sxy, VzN, c, nx, 1, d, ny, sx);
            }
        }
        // ---- Vx, Vy at plane k-1
        const int km = k - 1;
        if (own && km >= kb && km < ke && km >= 1 && km <= nz - 2) {
            const int q = o - sxy;
            if (vx_in) {
                double R = dfd<MUL>(txx1P - txx1PL, dx, _dx) + dfd<MUL>(txy1PD - txy1P, dy, _dy) +
                           dfd<MUL>(txz1n - txz1P, dz, _dz) - dfd<MUL>(Pt1P - Pt1PL, dx, _dx);
                put_bc(Vx1, q, Vx0P + R * ((RhPL + RhP) * 0.5), d, ny, sx, km, nz, sxy);
            } else if (vx_dir) {
                put_bc(Vx1, q, Vx0P, d, ny, sx, km, nz, sxy);
            }
            if (vy_in) {
                double R = dfd<MUL>(tyy1P - tyy1PU, dy, _dy) + dfd<MUL>(txy1PR - txy1P, dx, _dx) +
                           dfd<MUL>(tyz1n - tyz1P, dz, _dz) - dfd<MUL>(Pt1P - Pt1PU, dy, _dy);
                put_bc(Vy1, q, Vy0P + R * ((RhUP + RhP) * 0.5), c, nx, 1, km, nz, sxy);
            } else if (vy_dir) {
                put_bc(Vy1, q, Vy0P, c, nx, 1, km, nz, sxy);
            }
        }

        // ---- carry to plane k+1
        Vx0P = Vx0c;
        Vy0P = Vy0c;
        VzC = VzN;
        VzUC = VzUN;
        GdLP = GdL;
        MuLP = MuL;
        GdUP = GdU;
        MuUP = MuU;
        GdP = Gd;
        MuP = Mu;
        RhUP = RhU;
        RhP = Rh;
        rgP = rg;
        Pt1P = Pt1n;
        txx1P = txx1n;
        tyy1P = tyy1n;
        tzz1P = tzz1n;
        txy1P = txy1n;
        txz1P = txz1n;
        tyz1P = tyz1n;
        Pt1PU = Pt1U;
        tyy1PU = tyy1U;
        txy1PD = txy1D;
    }
}

// Resident blocks of one k3_iter instance on the whole device.
template <bool VE, bool MUL, bool BF, int BY, bool DP>
int concurrency3() {
    static int cached = 0;
    if (!cached) {
        int dev = 0, cus = 0, per = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k3_iter<VE, MUL, BF, BY, DP>, 64 * BY, 0);
        cached = std::max(1, cus) * std::max(1, per);
    }
    return cached;
}

// Lz planes per block; Lz < 0: about one round of resident blocks, within
// [4, -Lz] planes (small grids get short columns, large ones long marches).
template <bool VE, bool MUL, bool BF, bool DP>
void go3(hipStream_t s, const Fields& in, const Fields& out, const StepScalars& c, int Lz, int by) {
    const int ntx = (in.nx + 1 + 61) / 62;
    auto launch = [&](auto byc) {
        constexpr int BY = decltype(byc)::value;
        const int nty = (in.ny + 1 + (BY - 2) - 1) / (BY - 2);
        int L = Lz;
        if (L < 0) {
            const int nch = std::max(1, concurrency3<VE, MUL, BF, BY, DP>() / (ntx * nty));
            L = std::max(4, std::min((in.nz + nch - 1) / nch, -Lz));
        }
        L = std::max(1, L);
        const int nch = (in.nz + L - 1) / L;
        k3_iter<VE, MUL, BF, BY, DP><<<ntx * nty * nch, dim3(64, BY), 0, s>>>(in, out, c, L, ntx, nty);
    };
    if constexpr (DP) {
        launch(std::integral_constant<int, 12>{});
    } else {
        if (by == 8) launch(std::integral_constant<int, 8>{});
        else if (by == 16) launch(std::integral_constant<int, 16>{});
        else launch(std::integral_constant<int, 12>{});
    }
}

}  // namespace

hipError_t launch3_iter(hipStream_t s, const Fields& in, const Fields& out, const StepScalars& c, bool ve, bool mul,
                        bool bf, int Lz, int by, bool dp) {
    if (Lz == 0) Lz = 1;
    // dp: derive Gdτ, dτ_Rho from Musτ (tile rows 12 only; the others read them)
    if (dp && (by != 8 && by != 16)) {
        if (ve) {
            if (mul) { if (bf) go3<true, true, true, true>(s, in, out, c, Lz, 12); else go3<true, true, false, true>(s, in, out, c, Lz, 12); }
            else { if (bf) go3<true, false, true, true>(s, in, out, c, Lz, 12); else go3<true, false, false, true>(s, in, out, c, Lz, 12); }
        } else {
            if (mul) { if (bf) go3<false, true, true, true>(s, in, out, c, Lz, 12); else go3<false, true, false, true>(s, in, out, c, Lz, 12); }
            else { if (bf) go3<false, false, true, true>(s, in, out, c, Lz, 12); else go3<false, false, false, true>(s, in, out, c, Lz, 12); }
        }
        return hipGetLastError();
    }
    if (ve) {
        if (mul) { if (bf) go3<true, true, true, false>(s, in, out, c, Lz, by); else go3<true, true, false, false>(s, in, out, c, Lz, by); }
        else { if (bf) go3<true, false, true, false>(s, in, out, c, Lz, by); else go3<true, false, false, false>(s, in, out, c, Lz, by); }
    } else {
        if (mul) { if (bf) go3<false, true, true, false>(s, in, out, c, Lz, by); else go3<false, true, false, false>(s, in, out, c, Lz, by); }
        else { if (bf) go3<false, false, true, false>(s, in, out, c, Lz, by); else go3<false, false, false, false>(s, in, out, c, Lz, by); }
    }
    return hipGetLastError();
}

}  // namespace pts
