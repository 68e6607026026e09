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
//
// Two kernels share this structure: k3_iter (predicated loads, any layout)
// and k3_iter2 (the default: grouped allocations, no load predicates, LDS
// exchange of the row-neighbour inputs, Musτ formed from μ); launch3_iter at
// the end of the file picks one.
#include <cmath>
#include <cstdint>
#include <type_traits>

#include "pts_internal.hpp"

namespace pts {

// Planes per block (Lz > 0: Lz).  Lz < 0: the tiles of a z-chunk start
// together and march to its end, so a launch takes ceil(blocks / resident)
// rounds of about L + 3 plane steps (the redundant plane below the chunk and
// the start-up of a block); take the L in [4, -Lz] with the fewest such steps
// (ties: the longer march).  At 511^3 (468 tiles, 256 resident) that is 86
// planes: 2808 blocks in 11 nearly full rounds.
int planes_per_block(int Lz, int nz, int64_t ntile, int64_t conc) {
    if (Lz > 0) return Lz;
    const int lmax = std::max(1, std::min(-Lz, nz));
    int L = lmax;
    int64_t best = -1;
    for (int l = lmax; l >= std::min(4, lmax); --l) {
        const int64_t nch = (nz + l - 1) / l;
        const int64_t cost = (ntile * nch + conc - 1) / conc * (l + 3);
        if (best < 0 || cost < best) {
            best = cost;
            L = l;
        }
    }
    return L;
}


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

// Lz planes per block (Lz < 0: planes_per_block's choice within [4, -Lz]).
template <bool VE, bool MUL, bool BF, bool DP>
void go3(hipStream_t s, const Fields& in, const Fields& out, const StepScalars& c, int Lz, int by) {
    const int ntx = (in.nx + 1 + 61) / 62;
    auto launch = [&](auto byc) {
        constexpr int BY = decltype(byc)::value;
        const int nty = (in.ny + 1 + (BY - 2) - 1) / (BY - 2);
        const int L = planes_per_block(Lz, in.nz, (int64_t)ntx * nty, concurrency3<VE, MUL, BF, BY, DP>());
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


// ---------------------------------------------------------------------------
// k3_iter2 -- the same iteration and traversal, with less work per cell:
//   * one 32-bit byte offset per thread and plane addresses every load and
//     store: each field's pointer carries the field's fixed index shift (txx
//     at (c, d-1, k-1), edges at (c-1, d-1, k-1), ...) and the input, output
//     and old-stress buffers are slots of one allocation, so the loads use the
//     scalar-base form with 12 base pointers;
//   * loads carry no lane or z-range predicates: FP64 3D allocations have
//     guard zones (a plane and 32 rows) around the strided box, and lanes or
//     rows outside [0, nx] x [0, ny] load at the nearest position inside.
//     Such values only feed results that are not stored;
//   * the row-above inputs (Gdτ, dτ_Rho, μ, Vx, Vz) and the row-below Vy come
//     from LDS, written by the neighbouring row one plane ahead, instead of
//     being loaded again and (Gdτ, dτ_Rho) recomputed from Musτ;
//   * each row role runs its own loop: the top halo row only forms Pt and τyy
//     for its neighbour, the bottom halo row τxy and τyz;
//   * the tiles are dealt so that each XCD sweeps its own eighth of the tiles
//     chunk by chunk, and the march length minimises the rounds of blocks;
//   * MX: the own rows form Musτ from μ (the cross maximum of
//     compute_maxloc!), one array fewer per iteration;
//   * PF: the loads of plane k+1 are issued before the arithmetic of plane k;
//   * G2: with G·dt a power of two, τ·Gr uses Gdτ·(1/(G·dt)), which rounds
//     exactly as Gdτ/(G·dt).
//   * SL: rows 2 .. BY-3 of the tile (no other tile's halo rows) load the
//     lines that no neighbouring tile reads with the streaming policy, so
//     that the L2 keeps the lines the tiles share (3D VE 511^3: 5.95 ->
//     5.80 ms per iteration; the L2-to-memory read requests are unchanged,
//     the fewer replacements of shared lines shorten the waits).
// Bit-identical to k3_iter.

// Fields of k3_iter2.  Each state field's input and output buffers (and,
// for the visco-elastic stresses, τ_o) are slots of one allocation: base
// points at slot 0 and the slots are byte offsets from it.
struct Fields3g {
    const double* base[PTS_TYZ + 1];  // PT, VX, VY, VZ, TXX .. TYZ
    const double* g;                  // Musτ (dp) or Gdτ
    const double* mu;
    const double* rh;                 // dτ_Rho (without dp)
    const double* rg;                 // ρg (body force)
    uint64_t off_in, off_out, off_old;
    int sx, sxy, nx, ny, nz;
};

struct Plane3 {  // inputs of one plane at (c, d); g/h: Gdτ, dτ_Rho after the exchange
    double g, h, mu, pt, vx, vy, vzn, rg;
    double mun;  // MX: μ of the next plane
    double tx, txo, ty, tyo, tz, tzo, txy, txyo, txz, txzo, tyz, tyzo;
};
struct Nbr3 {  // the same plane from the neighbouring rows
    double gU, rhU, muU, vxU, vzUN, vyN;
    double muUn, muDn;  // MX: μ of the next plane in the rows above and below
};
struct Carry3 {  // plane k-1
    double vxP, vyP, gP, muP, gUP, muUP, rhP, rhUP, rgP;
    double pt1P, txx1P, tyy1P, tzz1P, txy1P, txz1P, tyz1P;
    double vzC, vzUC;  // Vz at z-face k (own row, row above)
};

template <bool NT = false>
__device__ __forceinline__ double ldb(const double* p, uint32_t boff) {
    // an empty asm makes the offset a value of the load's own block, so that
    // the address is formed there as scalar base + 32-bit offset (no 64-bit
    // per-lane address) also for loads under branches
    asm volatile("" : "+v"(boff));
    const double* q = reinterpret_cast<const double*>(reinterpret_cast<const char*>(p) + boff);
    if constexpr (NT) return __builtin_nontemporal_load(q);
    else return *q;
}
__device__ __forceinline__ void stb(double* p, uint32_t boff, double v) {
    *reinterpret_cast<double*>(reinterpret_cast<char*>(p) + boff) = v;
}
// put_bc with a byte offset (same copies)
__device__ __forceinline__ void put_bcb(double* V, uint32_t q, double v, int a, int na, int sa, int b, int nb, int sb) {
    stb(V, q, v);
    const int da = (a == 1) ? -1 : ((a == na - 2) ? 1 : 0);
    const int db = (b == 1) ? -1 : ((b == nb - 2) ? 1 : 0);
    const uint32_t ea = (uint32_t)(da * sa * 8), eb = (uint32_t)(db * sb * 8);
    if (da) stb(V, q + ea, v);
    if (db) stb(V, q + eb, v);
    if (da && db) stb(V, q + ea + eb, v);
    if (na == 3 || nb == 3) {  // grid-uniform: the far-side copies of extent 3
        const bool a3 = na == 3 && a == 1, b3 = nb == 3 && b == 1;
        const uint32_t fa = (uint32_t)sa * 8u, fb = (uint32_t)sb * 8u;
        if (a3) {
            stb(V, q + fa, v);
            if (db) stb(V, q + fa + eb, v);
        }
        if (b3) {
            stb(V, q + fb, v);
            if (da) stb(V, q + ea + fb, v);
            if (a3) stb(V, q + fa + fb, v);
        }
    }
}

template <bool VE, bool G2>
__device__ __forceinline__ double tau_upd2(double t, double to, double g, double m, const StepScalars& cs, double e) {
    if constexpr (VE) {
        const double Gr = G2 ? g * cs.iGdt : g / cs.Gdt;
        return (t + to * Gr + 2.0 * g * e) / (1.0 + g / m + Gr);
    } else {
        return (t + 2.0 * g * e) / (1.0 + g / m);
    }
}

template <bool VE, bool MUL, bool BF, int BY, bool DP, bool PF, bool G2, bool MX, bool W64, bool SL>
__global__ __launch_bounds__(64 * BY) void k3_iter2(const Fields3g f, const StepScalars cs, int Lz,
                                                    int ntx, int nty, int nchunk) {
    // new Pt1, τyy1 (read by row +1) and τxy1, τyz1 (read by row -1), by plane
    // parity: the V update of plane k-1 re-reads plane k-1's from the other half
    __shared__ double shA[2][4][BY][64];
    // next-plane inputs: Gdτ, dτ_Rho, μ (MX: μ of the plane after), Vx, Vz (row +1), Vy (row -1); Vz, μ (start)
    __shared__ double shB[8][BY][64];
    static_assert(!(MX && !DP), "MX: derived parameters");

    // Block b runs on XCD b % 8.  Each XCD owns a contiguous eighth of the
    // tiles (neighbouring tiles share halo rows in its L2) and takes the
    // z-chunks in order, so all XCDs sweep the same chunk at the same time on
    // different rows of the planes (rather than eight chunks whose planes lie
    // a multiple of the plane size apart).  Blocks beyond an XCD's share exit.
    const int ntile = ntx * nty;
    int tile, chunk;
    if (nchunk > 0) {
        const int xcd = blockIdx.x & 7, seq = blockIdx.x >> 3;
        const int lo = (int)((int64_t)ntile * xcd / 8), cnt = (int)((int64_t)ntile * (xcd + 1) / 8) - lo;
        if (seq >= cnt * nchunk) return;
        tile = lo + seq % cnt;
        chunk = seq / cnt;
    } else {  // nchunk = 0: contiguous runs of the (chunk, tile) order per XCD
        const int b = xcd_tile(blockIdx.x, gridDim.x);
        tile = b % ntile;
        chunk = b / ntile;
    }
    const int tx = tile % ntx, ty = tile / ntx;
    const int lane = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int c = tx * 62 - 1 + lane;
    const int d = ty * (BY - 2) - 1 + w;
    const int nx = f.nx, ny = f.ny, nz = f.nz, sx = f.sx, sxy = f.sxy;
    const int kb = chunk * Lz;
    const int ke = min(kb + Lz, nz);
    // Byte offset of the loads of (c, d) within a plane.  Lanes and rows
    // outside [0, nx] x [0, ny] (halo and edge tiles) load at the nearest
    // position inside: their values feed no stored result, and the clamp keeps
    // them on lines that the neighbouring lanes and rows fetch anyway (instead
    // of the next row's or plane's).  The bias keeps every shifted field offset
    // non-negative.
    const int bias = sx + 1;
    const uint32_t tb = (uint32_t)(min(max(c, 0), nx) + sx * min(max(d, 0), ny) + bias) * 8u;


    // element shifts of the fields relative to (c, d, k)
    const long long sTXX = -(long long)sx - sxy, sTYY = -1 - (long long)sxy, sTZZ = -1 - (long long)sx;
    const long long sE = -1 - (long long)sx - sxy;
    // one pointer per field (slot 0 of its allocation, moved by the field's
    // shift and the bias); the input, output and τ_o slots are byte offsets
    const double* __restrict__ Ptb = f.base[PTS_PT] - bias;
    const double* __restrict__ txxb = f.base[PTS_TXX] + sTXX - bias;
    const double* __restrict__ tyyb = f.base[PTS_TYY] + sTYY - bias;
    const double* __restrict__ tzzb = f.base[PTS_TZZ] + sTZZ - bias;
    const double* __restrict__ txyb = f.base[PTS_TXY] + sE - bias;
    const double* __restrict__ txzb = f.base[PTS_TXZ] + sE - bias;
    const double* __restrict__ tyzb = f.base[PTS_TYZ] + sE - bias;
    const double* __restrict__ Vxb = f.base[PTS_VX] - bias;
    const double* __restrict__ Vyb = f.base[PTS_VY] - bias;
    const double* __restrict__ Vzb = f.base[PTS_VZ] - bias;
    const int k0 = max(kb - 1, 0), k1 = min(ke, nz - 1);
    // W64: the slots of a field lie more than 4 GiB apart, so each slot gets
    // its own pointer, moved to the block's first plane k0, and the per-lane
    // offset counts planes from k0 (a march of at most a few hundred planes
    // stays within 32 bits).  Otherwise one pointer per field and the slots
    // as 32-bit offsets (the same pointer value for every slot).
    const int64_t kshift = W64 ? (int64_t)k0 * sxy : 0;
    auto slot = [&](const double* b, uint64_t off) __attribute__((always_inline)) {
        return W64 ? b + (int64_t)(off >> 3) + kshift : b;
    };
    const double* __restrict__ Gp = f.g - bias + kshift;
    const double* __restrict__ Mup = f.mu - bias + kshift;
    const double* __restrict__ Rhp = f.rh - bias + kshift;
    const double* __restrict__ rgp = f.rg - bias + kshift;
    const uint32_t oin = W64 ? 0u : (uint32_t)f.off_in, oout = W64 ? 0u : (uint32_t)f.off_out;
    const uint32_t oold = W64 ? 0u : (uint32_t)f.off_old;
    // inputs (i), τ_o (l) and outputs (o) of the state fields
    const double* __restrict__ Pt_i = slot(Ptb, f.off_in);
    const double* __restrict__ Vx_i = slot(Vxb, f.off_in);
    const double* __restrict__ Vy_i = slot(Vyb, f.off_in);
    const double* __restrict__ Vz_i = slot(Vzb, f.off_in);
    const double* __restrict__ txx_i = slot(txxb, f.off_in);
    const double* __restrict__ tyy_i = slot(tyyb, f.off_in);
    const double* __restrict__ tzz_i = slot(tzzb, f.off_in);
    const double* __restrict__ txy_i = slot(txyb, f.off_in);
    const double* __restrict__ txz_i = slot(txzb, f.off_in);
    const double* __restrict__ tyz_i = slot(tyzb, f.off_in);
    const double* __restrict__ txx_l = slot(txxb, f.off_old);
    const double* __restrict__ tyy_l = slot(tyyb, f.off_old);
    const double* __restrict__ tzz_l = slot(tzzb, f.off_old);
    const double* __restrict__ txy_l = slot(txyb, f.off_old);
    const double* __restrict__ txz_l = slot(txzb, f.off_old);
    const double* __restrict__ tyz_l = slot(tyzb, f.off_old);
    double* __restrict__ Ptm = const_cast<double*>(slot(Ptb, f.off_out));
    double* __restrict__ txxm = const_cast<double*>(slot(txxb, f.off_out));
    double* __restrict__ tyym = const_cast<double*>(slot(tyyb, f.off_out));
    double* __restrict__ tzzm = const_cast<double*>(slot(tzzb, f.off_out));
    double* __restrict__ txym = const_cast<double*>(slot(txyb, f.off_out));
    double* __restrict__ txzm = const_cast<double*>(slot(txzb, f.off_out));
    double* __restrict__ tyzm = const_cast<double*>(slot(tyzb, f.off_out));
    double* __restrict__ Vxm = const_cast<double*>(slot(Vxb, f.off_out));
    double* __restrict__ Vym = const_cast<double*>(slot(Vyb, f.off_out));
    double* __restrict__ Vzm = const_cast<double*>(slot(Vzb, f.off_out));

    const double dx = cs.dx, dy = cs.dy, dz = cs.dz, _dx = cs._dx, _dy = cs._dy, _dz = cs._dz;
    const double r = cs.r;
    // byte offset of (c, d, k): one 32-bit value per plane for every field
    // (W64: planes counted from k0)
    auto tko = [&](int k) __attribute__((always_inline)) {
        return tb + (uint32_t)(W64 ? k - k0 : k) * (uint32_t)sxy * 8u;
    };

    const uint32_t sxy8 = (uint32_t)sxy * 8u;
    // lanes whose column shares a 128-byte line with the tile's first or last
    // column (lines the neighbouring tiles fetch too)
    const int c0 = tx * 62 - 1;
    const bool edge_line = (c >> 4) == (c0 >> 4) || (c >> 4) == ((c0 + 63) >> 4);
    (void)edge_line;

    // Each row role runs its own loop (R = 0: top halo row, 1: own rows,
    // 2: bottom halo row; w is wave-uniform), loading only what it uses.
    // Role 3 (SL) is an own row that no neighbouring tile loads as a halo row.  The
    // loads need no z-range conditions either: the guard zones hold a plane.
    // MX (own rows): Musτ of plane k is formed from μ (the cross maximum of
    // compute_maxloc!, Stokes3D_ve_perf.jl:29-32) except on the boundary
    // planes and rows and at the first plane of the march, where it is loaded.
    auto mt_loaded = [&](auto role, int k, bool start) __attribute__((always_inline)) {
        constexpr int R = decltype(role)::value == 3 ? 1 : decltype(role)::value;
        if constexpr (!MX || R != 1) return true;
        else return start || k == 0 || k == nz - 1 || d == 0 || d == ny - 1;
    };
    auto load = [&](auto role, int k, Plane3& I, bool start) __attribute__((always_inline)) {
        constexpr int R = decltype(role)::value == 3 ? 1 : decltype(role)::value;
        // role 3 (SL): lines that no neighbouring tile reads are loaded with
        // the streaming (non-temporal) policy, so that they leave the L2
        // first and the lines the neighbours share stay
        auto ldA = [&](const double* p, uint32_t o) __attribute__((always_inline)) {
            if constexpr (decltype(role)::value == 3) {
                return edge_line ? ldb<false>(p, o) : ldb<true>(p, o);
            } else {
                return ldb<false>(p, o);
            }
        };
        const uint32_t t = tko(k), ti = t + oin, to = t + oold;
        if (mt_loaded(role, k, start)) {
            I.g = ldA(Gp, t);
        } else if (lane == 0 || lane == 63) {
            // MX: μ of the x-neighbour outside the wave (c - 1, c + 1, clamped
            // like tb), held where the loaded Musτ would be
            int cl = c;
            asm volatile("" : "+v"(cl));
            const int cc = min(max(cl, 0), nx), cf = lane == 0 ? max(cl - 1, 0) : min(cl + 1, nx);
            I.g = ldb(Mup, t + (uint32_t)((cf - cc) * 8));  // never streaming
        }
        if constexpr (!DP && R != 2) I.h = ldA(Rhp, t);
        if constexpr (MX) {
            I.mun = ldA(Mup, t + sxy8);  // μ(k) comes from the previous plane's load
            if (start) I.mu = ldA(Mup, t);
        } else {
            I.mu = ldA(Mup, t);
        }
        I.vx = ldA(Vx_i, ti);
        I.vy = ldA(Vy_i, ti);
        I.vzn = ldA(Vz_i, ti + sxy8);
        if constexpr (R != 2) I.pt = ldA(Pt_i, ti);
        if constexpr (BF && R == 1) I.rg = ldA(rgp, t);
        if constexpr (R == 1) {
            I.tx = ldA(txx_i, ti);
            if constexpr (VE) I.txo = ldA(txx_l, to);
            I.tz = ldA(tzz_i, ti);
            if constexpr (VE) I.tzo = ldA(tzz_l, to);
            I.txz = ldA(txz_i, ti);
            if constexpr (VE) I.txzo = ldA(txz_l, to);
        }
        if constexpr (R != 2) {
            I.ty = ldA(tyy_i, ti);
            if constexpr (VE) I.tyo = ldA(tyy_l, to);
        }
        if constexpr (R != 0) {
            I.txy = ldA(txy_i, ti);
            if constexpr (VE) I.txyo = ldA(txy_l, to);
            I.tyz = ldA(tyz_i, ti);
            if constexpr (VE) I.tyzo = ldA(tyz_l, to);
        }
    };
    // Gdτ, dτ_Rho of plane k (DP: from Musτ) and the exchange with the
    // neighbouring rows; `start`: also pass Vz (and, MX, μ) of the first plane
    auto exchange = [&](auto role, int k, Plane3& I, Nbr3& U, const Carry3& P, bool start, double vzc,
                        double& vzUC) __attribute__((always_inline)) {
        constexpr int R = decltype(role)::value == 3 ? 1 : decltype(role)::value;
        if constexpr (DP) {
            double M = I.g;
            if constexpr (MX && R == 1) {
                if (!mt_loaded(role, k, start)) {
                    // cross maximum in compute_maxloc!'s order: x, centre, y, z;
                    // the row neighbours came one plane ahead, μ(k-1) is carried.
                    // Lanes 0 and 63 take the x-neighbour outside the wave from
                    // their own load (their Gdτ feeds the edges of lanes 1 and
                    // 62); a maximum is exact, so the order of the terms does
                    // not change the result.
                    double xl = from_left(I.mu), xr = from_right(I.mu);
                    if (lane == 0) xl = I.g;
                    if (lane == 63) xr = I.g;
                    double m = fmax(xl, xr);
                    m = fmax(m, I.mu);
                    m = fmax(m, U.muUn);
                    m = fmax(m, U.muDn);
                    m = fmax(m, P.muP);
                    m = fmax(m, I.mun);
                    const double m1 = m;
                    // free-slip copies (bc_x!): the boundary columns take their neighbour's
                    const double mr = from_right(m1), ml = from_left(m1);
                    M = c == 0 ? mr : (c == nx - 1 ? ml : m1);
                }
            }
            double rh, gd;
            iter_params_of(M, cs, rh, gd);
            I.g = gd;
            I.h = rh;
        }
        if constexpr (R != 2) {  // read by the row below
            shB[0][w][lane] = I.g;
            shB[1][w][lane] = I.h;
            if constexpr (!MX) shB[2][w][lane] = I.mu;
            shB[3][w][lane] = I.vx;
            shB[4][w][lane] = I.vzn;
            if (start) shB[6][w][lane] = vzc;
        }
        if constexpr (MX) {
            shB[2][w][lane] = I.mun;  // read by the rows above and below
            if (start) shB[7][w][lane] = I.mu;
        }
        if constexpr (R != 0) shB[5][w][lane] = I.vy;  // read by the row above
        __syncthreads();
        if constexpr (R != 0) {
            U.gU = shB[0][w - 1][lane];
            U.rhU = shB[1][w - 1][lane];
            if constexpr (MX) {
                U.muU = start ? shB[7][w - 1][lane] : U.muUn;
                U.muUn = shB[2][w - 1][lane];
            } else {
                U.muU = shB[2][w - 1][lane];
            }
            U.vxU = shB[3][w - 1][lane];
            U.vzUN = shB[4][w - 1][lane];
            if (start) vzUC = shB[6][w - 1][lane];
        }
        if constexpr (R != 2) U.vyN = shB[5][w + 1][lane];
        if constexpr (MX && R == 1) U.muDn = shB[2][w + 1][lane];
    };

    auto step = [&](auto role, int k, Plane3& I, Plane3& N, Nbr3& U, Carry3& P) __attribute__((always_inline)) {
        constexpr int R = decltype(role)::value == 3 ? 1 : decltype(role)::value;
        if constexpr (PF) {
            if (k < k1) load(role, k + 1, N, false);
        }
        const double Gd = I.g, Rh = I.h, Mu = I.mu;
        double Pt1n = 0.0, txx1n = 0.0, tyy1n = 0.0, tzz1n = 0.0, txy1n = 0.0, txz1n = 0.0, tyz1n = 0.0;
        double GdL = 0.0, MuL = 0.0;
        if constexpr (R != 2) {  // cell (c, d, k)
            const double VxR = from_right(I.vx);
            const double ex = dfd<MUL>(VxR - I.vx, dx, _dx);
            const double ey = dfd<MUL>(U.vyN - I.vy, dy, _dy);
            const double ez = dfd<MUL>(I.vzn - P.vzC, dz, _dz);
            Pt1n = I.pt - r * Gd * (ex + ey + ez);
            tyy1n = tau_upd2<VE, G2>(I.ty, I.tyo, Gd, Mu, cs, ey);
            if constexpr (R == 1) {
                txx1n = tau_upd2<VE, G2>(I.tx, I.txo, Gd, Mu, cs, ex);
                tzz1n = tau_upd2<VE, G2>(I.tz, I.tzo, Gd, Mu, cs, ez);
            }
        }
        if constexpr (R != 0) {  // edges
            GdL = from_left(Gd);
            MuL = from_left(Mu);
            const double GdUL = from_left(U.gU), MuUL = from_left(U.muU);
            const double VyL = from_left(I.vy);
            {   // τxy at (x-face c, y-face d, cell k)
                const double ag = (GdUL + U.gU + GdL + Gd) * 0.25;
                const double am = (MuUL + U.muU + MuL + Mu) * 0.25;
                const double e = 0.5 * (dfd<MUL>(I.vx - U.vxU, dy, _dy) + dfd<MUL>(I.vy - VyL, dx, _dx));
                txy1n = tau_upd2<VE, G2>(I.txy, I.txyo, ag, am, cs, e);
            }
            if constexpr (R == 1) {  // τxz at (x-face c, cell d, z-face k)
                const double VzL = from_left(P.vzC);
                const double ag = (from_left(P.gP) + P.gP + GdL + Gd) * 0.25;
                const double am = (from_left(P.muP) + P.muP + MuL + Mu) * 0.25;
                const double e = 0.5 * (dfd<MUL>(I.vx - P.vxP, dz, _dz) + dfd<MUL>(P.vzC - VzL, dx, _dx));
                txz1n = tau_upd2<VE, G2>(I.txz, I.txzo, ag, am, cs, e);
            }
            {   // τyz at (cell c, y-face d, z-face k)
                const double ag = (P.gUP + P.gP + U.gU + Gd) * 0.25;
                const double am = (P.muUP + P.muP + U.muU + Mu) * 0.25;
                const double e = 0.5 * (dfd<MUL>(I.vy - P.vyP, dz, _dz) + dfd<MUL>(P.vzC - P.vzUC, dy, _dy));
                tyz1n = tau_upd2<VE, G2>(I.tyz, I.tyzo, ag, am, cs, e);
            }
        }

        // y-neighbour exchange of the new values
        const int cb = k & 1, pb = cb ^ 1;
        if constexpr (R != 2) {
            shA[cb][0][w][lane] = Pt1n;
            shA[cb][1][w][lane] = tyy1n;
        }
        if constexpr (R != 0) {
            shA[cb][2][w][lane] = txy1n;
            shA[cb][3][w][lane] = tyz1n;
        }
        __syncthreads();
        if constexpr (R == 1) {
            const double tyz1D = shA[cb][3][w + 1][lane];  // τyz (c, y-face d+1, z-face k)

            const uint32_t t = tko(k) + oout;
            const double txz1R = from_right(txz1n);
            const bool kin2 = k >= 1 && k <= nz - 2, kin1 = k >= 1 && k <= nz - 1;
            // Store predicates: a lane part from the column and a wave-uniform
            // part from the row.  The column goes through an empty asm so that
            // the masks are formed here each step (a few compares) instead of
            // being held in scalar registers across the loop.
            int cl = c;
            asm volatile("" : "+v"(cl));
            const bool ol = lane >= 1 && lane <= 62;
            const bool l_cell = ol && cl < nx;                 // cells, τxx
            const bool l_in2 = ol && cl >= 1 && cl <= nx - 2;  // τyy, τzz, τyz, Vy, Vz
            const bool l_in1 = ol && cl >= 1 && cl <= nx - 1;  // τxy, τxz, Vx
            const bool l_dir = ol && (cl == 0 || cl == nx);    // Dirichlet Vx faces
            const bool r_in = d >= 1 && d <= ny - 2, r_in1 = d >= 1 && d <= ny - 1;
            const bool s_cell = l_cell && d < ny, s_txx = l_cell && r_in, s_tyy = l_in2 && d <= ny - 1;
            const bool s_tzz = l_in2 && r_in, s_txy = l_in1 && r_in1, s_txz = l_in1 && r_in, s_tyz = l_in2 && r_in1;
            const bool vx_in = l_in1 && r_in, vx_dir = l_dir && r_in;
            const bool vy_in = l_in2 && r_in1, vy_dir = l_in2 && (d == 0 || d == ny);
            const bool vz_in = l_in2 && r_in;
            if (k >= kb && k < ke) {
                if (s_cell) stb(Ptm, t, Pt1n);
                if (kin2 && s_txx) stb(txxm, t, txx1n);
                if (kin2 && s_tyy) stb(tyym, t, tyy1n);
                if (s_tzz) stb(tzzm, t, tzz1n);
                if (kin2 && s_txy) stb(txym, t, txy1n);
                if (kin1 && s_txz) stb(txzm, t, txz1n);
                if (kin1 && s_tyz) stb(tyzm, t, tyz1n);
                // Vz at z-face k
                if (vz_in) {
                    if (k >= 1) {
                        double Rz = dfd<MUL>(tzz1n - P.tzz1P, dz, _dz) + dfd<MUL>(txz1R - txz1n, dx, _dx) +
                                    dfd<MUL>(tyz1D - tyz1n, dy, _dy) - dfd<MUL>(Pt1n - P.pt1P, dz, _dz);
                        if constexpr (BF) Rz = Rz - (P.rgP + I.rg) * 0.5;
                        put_bcb(Vzm, t, P.vzC + Rz * ((P.rhP + Rh) * 0.5), cl, nx, 1, d, ny, sx);
                    } else {  // Dirichlet face z = 0
                        put_bcb(Vzm, t, P.vzC, cl, nx, 1, d, ny, sx);
                    }
                    if (k == nz - 1) put_bcb(Vzm, t + sxy8, I.vzn, cl, nx, 1, d, ny, sx);  // face z = nz
                }
            }
            // Vx, Vy at plane k-1
            const int km = k - 1;
            if (km >= kb && km < ke && km >= 1 && km <= nz - 2) {
                const double Pt1PL = from_left(P.pt1P), txx1PL = from_left(P.txx1P), RhPL = from_left(P.rhP);
                const double txy1PR = from_right(P.txy1P);
                const double Pt1PU = shA[pb][0][w - 1][lane];   // (c, d-1, k-1)
                const double tyy1PU = shA[pb][1][w - 1][lane];
                const double txy1PD = shA[pb][2][w + 1][lane];  // τxy (c, y-face d+1, k-1)
                const uint32_t tq = t - sxy8;
                if (vx_in) {
                    const double Rx = dfd<MUL>(P.txx1P - txx1PL, dx, _dx) + dfd<MUL>(txy1PD - P.txy1P, dy, _dy) +
                                      dfd<MUL>(txz1n - P.txz1P, dz, _dz) - dfd<MUL>(P.pt1P - Pt1PL, dx, _dx);
                    put_bcb(Vxm, tq, P.vxP + Rx * ((RhPL + P.rhP) * 0.5), d, ny, sx, km, nz, sxy);
                } else if (vx_dir) {
                    put_bcb(Vxm, tq, P.vxP, d, ny, sx, km, nz, sxy);
                }
                if (vy_in) {
                    const double Ry = dfd<MUL>(P.tyy1P - tyy1PU, dy, _dy) + dfd<MUL>(txy1PR - P.txy1P, dx, _dx) +
                                      dfd<MUL>(tyz1n - P.tyz1P, dz, _dz) - dfd<MUL>(P.pt1P - Pt1PU, dy, _dy);
                    put_bcb(Vym, tq, P.vyP + Ry * ((P.rhUP + P.rhP) * 0.5), cl, nx, 1, km, nz, sxy);
                } else if (vy_dir) {
                    put_bcb(Vym, tq, P.vyP, cl, nx, 1, km, nz, sxy);
                }
            }
        }

        // carry to plane k+1 (what the role uses)
        P.vzC = I.vzn;
        if constexpr (R != 0) {
            P.vyP = I.vy;
            P.vzUC = U.vzUN;
            P.gUP = U.gU;
            P.muUP = U.muU;
            P.gP = Gd;
            P.muP = Mu;
        }
        if constexpr (R == 1) {
            P.vxP = I.vx;
            P.rhUP = U.rhU;
            P.rhP = Rh;
            if constexpr (BF) P.rgP = I.rg;
            P.pt1P = Pt1n;
            P.txx1P = txx1n;
            P.tyy1P = tyy1n;
            P.tzz1P = tzz1n;
            P.txy1P = txy1n;
            P.txz1P = txz1n;
            P.tyz1P = tyz1n;
        }

        if (k < k1) {
            if constexpr (!PF) load(role, k + 1, N, false);
            if constexpr (MX) N.mu = I.mun;
            double unused = 0.0;
            exchange(role, k + 1, N, U, P, false, 0.0, unused);
        }
    };

    auto run = [&](auto role) __attribute__((always_inline)) {
        Plane3 A{}, B{};
        Nbr3 U{};
        Carry3 P{};
        load(role, k0, A, true);
        const double vz0 = ldb(Vz_i, tko(k0) + oin);  // Vz at z-face k0
        P.vzC = vz0;
        exchange(role, k0, A, U, P, true, vz0, P.vzUC);
        for (int k = k0;;) {
            step(role, k, A, B, U, P);
            if (k == k1) break;
            ++k;
            step(role, k, B, A, U, P);
            if (k == k1) break;
            ++k;
        }
    };
    if (w == 0) run(std::integral_constant<int, 0>{});
    else if (w == BY - 1) run(std::integral_constant<int, 2>{});
    else {
        if constexpr (SL) {
            // rows 2 .. BY-3 are no neighbouring tile's halo rows
            if (w >= 2 && w <= BY - 3) {
                run(std::integral_constant<int, 3>{});
                return;
            }
        }
        run(std::integral_constant<int, 1>{});
    }
}

template <bool VE, bool MUL, bool BF, int BY, bool DP, bool PF, bool G2, bool MX, bool W64, bool SL>
int concurrency3b() {
    static int cached = 0;
    if (!cached) {
        int dev = 0, cus = 0, per = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k3_iter2<VE, MUL, BF, BY, DP, PF, G2, MX, W64, SL>, 64 * BY, 0);
        cached = std::max(1, cus) * std::max(1, per);
    }
    return cached;
}

template <bool VE, bool MUL, bool BF, bool DP, bool PF, bool G2, bool MX = false, bool W64 = false, bool SL = false>
void go3b(hipStream_t s, const Fields3g& in, const StepScalars& c, int Lz, bool map_flat) {
    constexpr int BY = 12;
    if constexpr (W64) {  // the per-lane offset spans the march: (L + 4) planes within 2^31 bytes
        const int64_t cap = std::max<int64_t>(1, ((int64_t)1 << 31) / ((int64_t)in.sxy * 8) - 4);
        Lz = Lz > 0 ? (int)std::min<int64_t>(Lz, cap) : -(int)std::min<int64_t>(-(int64_t)Lz, cap);
    }
    const int ntx = (in.nx + 1 + 61) / 62;
    const int nty = (in.ny + 1 + (BY - 2) - 1) / (BY - 2);
    const int64_t conc = concurrency3b<VE, MUL, BF, BY, DP, PF, G2, MX, W64, SL>();
    if (map_flat) {
        const int L = planes_per_block(Lz, in.nz, (int64_t)ntx * nty, conc);
        const int nch = (in.nz + L - 1) / L;
        k3_iter2<VE, MUL, BF, BY, DP, PF, G2, MX, W64, SL><<<(unsigned)(ntx * nty * nch), dim3(64, BY), 0, s>>>(in, c, L, ntx, nty, 0);
        return;
    }
    // per XCD: ceil(tiles / 8) tiles per chunk on an eighth of the resident blocks
    const int64_t per = ((int64_t)ntx * nty + 7) / 8;
    const int L = planes_per_block(Lz, in.nz, per, std::max<int64_t>(1, conc / 8));
    const int nch = (in.nz + L - 1) / L;
    k3_iter2<VE, MUL, BF, BY, DP, PF, G2, MX, W64, SL><<<(unsigned)(8 * per * nch), dim3(64, BY), 0, s>>>(in, c, L, ntx, nty, nch);
}

template <bool VE, bool MUL, bool BF, bool DP>
void go3b_pf(hipStream_t s, const Fields3g& in, const StepScalars& c, int Lz, bool pf, bool g2, bool mx,
             bool map_flat, bool w64, bool sl) {
    if constexpr (DP) {
        // slots more than 4 GiB apart: per-slot pointers (without prefetch)
        if (w64) {
            if (mx) {
                // sl: streaming loads for the lines no neighbouring tile reads
                if (sl) {
                    if (g2) go3b<VE, MUL, BF, DP, false, true, true, true, true>(s, in, c, Lz, map_flat);
                    else go3b<VE, MUL, BF, DP, false, false, true, true, true>(s, in, c, Lz, map_flat);
                } else {
                    if (g2) go3b<VE, MUL, BF, DP, false, true, true, true>(s, in, c, Lz, map_flat);
                    else go3b<VE, MUL, BF, DP, false, false, true, true>(s, in, c, Lz, map_flat);
                }
            } else {
                if (g2) go3b<VE, MUL, BF, DP, false, true, false, true>(s, in, c, Lz, map_flat);
                else go3b<VE, MUL, BF, DP, false, false, false, true>(s, in, c, Lz, map_flat);
            }
            return;
        }
        if (mx && pf) {
            if (g2) go3b<VE, MUL, BF, DP, true, true, true>(s, in, c, Lz, map_flat);
            else go3b<VE, MUL, BF, DP, true, false, true>(s, in, c, Lz, map_flat);
            return;
        }
        if (mx && !pf) {
            if (g2) go3b<VE, MUL, BF, DP, false, true, true>(s, in, c, Lz, map_flat);
            else go3b<VE, MUL, BF, DP, false, false, true>(s, in, c, Lz, map_flat);
            return;
        }
    }
    if (pf) {
        if (g2) go3b<VE, MUL, BF, DP, true, true>(s, in, c, Lz, map_flat);
        else go3b<VE, MUL, BF, DP, true, false>(s, in, c, Lz, map_flat);
    } else {
        if (g2) go3b<VE, MUL, BF, DP, false, true>(s, in, c, Lz, map_flat);
        else go3b<VE, MUL, BF, DP, false, false>(s, in, c, Lz, map_flat);
    }
}

// The grouped view of a ping-pong pair, if its buffers are laid out as slots
// of one allocation `slot` bytes apart (in and out in slots 0/1 in either
// order, τ_o in slot 2).  0: not grouped; 1: every offset fits in 32 bits;
// 2: per-slot pointers are needed (or asked for by w64), which only asks
// that a few planes fit in 31 bits.
int group_fields(const Fields& in, const Fields& out, bool ve, bool dp, bool bf, int64_t slot, bool w64,
                 Fields3g& g) {
    if (slot <= 0) return 0;
    g = Fields3g{};
    int par = -1;
    for (int f = PTS_PT; f <= PTS_TYZ; ++f) {
        const char* a = (const char*)in.p[f];
        const char* b = (const char*)out.p[f];
        if (!a || !b) return 0;
        int p;
        if (b - a == slot) p = 0;
        else if (a - b == slot) p = 1;
        else return 0;
        if (par >= 0 && p != par) return 0;
        par = p;
        const char* base = p ? b : a;
        g.base[f] = (const double*)base;
        if (ve && f >= PTS_TXX && (const char*)in.p[f + (PTS_TXX_O - PTS_TXX)] != base + 2 * slot) return 0;
    }
    g.g = dp ? in.p[PTS_MUSTAU] : in.p[PTS_GDTAU];
    g.mu = in.p[PTS_MUS];
    g.rh = in.p[PTS_DTAU_RHO];
    g.rg = in.p[PTS_RHOG];
    if (!g.g || !g.mu || (!dp && !g.rh) || (bf && !g.rg)) return 0;
    // largest byte offset: the last slot plus the box, one plane of Vz beyond
    // and the guard zone (k3_iter2 stays inside the guard zones)
    const int64_t box = ((int64_t)in.sxy * (in.nz + 1) + 32 * (int64_t)in.sx + 256) * (int64_t)sizeof(double);
    const bool fits32 = (ve ? 2 : 1) * slot + box < ((int64_t)1 << 32);
    if (!fits32 && 5 * (int64_t)in.sxy * (int64_t)sizeof(double) >= ((int64_t)1 << 31)) return 0;
    g.off_in = (uint64_t)(par ? slot : 0);
    g.off_out = (uint64_t)(par ? 0 : slot);
    g.off_old = (uint64_t)(2 * slot);
    g.sx = in.sx;
    g.sxy = in.sxy;
    g.nx = in.nx;
    g.ny = in.ny;
    g.nz = in.nz;
    return (fits32 && !w64) ? 1 : 2;
}

}  // namespace

hipError_t launch3_iter(hipStream_t s, const Fields& in, const Fields& out, const StepScalars& c, bool ve, bool mul,
                        bool bf, int Lz, int by, bool dp, int variant, int64_t slot_bytes, bool mx,
                        bool map_flat, bool w64, bool sl) {
    if (Lz == 0) Lz = 1;
    Fields3g g;
    const int gm = (variant == 1 || variant == 2) ? group_fields(in, out, ve, dp, bf, slot_bytes, w64 && dp, g) : 0;
    // per-slot pointers are instantiated for the derived-parameter kernels
    if (gm == 1 || (gm == 2 && dp)) {
        const bool pf = variant == 2;
        const bool wide = gm == 2;
        int e2 = 0;
        const bool g2 = ve && std::frexp(c.Gdt, &e2) == 0.5;  // G·dt a power of two
        StepScalars c2 = c;
        c2.iGdt = g2 ? 1.0 / c.Gdt : 0.0;
        if (ve) {
            if (mul) {
                if (dp) { if (bf) go3b_pf<true, true, true, true>(s, g, c2, Lz, pf, g2, mx, map_flat, wide, sl); else go3b_pf<true, true, false, true>(s, g, c2, Lz, pf, g2, mx, map_flat, wide, sl); }
                else { if (bf) go3b_pf<true, true, true, false>(s, g, c2, Lz, pf, g2, mx, map_flat, wide, sl); else go3b_pf<true, true, false, false>(s, g, c2, Lz, pf, g2, mx, map_flat, wide, sl); }
            } else {
                if (dp) { if (bf) go3b_pf<true, false, true, true>(s, g, c2, Lz, pf, g2, mx, map_flat, wide, sl); else go3b_pf<true, false, false, true>(s, g, c2, Lz, pf, g2, mx, map_flat, wide, sl); }
                else { if (bf) go3b_pf<true, false, true, false>(s, g, c2, Lz, pf, g2, mx, map_flat, wide, sl); else go3b_pf<true, false, false, false>(s, g, c2, Lz, pf, g2, mx, map_flat, wide, sl); }
            }
        } else {
            if (mul) {
                if (dp) { if (bf) go3b_pf<false, true, true, true>(s, g, c2, Lz, pf, false, mx, map_flat, wide, sl); else go3b_pf<false, true, false, true>(s, g, c2, Lz, pf, false, mx, map_flat, wide, sl); }
                else { if (bf) go3b_pf<false, true, true, false>(s, g, c2, Lz, pf, false, mx, map_flat, wide, sl); else go3b_pf<false, true, false, false>(s, g, c2, Lz, pf, false, mx, map_flat, wide, sl); }
            } else {
                if (dp) { if (bf) go3b_pf<false, false, true, true>(s, g, c2, Lz, pf, false, mx, map_flat, wide, sl); else go3b_pf<false, false, false, true>(s, g, c2, Lz, pf, false, mx, map_flat, wide, sl); }
                else { if (bf) go3b_pf<false, false, true, false>(s, g, c2, Lz, pf, false, mx, map_flat, wide, sl); else go3b_pf<false, false, false, false>(s, g, c2, Lz, pf, false, mx, map_flat, wide, sl); }
            }
        }
        return hipGetLastError();
    }
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
