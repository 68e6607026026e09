// pts_adjoint.hip -- adjoint of the converged 2D PT iteration and the
// gradient with respect to log viscosity (no counterpart in the reference).
//
// One forward PT iteration (multiply form) is the affine map
//   Φ = B ∘ S_V ∘ S_Pτ :
//   S_Pτ:  P' = P - r Gdτ ⊙ (D V),  τ' = a ⊙ (τ + Gr τ_o) + b ⊙ (E V),
//          a = 1/(1 + Gdτ/μ + Gr), b = 2 Gdτ a, Gr = Gdτ/(G dt) (0 viscous)
//          (vertex means of Gdτ and μ for τxy)
//   S_V :  V_int' = V_int + Rh ⊙ (Dτ τ' - ∇P' [- ρg])   (Rh: face mean of dτ_Rho)
//   B   :  free-slip copies V[:,0] = V[:,1], ... (bc_y!(Vx), bc_x!(Vy))
// D is the discrete divergence, E the strain rate (ε̇xx, ε̇yy, ε̇xy), Dτ the
// stress divergence.  At the converged state x* = Φ(x*), the adjoint of
// J(x*) is the fixed point of
//   λ ← Φᵀ λ + ∂J/∂x = S_Pτᵀ S_Vᵀ Bᵀ λ + g,
// which contracts at the rate of the forward iteration; no iteration history
// is needed.  dJ/dm = (S_Vᵀ Bᵀ λ)_τ · ∂τ'/∂m.  Over visco-elastic physical
// steps the converged τ of one step is the next step's τ_o, which adds
// a·Gr ⊙ (S_Vᵀ Bᵀ λ_next)_τ to the previous step's source.
//
// pass 1 (Bᵀ, S_Vᵀ): fold the boundary-copy adjoints onto their sources,
//   w = Rh ⊙ λV_int,  λP += D(w)-like divergence,  λτ += -Dτᵀ w.
// pass 2 (S_Pτᵀ, + g): λτ ← a ⊙ λτ,  λV ← λV - Dᵀ(r Gdτ ⊙ λP) + Eᵀ(b ⊙ λτ) + g,
//   with λ = 0 on the Dirichlet (never-updated) velocity entries.
// The adjoint fields use the field ids of the forward ones.
#include <algorithm>

#include "pts_internal.hpp"

namespace pts {

namespace {

constexpr int BX = 64, BY = 4;

// adjoint weight w at a Vx face (I,J): Rh ⊙ folded λVx, zero off the interior
__device__ __forceinline__ double wx_at(const Fields& f, const Fields& lam, int I, int J) {
    const int nx = f.nx, ny = f.ny, sx = f.sx;
    if (I < 1 || I > nx - 1 || J < 1 || J > ny - 2) return 0.0;
    const int o = I + sx * J;
    const double* L = lam.p[PTS_VX];
    double l = L[o];
    if (J == 1) l += L[o - sx];
    if (J == ny - 2) l += L[o + sx];
    const double* Rh = f.p[PTS_DTAU_RHO];
    return l * (0.5 * (Rh[o - 1] + Rh[o]));
}
__device__ __forceinline__ double wy_at(const Fields& f, const Fields& lam, int I, int J) {
    const int nx = f.nx, ny = f.ny, sx = f.sx;
    if (I < 1 || I > nx - 2 || J < 1 || J > ny - 1) return 0.0;
    const int o = I + sx * J;
    const double* L = lam.p[PTS_VY];
    double l = L[o];
    if (I == 1) l += L[o - 1];
    if (I == nx - 2) l += L[o + 1];
    const double* Rh = f.p[PTS_DTAU_RHO];
    return l * (0.5 * (Rh[o - sx] + Rh[o]));
}

// pass 1: in = λ, out = (λP', λτ', λV folded)
__global__ __launch_bounds__(BX* BY) void k2_adj1(Fields f, Fields lam, Fields out, StepScalars c) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    const int nx = f.nx, ny = f.ny, sx = f.sx;
    if (i > nx || j > ny) return;
    const int o = i + sx * j;
    if (i < nx && j < ny) {
        const double wx0 = wx_at(f, lam, i, j), wx1 = wx_at(f, lam, i + 1, j);
        const double wy0 = wy_at(f, lam, i, j), wy1 = wy_at(f, lam, i, j + 1);
        out.p[PTS_PT][o] = lam.p[PTS_PT][o] + c._dx * (wx1 - wx0) + c._dy * (wy1 - wy0);
        out.p[PTS_TXX][o] = lam.p[PTS_TXX][o] + c._dx * (wx0 - wx1);
        out.p[PTS_TYY][o] = lam.p[PTS_TYY][o] + c._dy * (wy0 - wy1);
    }
    if (i < nx - 1 && j < ny - 1) {  // vertex (i,j): x-face i+1, y-face j+1
        out.p[PTS_TXY][o] = lam.p[PTS_TXY][o] + c._dy * (wx_at(f, lam, i + 1, j) - wx_at(f, lam, i + 1, j + 1)) +
                            c._dx * (wy_at(f, lam, i, j + 1) - wy_at(f, lam, i + 1, j + 1));
    }
    // folded velocity adjoints: sources get their copies, copies are zeroed
    if (j < ny) {
        const double* L = lam.p[PTS_VX];
        double v = 0.0;
        if (i >= 1 && i <= nx - 1 && j >= 1 && j <= ny - 2) {
            v = L[o];
            if (j == 1) v += L[o - sx];
            if (j == ny - 2) v += L[o + sx];
        }
        out.p[PTS_VX][o] = v;
    }
    if (i < nx) {
        const double* L = lam.p[PTS_VY];
        double v = 0.0;
        if (i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 1) {
            v = L[o];
            if (i == 1) v += L[o - 1];
            if (i == nx - 2) v += L[o + 1];
        }
        out.p[PTS_VY][o] = v;
    }
}

// a = 1/(1 + Gdτ/μ + Gr) of the τ update, Gr = Gdτ/(G·dt) (visco-elastic;
// igdt = 1/(G·dt)) or 0 (viscous, igdt = 0); b = 2·Gdτ·a.
__device__ __forceinline__ double acell(const Fields& f, int o, double igdt) {
    const double g = f.p[PTS_GDTAU][o], m = f.p[PTS_MUS][o];
    return 1.0 / (1.0 + g / m + g * igdt);
}
// b ⊙ λτ at cells (normal) and vertices (shear), zero off range
__device__ __forceinline__ double pxx(const Fields& f, const Fields& l1, double igdt, int i, int j) {
    if (i < 0 || i >= f.nx || j < 0 || j >= f.ny) return 0.0;
    const int o = i + f.sx * j;
    return 2.0 * f.p[PTS_GDTAU][o] * acell(f, o, igdt) * l1.p[PTS_TXX][o];
}
__device__ __forceinline__ double pyy(const Fields& f, const Fields& l1, double igdt, int i, int j) {
    if (i < 0 || i >= f.nx || j < 0 || j >= f.ny) return 0.0;
    const int o = i + f.sx * j;
    return 2.0 * f.p[PTS_GDTAU][o] * acell(f, o, igdt) * l1.p[PTS_TYY][o];
}
// vertex means of Gdτ and μ, and the vertex coefficients a, b
__device__ __forceinline__ void vert_coef(const Fields& f, int o, double igdt, double& ag, double& am, double& a,
                                          double& b) {
    const int sx = f.sx;
    const double* Gd = f.p[PTS_GDTAU];
    const double* Mu = f.p[PTS_MUS];
    ag = 0.25 * (Gd[o] + Gd[o + 1] + Gd[o + sx] + Gd[o + sx + 1]);
    am = 0.25 * (Mu[o] + Mu[o + 1] + Mu[o + sx] + Mu[o + sx + 1]);
    a = 1.0 / (1.0 + ag / am + ag * igdt);
    b = 2.0 * ag * a;
}
__device__ __forceinline__ double pxy(const Fields& f, const Fields& l1, double igdt, int i, int j) {
    if (i < 0 || i >= f.nx - 1 || j < 0 || j >= f.ny - 1) return 0.0;
    const int o = i + f.sx * j;
    double ag, am, a, b;
    vert_coef(f, o, igdt, ag, am, a, b);
    return b * l1.p[PTS_TXY][o];
}
__device__ __forceinline__ double qP(const Fields& f, const Fields& l1, double r, int i, int j) {
    if (i < 0 || i >= f.nx || j < 0 || j >= f.ny) return 0.0;
    const int o = i + f.sx * j;
    return r * f.p[PTS_GDTAU][o] * l1.p[PTS_PT][o];
}

// pass 2: in = pass-1 output, out = new λ; g holds the sources (∂J/∂x) under
// the field ids VX, VY, TXX, TYY, TXY (null: none)
__global__ __launch_bounds__(BX* BY) void k2_adj2(Fields f, Fields l1, Fields out, StepScalars c, double igdt,
                                                  Fields g) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    const int nx = f.nx, ny = f.ny, sx = f.sx;
    if (i > nx || j > ny) return;
    const int o = i + sx * j;
    if (i < nx && j < ny) {
        out.p[PTS_PT][o] = l1.p[PTS_PT][o];
        const double a = acell(f, o, igdt);
        double txx = a * l1.p[PTS_TXX][o], tyy = a * l1.p[PTS_TYY][o];
        if (g.p[PTS_TXX]) txx += g.p[PTS_TXX][o];
        if (g.p[PTS_TYY]) tyy += g.p[PTS_TYY][o];
        out.p[PTS_TXX][o] = txx;
        out.p[PTS_TYY][o] = tyy;
    }
    if (i < nx - 1 && j < ny - 1) {
        double ag, am, a, b;
        vert_coef(f, o, igdt, ag, am, a, b);
        double txy = a * l1.p[PTS_TXY][o];
        if (g.p[PTS_TXY]) txy += g.p[PTS_TXY][o];
        out.p[PTS_TXY][o] = txy;
    }
    if (j < ny) {  // Vx face (i, j); Dirichlet faces i = 0, nx carry no adjoint
        double v = 0.0;
        if (i >= 1 && i <= nx - 1) {
            v = l1.p[PTS_VX][o] - c._dx * (qP(f, l1, c.r, i - 1, j) - qP(f, l1, c.r, i, j)) +
                c._dx * (pxx(f, l1, igdt, i - 1, j) - pxx(f, l1, igdt, i, j)) +
                0.5 * c._dy * (pxy(f, l1, igdt, i - 1, j - 1) - pxy(f, l1, igdt, i - 1, j));
            if (g.p[PTS_VX]) v += g.p[PTS_VX][o];
        }
        out.p[PTS_VX][o] = v;
    }
    if (i < nx) {  // Vy face (i, j); Dirichlet rows j = 0, ny
        double v = 0.0;
        if (j >= 1 && j <= ny - 1) {
            v = l1.p[PTS_VY][o] - c._dy * (qP(f, l1, c.r, i, j - 1) - qP(f, l1, c.r, i, j)) +
                c._dy * (pyy(f, l1, igdt, i, j - 1) - pyy(f, l1, igdt, i, j)) +
                0.5 * c._dx * (pxy(f, l1, igdt, i - 1, j - 1) - pxy(f, l1, igdt, i, j - 1));
            if (g.p[PTS_VY]) v += g.p[PTS_VY][o];
        }
        out.p[PTS_VY][o] = v;
    }
}

// ∂J/∂ln μ at cells from the pass-1 adjoint of the converged λ and the
// converged stresses τ* (= τ' at the fixed point), added to grad when acc.
// τ' = (τ + τ_o·Gr + 2Gdτ·ε̇)/D with D = 1 + Gdτ/μ + Gr gives
// ∂τ'/∂ln μ = τ'·(Gdτ/μ)/D; the vertex mean of μ contributes μ/4 per cell.
__global__ __launch_bounds__(BX* BY) void k2_adj_grad(Fields f, Fields l1, double igdt, double* grad, int acc) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    const int nx = f.nx, ny = f.ny, sx = f.sx;
    if (i >= nx || j >= ny) return;
    const int o = i + sx * j;
    const double* Gd = f.p[PTS_GDTAU];
    const double* Mu = f.p[PTS_MUS];
    const double gm = Gd[o] / Mu[o];
    double g = (l1.p[PTS_TXX][o] * f.p[PTS_TXX][o] + l1.p[PTS_TYY][o] * f.p[PTS_TYY][o]) *
               (gm / (1.0 + gm + Gd[o] * igdt));
    // shear: the (up to) four vertices around cell (i,j)
    for (int dj = -1; dj <= 0; ++dj)
        for (int di = -1; di <= 0; ++di) {
            const int vi = i + di, vj = j + dj;
            if (vi < 0 || vi >= nx - 1 || vj < 0 || vj >= ny - 1) continue;
            const int v = vi + sx * vj;
            double ag, am, a, b;
            vert_coef(f, v, igdt, ag, am, a, b);
            g += l1.p[PTS_TXY][v] * f.p[PTS_TXY][v] * 0.25 * Mu[o] * (ag / (am * am)) * a;
        }
    const int64_t q = i + (int64_t)nx * j;
    grad[q] = acc ? grad[q] + g : g;
}

// The τ_o dependence of the previous physical step: its converged τ is this
// step's τ_o, entering τ' with the factor a·Gr.  gτ = a·Gr ⊙ (pass-1 λ)_τ is
// the previous step's adjoint source on τ.
__global__ __launch_bounds__(BX* BY) void k2_adj_couple(Fields f, Fields l1, double igdt, Fields g) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    const int nx = f.nx, ny = f.ny, sx = f.sx;
    if (i >= nx || j >= ny) return;
    const int o = i + sx * j;
    const double a = acell(f, o, igdt), Gr = f.p[PTS_GDTAU][o] * igdt;
    g.p[PTS_TXX][o] = a * Gr * l1.p[PTS_TXX][o];
    g.p[PTS_TYY][o] = a * Gr * l1.p[PTS_TYY][o];
    if (i < nx - 1 && j < ny - 1) {
        double ag, am, av, b;
        vert_coef(f, o, igdt, ag, am, av, b);
        g.p[PTS_TXY][o] = av * (ag * igdt) * l1.p[PTS_TXY][o];
    }
}

// NaN-propagating max (fmax drops a NaN operand)
__device__ __forceinline__ double nmax(double a, double b) { return (a != a) ? a : ((b != b) ? b : fmax(a, b)); }

// max |λV - λV_prev| and max |λV| over both velocity components
__global__ __launch_bounds__(256) void k2_adj_change(Fields f, Fields a, Fields b, double* part) {
    __shared__ double s0[256], s1[256];
    const int nx = f.nx, ny = f.ny, sx = f.sx;
    const int64_t tot = (int64_t)(nx + 1) * (ny + 1);
    double d = 0.0, m = 0.0;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % (nx + 1)), j = (int)(q / (nx + 1));
        const int o = i + sx * j;
        if (j < ny) {
            d = nmax(d, fabs(a.p[PTS_VX][o] - b.p[PTS_VX][o]));
            m = nmax(m, fabs(a.p[PTS_VX][o]));
        }
        if (i < nx) {
            d = nmax(d, fabs(a.p[PTS_VY][o] - b.p[PTS_VY][o]));
            m = nmax(m, fabs(a.p[PTS_VY][o]));
        }
    }
    s0[threadIdx.x] = d;
    s1[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s0[threadIdx.x] = nmax(s0[threadIdx.x], s0[threadIdx.x + w]);
            s1[threadIdx.x] = nmax(s1[threadIdx.x], s1[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s0[0];
        part[2 * blockIdx.x + 1] = s1[0];
    }
}

__global__ __launch_bounds__(256) void k_max2_final(const double* part, int nb, double* out) {
    __shared__ double s0[256], s1[256];
    const int t = threadIdx.x;
    double d = 0.0, m = 0.0;
    for (int b = t; b < nb; b += 256) {
        d = nmax(d, part[2 * b]);
        m = nmax(m, part[2 * b + 1]);
    }
    s0[t] = d;
    s1[t] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            s0[t] = nmax(s0[t], s0[t + w]);
            s1[t] = nmax(s1[t], s1[t + w]);
        }
        __syncthreads();
    }
    if (t == 0) {
        out[0] = s0[0];
        out[1] = s1[0];
    }
}

// ---------------------------------------------------------------------------
// One adjoint iteration in a single pass (k2_adj1 and k2_adj2 fused): one
// 64-lane wave per strip of 60 columns (lanes 0, 1, 62, 63 are halo: a
// velocity entry of λ' needs pass-1 values two columns to its left), marching
// a chunk of rows in y after two warm-up rows.  At row step J lane c forms
//   wx at face (c, J) and wy at face (c, J+1)   (pass 1, folded copies)
//   l1 of cell (c, J) and of vertex (c, J-1)    (pass 1)
//   λ' of cell (c, J), vertex (c, J-1), Vx face (c, J-1), Vy face (c, J)
// with the expressions of k2_adj1 / k2_adj2 operation for operation, so both
// forms give the same bits.  Values of the previous row are carried in
// registers, x-neighbours come from DPP lane shifts.
// ---------------------------------------------------------------------------
constexpr int kAdjW = 60;

__global__ __launch_bounds__(64) void k2_adj_iter(Fields f, Fields lam, Fields out, StepScalars cs, double igdt,
                                                  Fields g, int L, int nstrips) {
    const int lane = threadIdx.x;
    const int b = xcd_tile(blockIdx.x, gridDim.x);
    const int strip = b % nstrips, chunk = b / nstrips;
    const int c = strip * kAdjW - 2 + lane;
    const int nx = f.nx, ny = f.ny, sx = f.sx;
    const int jb = chunk * L, je = min(jb + L, ny);
    const bool own = lane >= 2 && lane <= kAdjW + 1;
    const int cc = min(max(c, 0), nx - 1), cf = min(max(c, 0), nx), cv = min(max(c, 0), nx - 2);
    const double _dx = cs._dx, _dy = cs._dy, r = cs.r;
    const double* __restrict__ LP = lam.p[PTS_PT];
    const double* __restrict__ Lxx = lam.p[PTS_TXX];
    const double* __restrict__ Lyy = lam.p[PTS_TYY];
    const double* __restrict__ Lxy = lam.p[PTS_TXY];
    const double* __restrict__ LVx = lam.p[PTS_VX];
    const double* __restrict__ LVy = lam.p[PTS_VY];
    const double* __restrict__ Gdp = f.p[PTS_GDTAU];
    const double* __restrict__ Mup = f.p[PTS_MUS];
    const double* __restrict__ Rhp = f.p[PTS_DTAU_RHO];
    const double *gx = g.p[PTS_VX], *gy = g.p[PTS_VY], *gxx = g.p[PTS_TXX], *gyy = g.p[PTS_TYY],
                 *gxy = g.p[PTS_TXY];
    const bool cx_in = c >= 1 && c <= nx - 1, cy_in = c >= 1 && c <= nx - 2;
    const bool cell_c = c >= 0 && c < nx, vert_c = c >= 0 && c <= nx - 2;

    // carried from row J-1 (face rows of wy/fy: row J)
    double wyC = 0.0, fyC = 0.0, wxP = 0.0, fxP = 0.0, lVxP = 0.0;
    double qP = 0.0, pxxP = 0.0, pyyP = 0.0, pxyP = 0.0, GdP = 0.0, MuP = 0.0;
    const int J0 = max(jb - 2, 0), J1 = min(je, ny);
    double RhC = Rhp[cc + sx * min(J0, ny - 1)];
    for (int J = J0; J <= J1; ++J) {
        const int jc = min(J, ny - 1);
        const int oc = cc + sx * jc;
        const double lP = LP[oc], lxx = Lxx[oc], lyy = Lyy[oc], Gd = Gdp[oc], Mu = Mup[oc];
        const double lVx = LVx[cf + sx * jc];
        const double lVyN = LVy[cc + sx * min(J + 1, ny)];
        const double RhN = Rhp[cc + sx * min(J + 1, ny - 1)];
        const double lxy = Lxy[cv + sx * min(max(J - 1, 0), ny - 2)];
        const double lVxNx = (J == ny - 2) ? LVx[cf + sx * (ny - 1)] : 0.0;

        // pass 1: weights at face (c, J) of Vx and face (c, J+1) of Vy
        const bool xin = cx_in && J >= 1 && J <= ny - 2;
        double fx = 0.0;
        if (xin) {
            fx = lVx;
            if (J == 1) fx += lVxP;
            if (J == ny - 2) fx += lVxNx;
        }
        const double RhL = from_left(RhC);
        const double wx = xin ? fx * (0.5 * (RhL + RhC)) : 0.0;
        const bool yin = cy_in && J + 1 <= ny - 1;
        const double lVyNL = from_left(lVyN), lVyNR = from_right(lVyN);
        double fyN = 0.0;
        if (yin) {
            fyN = lVyN;
            if (c == 1) fyN += lVyNL;
            if (c == nx - 2) fyN += lVyNR;
        }
        const double wyN = yin ? fyN * (0.5 * (RhC + RhN)) : 0.0;
        // pass 1: cell (c, J) and vertex (c, J-1)
        const double wxR = from_right(wx), wxPR = from_right(wxP), wyCR = from_right(wyC);
        const double l1P = lP + _dx * (wxR - wx) + _dy * (wyN - wyC);
        const double l1xx = lxx + _dx * (wx - wxR);
        const double l1yy = lyy + _dy * (wyC - wyN);
        const double l1xy = lxy + _dy * (wxPR - wxR) + _dx * (wyC - wyCR);
        // pass 2 coefficients: cell (c, J) and vertex (c, J-1)
        const double a = 1.0 / (1.0 + Gd / Mu + Gd * igdt);
        const double GdR = from_right(Gd), MuR = from_right(Mu), GdPR = from_right(GdP), MuPR = from_right(MuP);
        const double ag = 0.25 * (GdP + GdPR + Gd + GdR);
        const double am = 0.25 * (MuP + MuPR + Mu + MuR);
        const double av = 1.0 / (1.0 + ag / am + ag * igdt);
        const double bv = 2.0 * ag * av;
        const bool cell = cell_c && J < ny;
        const bool vert = vert_c && J >= 1 && J - 1 <= ny - 2;
        const double q = cell ? r * Gd * l1P : 0.0;
        const double pxx = cell ? 2.0 * Gd * a * l1xx : 0.0;
        const double pyy = cell ? 2.0 * Gd * a * l1yy : 0.0;
        const double pxy = vert ? bv * l1xy : 0.0;
        const double qPL = from_left(qP), pxxPL = from_left(pxxP), pxyPL = from_left(pxyP), pxyL = from_left(pxy);

        if (own) {
            if (cell && J >= jb && J < je) {
                const int o = c + sx * J;
                double txx = a * l1xx, tyy = a * l1yy;
                if (gxx) txx += gxx[o];
                if (gyy) tyy += gyy[o];
                out.p[PTS_PT][o] = l1P;
                out.p[PTS_TXX][o] = txx;
                out.p[PTS_TYY][o] = tyy;
            }
            if (vert && J - 1 >= jb && J - 1 < je) {
                const int o = c + sx * (J - 1);
                double txy = av * l1xy;
                if (gxy) txy += gxy[o];
                out.p[PTS_TXY][o] = txy;
            }
            if (J >= 1 && J - 1 >= jb && J - 1 < je && c >= 0 && c <= nx) {  // Vx face (c, J-1)
                const int o = c + sx * (J - 1);
                double v = 0.0;
                if (c >= 1 && c <= nx - 1) {
                    v = fxP - _dx * (qPL - qP) + _dx * (pxxPL - pxxP) + 0.5 * _dy * (pxyPL - pxyL);
                    if (gx) v += gx[o];
                }
                out.p[PTS_VX][o] = v;
            }
            if (((J >= jb && J < je) || (J == ny && je == ny)) && c >= 0 && c < nx) {  // Vy face (c, J)
                const int o = c + sx * J;
                double v = 0.0;
                if (J >= 1 && J <= ny - 1) {
                    v = fyC - _dy * (qP - q) + _dy * (pyyP - pyy) + 0.5 * _dx * (pxyL - pxy);
                    if (gy) v += gy[o];
                }
                out.p[PTS_VY][o] = v;
            }
        }
        // carry
        lVxP = lVx;
        RhC = RhN;
        wyC = wyN;
        fyC = fyN;
        wxP = wx;
        fxP = fx;
        qP = q;
        pxxP = pxx;
        pyyP = pyy;
        pxyP = pxy;
        GdP = Gd;
        MuP = Mu;
    }
}

inline dim3 grid2(int n1, int n2) { return dim3((n1 + BX - 1) / BX, (n2 + BY - 1) / BY, 1); }

}  // namespace

hipError_t launch2_adj_step(hipStream_t s, const Fields& f, const Fields& lam, const Fields& tmp, const Fields& out,
                            const StepScalars& c, double igdt, const Fields& g, bool fused) {
    if (fused) {
        static int conc = 0;
        if (!conc) {
            int dev = 0, cus = 0, per = 0;
            (void)hipGetDevice(&dev);
            (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
            (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k2_adj_iter, 64, 0);
            conc = std::max(1, cus) * std::max(1, per);
        }
        const int ns = (f.nx + 1 + kAdjW - 1) / kAdjW;
        const int nch = std::max(1, conc / ns);
        const int L = std::max(8, std::min(64, (f.ny + nch - 1) / nch));
        const int nb = ns * ((f.ny + L - 1) / L);
        k2_adj_iter<<<nb, 64, 0, s>>>(f, lam, out, c, igdt, g, L, ns);
        return hipGetLastError();
    }
    k2_adj1<<<grid2(f.nx + 1, f.ny + 1), dim3(BX, BY), 0, s>>>(f, lam, tmp, c);
    k2_adj2<<<grid2(f.nx + 1, f.ny + 1), dim3(BX, BY), 0, s>>>(f, tmp, out, c, igdt, g);
    return hipGetLastError();
}

hipError_t launch2_adj_pass1(hipStream_t s, const Fields& f, const Fields& lam, const Fields& tmp,
                             const StepScalars& c) {
    k2_adj1<<<grid2(f.nx + 1, f.ny + 1), dim3(BX, BY), 0, s>>>(f, lam, tmp, c);
    return hipGetLastError();
}

hipError_t launch2_adj_grad(hipStream_t s, const Fields& f, const Fields& l1, double igdt, double* grad, bool acc) {
    k2_adj_grad<<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s>>>(f, l1, igdt, grad, acc ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch2_adj_couple(hipStream_t s, const Fields& f, const Fields& l1, double igdt, const Fields& g) {
    k2_adj_couple<<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s>>>(f, l1, igdt, g);
    return hipGetLastError();
}

hipError_t launch2_adj_change(hipStream_t s, const Fields& f, const Fields& a, const Fields& b, double* part,
                              double* out) {
    k2_adj_change<<<kRedBlocks, 256, 0, s>>>(f, a, b, part);
    k_max2_final<<<1, 256, 0, s>>>(part, kRedBlocks, out);
    return hipGetLastError();
}

}  // namespace pts
