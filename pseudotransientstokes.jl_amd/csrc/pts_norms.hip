// pts_norms.hip -- residual-norm reductions for the convergence check
// (Stokes2D_ve.jl:164-170, Stokes3D_ve_perf.jl:211-218): min/max of Vx and
// Pt over the whole arrays, and Σ R² and max|R| of Rx, Ry, (Rz,) ∇V.
//
// Two passes with a fixed grid (kRedBlocks blocks, 256 threads), so the
// result is deterministic run to run: pass 1 writes one partial per block,
// pass 2 folds them in one block.  Within a block each wavefront reduces by
// cross-lane shuffles and the four wave totals meet in LDS.  "Live" mode
// computes R and ∇V on the fly from the state (no residual arrays are written
// or read); "stored" mode reduces the Rx/Ry/Rz/∇V arrays written by
// compute_Res!.
#include "pts_internal.hpp"

namespace pts {

namespace {

constexpr int kThreads = 256;

struct Acc {
    double v[kRedVals];
};

__device__ __forceinline__ double nanmax(double a, double b) { return (b != b || b > a) ? b : a; }
__device__ __forceinline__ double nanmin(double a, double b) { return (b != b || b < a) ? b : a; }

__device__ __forceinline__ void acc_init(Acc& a) {
    a.v[0] = INFINITY; a.v[1] = -INFINITY; a.v[2] = INFINITY; a.v[3] = -INFINITY;
    for (int q = 4; q < 8; ++q) a.v[q] = 0.0;
    for (int q = 8; q < 12; ++q) a.v[q] = 0.0;
}
__device__ __forceinline__ void acc_res(Acc& a, int comp, double r) {
    a.v[4 + comp] += r * r;
    a.v[8 + comp] = nanmax(a.v[8 + comp], fabs(r));
}
__device__ __forceinline__ void acc_merge(Acc& a, const Acc& b) {
    a.v[0] = nanmin(a.v[0], b.v[0]);
    a.v[1] = nanmax(a.v[1], b.v[1]);
    a.v[2] = nanmin(a.v[2], b.v[2]);
    a.v[3] = nanmax(a.v[3], b.v[3]);
    for (int q = 4; q < 8; ++q) a.v[q] += b.v[q];
    for (int q = 8; q < 12; ++q) a.v[q] = nanmax(a.v[q], b.v[q]);
}

// Butterfly over the 64 lanes of a wavefront (cross-lane shuffles, no LDS
// memory): every lane ends with the wave's total.  acc_merge is symmetric in
// its arguments, so both lanes of a pair compute the same bits.
__device__ __forceinline__ void wave_reduce(Acc& a) {
    for (int m = 32; m >= 1; m >>= 1) {
        Acc b;
        for (int q = 0; q < kRedVals; ++q) b.v[q] = __shfl_xor(a.v[q], m, 64);
        acc_merge(a, b);
    }
}

// Wave totals through LDS (one barrier), folded by thread 0 in wave order.
__device__ void block_reduce_store(Acc& a, double* dst) {
    constexpr int kWaves = kThreads / 64;
    __shared__ double sm[kWaves][kRedVals];
    wave_reduce(a);
    const int t = threadIdx.x, w = t / 64;
    if ((t & 63) == 0)
        for (int q = 0; q < kRedVals; ++q) sm[w][q] = a.v[q];
    __syncthreads();
    if (t == 0) {
        Acc r;
        for (int q = 0; q < kRedVals; ++q) r.v[q] = sm[0][q];
        for (int v = 1; v < kWaves; ++v) {
            Acc b;
            for (int q = 0; q < kRedVals; ++q) b.v[q] = sm[v][q];
            acc_merge(r, b);
        }
        for (int q = 0; q < kRedVals; ++q) dst[q] = r.v[q];
    }
}

template <bool MUL>
__device__ __forceinline__ double dv_at(const Fields& v, const StepScalars& c, int o, int ndim) {
    double d = dfd<MUL>(v.p[PTS_VX][o + 1] - v.p[PTS_VX][o], c.dx, c._dx) +
               dfd<MUL>(v.p[PTS_VY][o + v.sx] - v.p[PTS_VY][o], c.dy, c._dy);
    if (ndim == 3) d = d + dfd<MUL>(v.p[PTS_VZ][o + v.sxy] - v.p[PTS_VZ][o], c.dz, c._dz);
    return d;
}

// residual formulas, duplicated from the step kernels (kept identical)
template <bool MUL>
__device__ __forceinline__ double rx2(const Fields& f, const StepScalars& c, int o) {
    const int sx = f.sx;
    return dfd<MUL>(f.p[PTS_TXX][o + sx + 1] - f.p[PTS_TXX][o + sx], c.dx, c._dx) +
           dfd<MUL>(f.p[PTS_TXY][o + sx] - f.p[PTS_TXY][o], c.dy, c._dy) -
           dfd<MUL>(f.p[PTS_PT][o + sx + 1] - f.p[PTS_PT][o + sx], c.dx, c._dx);
}
template <bool MUL>
__device__ __forceinline__ double ry2(const Fields& f, const StepScalars& c, int o, bool bf) {
    const int sx = f.sx;
    double R = dfd<MUL>(f.p[PTS_TYY][o + sx + 1] - f.p[PTS_TYY][o + 1], c.dy, c._dy) +
               dfd<MUL>(f.p[PTS_TXY][o + 1] - f.p[PTS_TXY][o], c.dx, c._dx) -
               dfd<MUL>(f.p[PTS_PT][o + sx + 1] - f.p[PTS_PT][o + 1], c.dy, c._dy);
    if (bf) R = R - (f.p[PTS_RHOG][o + 1] + f.p[PTS_RHOG][o + sx + 1]) * 0.5;
    return R;
}
template <bool MUL>
__device__ __forceinline__ double rx3(const Fields& f, const StepScalars& c, int o) {
    const int sx = f.sx, sxy = f.sxy;
    return dfd<MUL>(f.p[PTS_TXX][o + 1] - f.p[PTS_TXX][o], c.dx, c._dx) +
           dfd<MUL>(f.p[PTS_TXY][o + sx] - f.p[PTS_TXY][o], c.dy, c._dy) +
           dfd<MUL>(f.p[PTS_TXZ][o + sxy] - f.p[PTS_TXZ][o], c.dz, c._dz) -
           dfd<MUL>(f.p[PTS_PT][o + 1 + sx + sxy] - f.p[PTS_PT][o + sx + sxy], c.dx, c._dx);
}
template <bool MUL>
__device__ __forceinline__ double ry3(const Fields& f, const StepScalars& c, int o) {
    const int sx = f.sx, sxy = f.sxy;
    return dfd<MUL>(f.p[PTS_TYY][o + sx] - f.p[PTS_TYY][o], c.dy, c._dy) +
           dfd<MUL>(f.p[PTS_TXY][o + 1] - f.p[PTS_TXY][o], c.dx, c._dx) +
           dfd<MUL>(f.p[PTS_TYZ][o + sxy] - f.p[PTS_TYZ][o], c.dz, c._dz) -
           dfd<MUL>(f.p[PTS_PT][o + 1 + sx + sxy] - f.p[PTS_PT][o + 1 + sxy], c.dy, c._dy);
}
template <bool MUL>
__device__ __forceinline__ double rz3(const Fields& f, const StepScalars& c, int o, bool bf) {
    const int sx = f.sx, sxy = f.sxy;
    double R = dfd<MUL>(f.p[PTS_TZZ][o + sxy] - f.p[PTS_TZZ][o], c.dz, c._dz) +
               dfd<MUL>(f.p[PTS_TXZ][o + 1] - f.p[PTS_TXZ][o], c.dx, c._dx) +
               dfd<MUL>(f.p[PTS_TYZ][o + sx] - f.p[PTS_TYZ][o], c.dy, c._dy) -
               dfd<MUL>(f.p[PTS_PT][o + 1 + sx + sxy] - f.p[PTS_PT][o + 1 + sx], c.dz, c._dz);
    if (bf) R = R - (f.p[PTS_RHOG][o + 1 + sx] + f.p[PTS_RHOG][o + 1 + sx + sxy]) * 0.5;
    return R;
}

// LIVE=false: reduce stored arrays; LIVE=true: compute R, ∇V on the fly.
template <bool LIVE, bool MUL>
__global__ __launch_bounds__(kThreads) void k_norms(Fields f, Fields fv, StepScalars c, int ndim, int bf,
                                                    double* partials) {
    Acc a;
    acc_init(a);
    const int nx = f.nx, ny = f.ny, nz = ndim == 3 ? f.nz : 1;
    const int64_t n1 = nx + 1, n2 = ny + 1, n3 = ndim == 3 ? nz + 1 : 1;
    const int64_t tot = n1 * n2 * n3;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % n1);
        const int64_t rr = q / n1;
        const int j = (int)(rr % n2);
        const int k = (int)(rr / n2);
        const int o = i + f.sx * j + f.sxy * k;
        // Vx over (nx+1) x ny (x nz)
        if (j < ny && k < nz) {
            double v = f.p[PTS_VX][o];
            a.v[0] = nanmin(a.v[0], v);
            a.v[1] = nanmax(a.v[1], v);
        }
        if (i < nx && j < ny && k < nz) {
            double p = f.p[PTS_PT][o];
            a.v[2] = nanmin(a.v[2], p);
            a.v[3] = nanmax(a.v[3], p);
            double d = LIVE ? dv_at<MUL>(fv, c, o, ndim) : f.p[PTS_DIVV][o];
            acc_res(a, 3, d);
        }
        if (ndim == 2) {
            if (i < nx - 1 && j < ny - 2) acc_res(a, 0, LIVE ? rx2<MUL>(f, c, o) : f.p[PTS_RX][o]);
            if (i < nx - 2 && j < ny - 1) acc_res(a, 1, LIVE ? ry2<MUL>(f, c, o, bf != 0) : f.p[PTS_RY][o]);
        } else {
            if (i < nx - 1 && j < ny - 2 && k < nz - 2) acc_res(a, 0, LIVE ? rx3<MUL>(f, c, o) : f.p[PTS_RX][o]);
            if (i < nx - 2 && j < ny - 1 && k < nz - 2) acc_res(a, 1, LIVE ? ry3<MUL>(f, c, o) : f.p[PTS_RY][o]);
            if (i < nx - 2 && j < ny - 2 && k < nz - 1)
                acc_res(a, 2, LIVE ? rz3<MUL>(f, c, o, bf != 0) : f.p[PTS_RZ][o]);
        }
    }
    block_reduce_store(a, partials + (int64_t)blockIdx.x * kRedVals);
}

__global__ __launch_bounds__(kThreads) void k_norms_final(const double* partials, int nparts, double* out) {
    Acc a;
    acc_init(a);
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
        Acc p;
        for (int q = 0; q < kRedVals; ++q) p.v[q] = partials[(int64_t)b * kRedVals + q];
        acc_merge(a, p);
    }
    block_reduce_store(a, out);
}

// The same reductions over dense arrays (the reference's own shapes, column
// major): min/max of Vx and Pt, Σ R² and max|R| of each residual; a null
// residual pointer contributes nothing.
__global__ __launch_bounds__(kThreads) void k_norms_dense(DenseNorms d, double* partials) {
    Acc a;
    acc_init(a);
    const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = t0; q < d.nvx; q += stride) {
        const double v = d.vx[q];
        a.v[0] = nanmin(a.v[0], v);
        a.v[1] = nanmax(a.v[1], v);
    }
    for (int64_t q = t0; q < d.npt; q += stride) {
        const double v = d.pt[q];
        a.v[2] = nanmin(a.v[2], v);
        a.v[3] = nanmax(a.v[3], v);
    }
    for (int c = 0; c < 4; ++c) {
        const double* r = d.r[c];
        if (!r) continue;
        for (int64_t q = t0; q < d.nr[c]; q += stride) acc_res(a, c, r[q]);
    }
    block_reduce_store(a, partials + (int64_t)blockIdx.x * kRedVals);
}

}  // namespace

hipError_t launch_norms_dense(hipStream_t s, const DenseNorms& d, double* partials, int nblocks, double* out) {
    k_norms_dense<<<nblocks, kThreads, 0, s>>>(d, partials);
    k_norms_final<<<1, kThreads, 0, s>>>(partials, nblocks, out);
    return hipGetLastError();
}

hipError_t launch_norms_stored(hipStream_t s, const Fields& f, int ndim, double* partials, double* out) {
    StepScalars c{};
    k_norms<false, false><<<kRedBlocks, kThreads, 0, s>>>(f, f, c, ndim, 0, partials);
    k_norms_final<<<1, kThreads, 0, s>>>(partials, kRedBlocks, out);
    return hipGetLastError();
}

hipError_t launch_norms_live(hipStream_t s, const Fields& f, const Fields& fv, const StepScalars& c, int ndim,
                             bool bf, bool mul, double* partials, double* out) {
    if (mul) k_norms<true, true><<<kRedBlocks, kThreads, 0, s>>>(f, fv, c, ndim, bf ? 1 : 0, partials);
    else k_norms<true, false><<<kRedBlocks, kThreads, 0, s>>>(f, fv, c, ndim, bf ? 1 : 0, partials);
    k_norms_final<<<1, kThreads, 0, s>>>(partials, kRedBlocks, out);
    return hipGetLastError();
}

}  // namespace pts
