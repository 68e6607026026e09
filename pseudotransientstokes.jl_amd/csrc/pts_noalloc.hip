// pts_noalloc.hip -- pts_norms_dense_noalloc: the residual-norm reductions
// of pts_norms_stored over caller-owned device arrays, for callers that
// cannot allocate device memory (code under HIP graph capture, serving
// loops).  It calls no hipMalloc/hipFree: the partial results of the first
// reduction pass live in a file-scope device array, one copy per GPU, whose
// address is looked up once per device.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "pts_internal.hpp"

namespace {

// Partial results: one slot of kRedVals doubles for each of the
// kScratchBlocks blocks of the first pass (24 KiB).  The reduction grid is
// fixed, so every input size fits; every call on a GPU shares this array.
constexpr int kScratchBlocks = 256;
__device__ __align__(256) double scratch[kScratchBlocks * pts::kRedVals];

// Device address of `scratch` on each GPU (the host-side name is only a
// stand-in; kernels receive these addresses).
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};

// Largest array the entry point accepts (elements), the library's grid limit.
constexpr int64_t kMaxElems = ((int64_t)1 << 31) - 1;

int arg_error(const std::string& m) {
    pts::set_global_error("pts_norms_dense_noalloc: " + m);
    return PTS_ERR_ARG;
}

}  // namespace

extern "C" int pts_norms_dense_noalloc(int32_t ndim, int32_t nx, int32_t ny, int32_t nz, const double* Vx,
                                       const double* Pt, const double* Rx, const double* Ry, const double* Rz,
                                       const double* divV, void* stream, double* d_out) {
    if (ndim != 2 && ndim != 3) return arg_error("ndim must be 2 or 3");
    if (nx < 3 || ny < 3 || (ndim == 3 && nz < 3)) return arg_error("each of nx, ny (, nz) must be >= 3");
    if (!Vx || !Pt || !Rx || !Ry || !divV || !d_out || (ndim == 3 && !Rz))
        return arg_error("null array (Rz may be null in 2D only)");
    const int64_t z = ndim == 3 ? nz : 1;
    pts::DenseNorms d{};
    d.vx = Vx;
    d.pt = Pt;
    d.r[0] = Rx;
    d.r[1] = Ry;
    d.r[2] = ndim == 3 ? Rz : nullptr;
    d.r[3] = divV;
    // reference shapes (pts_field_shape): Vx (nx+1)·ny·nz, Pt and ∇V nx·ny·nz,
    // Rx (nx-1)(ny-2)(nz-2), Ry (nx-2)(ny-1)(nz-2), Rz (nx-2)(ny-2)(nz-1); 2D
    // without the z factors
    d.nvx = (int64_t)(nx + 1) * ny * z;
    d.npt = (int64_t)nx * ny * z;
    d.nr[3] = d.npt;
    if (ndim == 2) {
        d.nr[0] = (int64_t)(nx - 1) * (ny - 2);
        d.nr[1] = (int64_t)(nx - 2) * (ny - 1);
        d.nr[2] = 0;
    } else {
        d.nr[0] = (int64_t)(nx - 1) * (ny - 2) * (nz - 2);
        d.nr[1] = (int64_t)(nx - 2) * (ny - 1) * (nz - 2);
        d.nr[2] = (int64_t)(nx - 2) * (ny - 2) * (nz - 1);
    }
    if (d.nvx > kMaxElems) return arg_error("grid too large (at most 2^31-1 elements per array)");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        pts::set_global_error(std::string("pts_norms_dense_noalloc: hipGetDevice: ") + hipGetErrorString(e));
        return PTS_ERR_HIP;
    }
    if (dev < 0 || dev >= kMaxDevices) return arg_error("device ordinal out of range");
    double* part = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (!g_scratch[dev]) {
            void* p = nullptr;
            e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) {
                pts::set_global_error(std::string("pts_norms_dense_noalloc: hipGetSymbolAddress: ") +
                                      hipGetErrorString(e));
                return PTS_ERR_HIP;
            }
            g_scratch[dev] = static_cast<double*>(p);
        }
        part = g_scratch[dev];
    }
    e = pts::launch_norms_dense(static_cast<hipStream_t>(stream), d, part, kScratchBlocks, d_out);
    if (e != hipSuccess) {
        pts::set_global_error(std::string("pts_norms_dense_noalloc: launch: ") + hipGetErrorString(e));
        return PTS_ERR_HIP;
    }
    return PTS_OK;
}
