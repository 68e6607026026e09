// pts_api.hip -- the C ABI (include/pts.h): context, device memory, host
// copies, per-kernel entry points and the device-resident PT loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <array>
#include <map>
#include <set>
#include <vector>
#include <string>

#include "pts_internal.hpp"

using namespace pts;

namespace {

// Fields that pts_pt_iterate advances with ping-pong buffers.
constexpr int kState[] = {PTS_PT, PTS_VX, PTS_VY, PTS_VZ, PTS_TXX, PTS_TYY, PTS_TZZ, PTS_TXY, PTS_TXZ, PTS_TYZ};
constexpr int kVel[] = {PTS_VX, PTS_VY, PTS_VZ};

std::string g_create_error;

}  // namespace

struct pts_ctx {
    pts_grid g{};
    int nd = 2, nx = 0, ny = 0, nz = 1;
    int sx = 0, sxy = 0;
    int64_t nalloc = 0;  // elements per field allocation
    double* cur[PTS_NFIELDS] = {};
    double* alt[PTS_NFIELDS] = {};
    float* cur32[PTS_NFIELDS] = {};  // FP32 correction state, constants and sources
    float* alt32[PTS_NFIELDS] = {};
    bool alt_synced = false;
    hipStream_t stream = nullptr;
    double* red_partials = nullptr;
    double* red_out = nullptr;
    double* h_red = nullptr;
    int64_t bytes = 0;
    int rows_per_block = -32;   // 2D single-pass: rows per block (0: one round of resident blocks; <0: that, capped)
    int variant_2d = 7;         // 2D single-pass code variant (prefetch, 8-wave strips)
    bool persist_2d = false;    // 2D single-pass: persistent waves looping over tiles
    // 2D small grids: pts_pt_solve runs each chunk of iterations between checks
    // as one cooperative launch with grid barriers (up to multi_max_cells cells)
    bool multi_2d = true;
    int multi_rows_2d = 2;              // rows per tile of that launch
    int64_t multi_max_cells = (int64_t)1 << 20;
    bool lds_2d = true;                 // 2D small grids: state in LDS for a check interval (k2_lds)
    int lds_rows_2d = 0;                // rows per block of that launch (0: the default choice)
    int lds_cols_2d = 0;                // columns per block of that launch (0: the whole width)
    int lds_threads_2d = 1024;          // threads per block of that launch
    int lds_halo_2d = 0;                // halo rows (1 to 4; 0: the default choice): rows exchanged every that many iterations
                                        // (127^2: 5.4 us per iteration with 2, 6.2 with 1)
    double* d_lds_xb = nullptr;         // its exchange rows
    unsigned long long* d_lds_flags = nullptr;
    int lds_blocks = 0;                 // blocks the counters are sized for
    size_t lds_xb_bytes = 0;            // size of the exchange buffer
    unsigned long long lds_gen = 0;     // the counters' value between launches
    int lds_nb = 0;                     // blocks of the last launch (whose counters all hold lds_gen)
    bool lds_fault_wait = false;        // test hook: the exchange waits target a generation never reached
    int lds_wait_spins = 1 << 22;       // polls before an exchange wait gives up
    unsigned* d_bar = nullptr;          // barrier count, generation, error flag
    int* h_bar_err = nullptr;           // pinned copy of the error flag
    int planes_per_block = -160; // 3D single-pass: z planes per block (<0: fewest block rounds, capped)
    int dc_planes_per_block = -24; // the same for the 3D FP32 correction sweep
    // 3D FP32 sweep (-1: when its arrays exceed the 256 MB Infinity Cache, 0 off, 1 on):
    int dc_dp = -1;              // Gdτ, dτ_Rho from Musτ when they are derived from it
    int dc_map_xcd = -1;         // each XCD sweeps its eighth of the tiles chunk by chunk
    int tile_rows_3d = 12;      // 3D single-pass: y rows per block (8, 12 or 16)
    bool two_kernel_3d = false; // 3D: in-place kernel pair instead of the single pass
    int variant_3d = 1;         // 3D single-pass kernel: 0 k3_iter, 1 k3_iter2, 2 k3_iter2 with prefetch
    bool group_3d = true;       // 3D: state fields, partners and τ_o in grouped allocations (ensure_group)
    bool map_flat_3d = false;   // k3_iter2: tiles in the flat (chunk, tile) order
    bool w64_3d = true;         // k3_iter2: per-slot pointers also where 32-bit slot offsets fit
    bool stream_3d = true;      // k3_iter2 (per-slot pointers, Musτ from μ): streaming loads of the lines only one tile reads
    // placement selection (place_fields): tries (3D, 2D), the least bytes of
    // field blocks for which it runs, a forced outcome for tests (-1: by time)
    int place_tries_3d = 31;
    int place_tries_2d = 31;
    int64_t place_budget_iters = 5000;  // stop trying after this many iterations' worth of time
    bool group_2d = false;      // 2D: state fields, partners and τ_o in grouped allocations
    int64_t place_min_bytes = (int64_t)4 << 30;
    int place_accept = -1;
    bool placed = false;
    float place_ms[32] = {};    // iteration time on each placement tried (0: not tried)
    // adjoint state λ (ping-pong by field id), the pass-1 scratch, the
    // source ∂J/∂Vx, ∂J/∂Vy and the gradient buffer
    double* adj[2][PTS_NFIELDS] = {};
    double* adj_tmp[PTS_NFIELDS] = {};
    double* adj_src[PTS_NFIELDS] = {};  // ∂J/∂x under VX, VY; step coupling under TXX, TYY, TXY
    double* adj_grad = nullptr;
    int adj_cur = 0;
    bool adj_fused = true;  // single-pass adjoint sweep (PTS_ADJ_TWO_PASS=1: the two passes)
    std::vector<std::array<double*, 3>> ckpt;  // converged τxx, τyy, τxy of each physical step
    // Gdτ and dτ_Rho as produced by the last pts_iter_params from the current
    // Musτ: the single-pass kernels may then derive them from Musτ (one array
    // read instead of two).  Any write to Musτ, Gdτ or dτ_Rho clears dp_ok.
    bool dp_ok = false;
    bool dp_enabled = true;
    StepScalars dp{};
    // 2D: Musτ is maxloc(μ) with the free-slip copies (checked on the device:
    // 0 unknown, 1 yes, -1 no), so the kernels may derive it from μ as well.
    // Any write to μ or Musτ resets the state to unknown.
    int mx_state = 0;
    bool mx_enabled = true;
    int* d_flag = nullptr;
    // field allocations: the pointer handed to kernels may sit `stagger`
    // bytes past the start of its hipMalloc block (see field_alloc)
    std::map<void*, void*> alloc_base;
    std::map<void*, size_t> alloc_len;  // hipMalloc'ed bytes of each block, by handed pointer
    std::set<void*> slot_ptrs;  // slots 1, 2 of grouped 3D fields (freed with slot 0)
    int64_t slot_bytes = 0;     // distance of the slots of a grouped field (bytes)
    int64_t stagger = 4352;
    int nallocs = 0;
    std::string err;
    // graphs of n single-pass iterations keyed by (n, parity of the start buffer)
    struct CachedGraph {
        hipGraphExec_t exec;
        Fields in, out;  // buffers and scalars the graph was captured with
        StepScalars s;
        bool dp, mx;
    };
    std::map<std::pair<int64_t, int>, CachedGraph> graphs;
    int parity = 0;  // flips each single-pass iteration (which physical buffer is `cur`)
};

namespace {

int fail(pts_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}

#define HIPCHK(ctx, call)                                                                             \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            return fail((ctx), PTS_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));       \
    } while (0)

// Elements of the zeroed guard zone before and after each 3D field array:
// k3_iter2 loads every field at its fixed shift without lane or z-range
// predicates, so the halo lanes and rows of the edge tiles (a row's worth of
// columns and up to 32 rows beyond the box) and the planes before the first
// and after the last read inside the allocation.  Those values only feed
// results that are not stored.
int64_t field_guard_elems(const pts_ctx* c) {
    return c->nd == 3 ? (int64_t)c->sxy + 32 * (int64_t)c->sx + 256 : 0;
}

// FP64 fields only: the FP32 correction state is not read by k3_iter2.
size_t field_guard_bytes(const pts_ctx* c, size_t elem) {
    if (elem != sizeof(double)) return 0;
    return ((size_t)field_guard_elems(c) * elem + 255) / 256 * 256;
}

// Distance of the slots of a grouped field: the array and one guard zone,
// plus one stagger step so that the slots do not align with each other.
size_t slot_stride(const pts_ctx* c, size_t elem) {
    return ((size_t)c->nalloc * elem + field_guard_bytes(c, elem) + 255) / 256 * 256 + (size_t)c->stagger;
}

// One field array of nalloc elements of `elem` bytes, zeroed (with `slots`
// > 1: that many arrays, slot_stride apart, in one allocation).  With a
// stagger of S bytes (a multiple of 256), the n-th array starts (n mod 16)·S
// bytes into its block, so equal offsets of different arrays do not all fall
// on the same HBM channel when the array size is a power of two.
int field_alloc(pts_ctx* c, size_t elem, void** out, int slots = 1) {
    const size_t bytes = (size_t)c->nalloc * elem;
    const size_t guard = field_guard_bytes(c, elem);
    const size_t off = guard + (size_t)((c->nallocs % 16) * c->stagger);
    const size_t total = bytes + 16 * (size_t)c->stagger + 2 * guard + (size_t)(slots - 1) * slot_stride(c, elem);
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, total);
    if (e != hipSuccess) return fail(c, PTS_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    void* q = (char*)p + off;
    HIPCHK(c, hipMemsetAsync(p, 0, total, c->stream));
    c->alloc_base[q] = p;
    c->alloc_len[q] = total;
    c->nallocs += 1;
    c->bytes += (int64_t)total;
    *out = q;
    return PTS_OK;
}

void field_free(pts_ctx* c, void* q) {
    if (!q || c->slot_ptrs.count(q)) return;
    auto it = c->alloc_base.find(q);
    (void)hipFree(it != c->alloc_base.end() ? it->second : q);
    if (it != c->alloc_base.end()) c->alloc_base.erase(it);
    c->alloc_len.erase(q);
}

bool field_valid(const pts_ctx* c, int f) {
    if (f < 0 || f >= PTS_NFIELDS) return false;
    if (c->nd == 2) {
        switch (f) {
            case PTS_VZ: case PTS_TZZ: case PTS_TXZ: case PTS_TYZ: case PTS_TZZ_O: case PTS_TXZ_O:
            case PTS_TYZ_O: case PTS_RZ: case PTS_DVZ:
                return false;
            default:
                return true;
        }
    }
    return true;
}

// 3D state fields are allocated as groups: slot 0 the field, slot 1 its
// ping-pong partner and, for the stresses of a visco-elastic context, slot 2
// the old-time stress τ_o, so that the single-pass kernel addresses all of
// them from one base pointer (k3_iter2).
int ensure_group(pts_ctx* c, int f) {
    const bool stress = f >= PTS_TXX && f <= PTS_TYZ;
    const int slots = (stress && c->g.ve) ? 3 : 2;
    void* p = nullptr;
    if (int rc = field_alloc(c, sizeof(double), &p, slots)) return rc;
    const size_t sb = slot_stride(c, sizeof(double));
    c->slot_bytes = (int64_t)sb;
    char* q = (char*)p;
    c->cur[f] = (double*)q;
    c->alt[f] = (double*)(q + sb);
    c->slot_ptrs.insert(c->alt[f]);
    c->alt_synced = false;
    if (slots == 3) {
        c->cur[f + (PTS_TXX_O - PTS_TXX)] = (double*)(q + 2 * sb);
        c->slot_ptrs.insert(q + 2 * sb);
    }
    return PTS_OK;
}

int ensure(pts_ctx* c, int f) {
    if (!field_valid(c, f)) return fail(c, PTS_ERR_STATE, "field " + std::to_string(f) + " not valid for this grid");
    if (c->cur[f]) return PTS_OK;
    if ((c->nd == 3 && c->group_3d) || (c->nd == 2 && c->group_2d)) {
        if (f >= PTS_PT && f <= PTS_TYZ) return ensure_group(c, f);
        if (f >= PTS_TXX_O && f <= PTS_TYZ_O && c->g.ve) {
            const int s = f - (PTS_TXX_O - PTS_TXX);
            if (!c->cur[s]) return ensure_group(c, s);
        }
    }
    void* p = nullptr;
    if (int rc = field_alloc(c, sizeof(double), &p)) return rc;
    c->cur[f] = (double*)p;
    return PTS_OK;
}

int ensure_list(pts_ctx* c, std::initializer_list<int> fs) {
    for (int f : fs) {
        int rc = ensure(c, f);
        if (rc) return rc;
    }
    return PTS_OK;
}

int ensure_alt(pts_ctx* c, int f) {
    if (c->alt[f]) return PTS_OK;
    void* p = nullptr;
    if (int rc = field_alloc(c, sizeof(double), &p)) return rc;
    c->alt[f] = (double*)p;
    c->alt_synced = false;
    return PTS_OK;
}

Fields fields_of(const pts_ctx* c, double* const* src) {
    Fields f;
    for (int q = 0; q < PTS_NFIELDS; ++q) f.p[q] = src[q];
    f.sx = c->sx;
    f.sxy = c->sxy;
    f.nx = c->nx;
    f.ny = c->ny;
    f.nz = c->nz;
    return f;
}
Fields F(const pts_ctx* c) { return fields_of(c, c->cur); }

// Any in-place change of the velocities invalidates the Dirichlet copy
// held by the ping-pong partner buffers.
void touched(pts_ctx* c) { c->alt_synced = false; }

// A write to field f (-1: any field) through the API.
void wrote(pts_ctx* c, int f) {
    if (f < 0 || f == PTS_MUSTAU || f == PTS_GDTAU || f == PTS_DTAU_RHO) c->dp_ok = false;
    if (f < 0 || f == PTS_MUSTAU || f == PTS_MUS) c->mx_state = 0;
}

int setdev(pts_ctx* c) {
    HIPCHK(c, hipSetDevice(c->g.device));
    (void)hipGetLastError();  // launch checks below report only their own errors
    return PTS_OK;
}

StepScalars scal(double r, double G, double dt, double dx, double dy, double dz) {
    StepScalars s{};
    s.r = r;
    s.Gdt = G * dt;
    s.dx = dx; s.dy = dy; s.dz = dz;
    s._dx = 1.0 / dx; s._dy = 1.0 / dy; s._dz = 1.0 / dz;
    return s;
}

int state_fields(const pts_ctx* c, int* out) {
    int n = 0;
    for (int f : kState)
        if (field_valid(c, f)) out[n++] = f;
    return n;
}

}  // namespace

namespace pts {
void set_global_error(const std::string& msg) { g_create_error = msg; }

void field_shape(int ndim, int nx, int ny, int nz, int f, int64_t s[3]) {
    s[0] = nx; s[1] = ny; s[2] = ndim == 3 ? nz : 1;
    if (ndim == 2) {
        switch (f) {
            case PTS_VX: s[0] = nx + 1; break;
            case PTS_VY: s[1] = ny + 1; break;
            case PTS_TXY: case PTS_TXY_O: s[0] = nx - 1; s[1] = ny - 1; break;
            case PTS_RX: case PTS_DVX: s[0] = nx - 1; s[1] = ny - 2; break;
            case PTS_RY: case PTS_DVY: s[0] = nx - 2; s[1] = ny - 1; break;
            default: break;
        }
        return;
    }
    switch (f) {
        case PTS_VX: s[0] = nx + 1; break;
        case PTS_VY: s[1] = ny + 1; break;
        case PTS_VZ: s[2] = nz + 1; break;
        case PTS_TXX: case PTS_TXX_O: s[1] = ny - 2; s[2] = nz - 2; break;
        case PTS_TYY: case PTS_TYY_O: s[0] = nx - 2; s[2] = nz - 2; break;
        case PTS_TZZ: case PTS_TZZ_O: s[0] = nx - 2; s[1] = ny - 2; break;
        case PTS_TXY: case PTS_TXY_O: s[0] = nx - 1; s[1] = ny - 1; s[2] = nz - 2; break;
        case PTS_TXZ: case PTS_TXZ_O: s[0] = nx - 1; s[1] = ny - 2; s[2] = nz - 1; break;
        case PTS_TYZ: case PTS_TYZ_O: s[0] = nx - 2; s[1] = ny - 1; s[2] = nz - 1; break;
        case PTS_RX: case PTS_DVX: s[0] = nx - 1; s[1] = ny - 2; s[2] = nz - 2; break;
        case PTS_RY: case PTS_DVY: s[0] = nx - 2; s[1] = ny - 1; s[2] = nz - 2; break;
        case PTS_RZ: case PTS_DVZ: s[0] = nx - 2; s[1] = ny - 2; s[2] = nz - 1; break;
        default: break;
    }
}
}  // namespace pts

extern "C" {

int pts_create(const pts_grid* g, pts_ctx** out) {
    if (!g || !out) return fail(nullptr, PTS_ERR_ARG, "null argument");
    *out = nullptr;
    if (g->ndim != 2 && g->ndim != 3) return fail(nullptr, PTS_ERR_ARG, "ndim must be 2 or 3");
    const int nz = g->ndim == 3 ? g->nz : 1;
    if (g->nx < 3 || g->ny < 3 || (g->ndim == 3 && g->nz < 3))
        return fail(nullptr, PTS_ERR_ARG, "each of nx, ny (, nz) must be >= 3");
    if (g->nx > 65000 || g->ny > 65000 || nz > 65000) return fail(nullptr, PTS_ERR_ARG, "grid dimension too large");
    if (g->fd_mode != PTS_FD_DIV && g->fd_mode != PTS_FD_MUL) return fail(nullptr, PTS_ERR_ARG, "bad fd_mode");
    pts_ctx* c = new pts_ctx();
    c->g = *g;
    c->nd = g->ndim;
    c->nx = g->nx;
    c->ny = g->ny;
    c->nz = nz;
    c->sx = ((g->nx + 1) + 31) / 32 * 32;
    int64_t sxy = (int64_t)c->sx * (g->ny + 1);
    // 3D: optional padding of the plane stride (elements, multiple of 32)
    if (g->ndim == 3)
        if (const char* e = std::getenv("PTS_PLANE_PAD")) sxy += std::max(0, std::atoi(e)) / 32 * 32;
    const int64_t na = sxy * (g->ndim == 3 ? nz + 1 : 1);
    if (na >= (int64_t)1 << 31) {
        delete c;
        return fail(nullptr, PTS_ERR_ARG, "grid too large for 32-bit element offsets");
    }
    c->sxy = (int)sxy;
    c->nalloc = na;
    if (const char* e = std::getenv("PTS_ROWS_PER_BLOCK")) c->rows_per_block = std::atoi(e);
    if (const char* e = std::getenv("PTS_2D_VARIANT")) c->variant_2d = std::atoi(e);
    if (const char* e = std::getenv("PTS_2D_PERSIST")) c->persist_2d = std::atoi(e) != 0;
    if (const char* e = std::getenv("PTS_2D_MULTI")) c->multi_2d = std::atoi(e) != 0;
    if (const char* e = std::getenv("PTS_2D_LDS")) c->lds_2d = std::atoi(e) != 0;
    if (const char* e = std::getenv("PTS_2D_LDS_ROWS")) c->lds_rows_2d = std::max(0, std::atoi(e));
    if (const char* e = std::getenv("PTS_2D_LDS_COLS")) c->lds_cols_2d = std::max(0, std::atoi(e));
    if (const char* e = std::getenv("PTS_2D_LDS_FAULT_WAIT")) c->lds_fault_wait = std::atoi(e) != 0;
    if (const char* e = std::getenv("PTS_2D_LDS_WAIT_SPINS")) c->lds_wait_spins = std::max(1, std::atoi(e));
    if (const char* e = std::getenv("PTS_2D_LDS_THREADS")) c->lds_threads_2d = std::min(1024, std::max(64, std::atoi(e) / 64 * 64));
    if (const char* e = std::getenv("PTS_2D_LDS_HALO")) c->lds_halo_2d = std::min(4, std::max(0, std::atoi(e)));
    if (const char* e = std::getenv("PTS_2D_MULTI_ROWS")) c->multi_rows_2d = std::max(1, std::atoi(e));
    if (const char* e = std::getenv("PTS_2D_MULTI_MAX_CELLS")) c->multi_max_cells = std::atoll(e);
    if (const char* e = std::getenv("PTS_DC_PLANES_PER_BLOCK")) c->dc_planes_per_block = std::atoi(e);
    if (const char* e = std::getenv("PTS_DC_DP")) c->dc_dp = std::max(-1, std::min(1, std::atoi(e)));
    if (const char* e = std::getenv("PTS_DC_MAP_XCD")) c->dc_map_xcd = std::max(-1, std::min(1, std::atoi(e)));
    if (const char* e = std::getenv("PTS_PLANES_PER_BLOCK")) c->planes_per_block = std::atoi(e);
    if (const char* e = std::getenv("PTS_3D_TWO_KERNEL")) c->two_kernel_3d = std::atoi(e) != 0;
    if (const char* e = std::getenv("PTS_3D_TILE_ROWS")) c->tile_rows_3d = std::atoi(e);
    if (const char* e = std::getenv("PTS_3D_VARIANT")) c->variant_3d = std::atoi(e);
    if (const char* e = std::getenv("PTS_3D_GROUPS")) c->group_3d = std::atoi(e) != 0;
    if (const char* e = std::getenv("PTS_3D_MAP_FLAT")) c->map_flat_3d = std::atoi(e) != 0;
    if (const char* e = std::getenv("PTS_3D_W64")) c->w64_3d = std::atoi(e) != 0;
    if (const char* e = std::getenv("PTS_3D_STREAM")) c->stream_3d = std::atoi(e) != 0;
    if (const char* e = std::getenv("PTS_3D_PLACE_TRIES")) c->place_tries_3d = std::min(31, std::max(0, std::atoi(e)));
    if (const char* e = std::getenv("PTS_2D_PLACE_TRIES")) c->place_tries_2d = std::min(31, std::max(0, std::atoi(e)));
    if (const char* e = std::getenv("PTS_2D_GROUPS")) c->group_2d = std::atoi(e) != 0;
    if (const char* e = std::getenv("PTS_PLACE_MIN_MB")) c->place_min_bytes = std::atoll(e) << 20;
    if (const char* e = std::getenv("PTS_PLACE_ACCEPT")) c->place_accept = std::atoi(e) != 0 ? 1 : 0;
    if (const char* e = std::getenv("PTS_PLACE_BUDGET_ITERS")) c->place_budget_iters = std::max<int64_t>(0, std::atoll(e));
    if (const char* e = std::getenv("PTS_DERIVED_PARAMS")) c->dp_enabled = std::atoi(e) != 0;
    if (const char* e = std::getenv("PTS_DERIVED_MUSTAU")) c->mx_enabled = std::atoi(e) != 0;
    if (const char* e = std::getenv("PTS_ADJ_TWO_PASS")) c->adj_fused = std::atoi(e) == 0;
    if (const char* e = std::getenv("PTS_STAGGER_BYTES")) c->stagger = std::max<int64_t>(0, std::atoll(e)) / 256 * 256;
    hipError_t e = hipSetDevice(g->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void**)&c->red_partials, sizeof(double) * kRedBlocks * kRedVals);
    if (e == hipSuccess) e = hipMalloc((void**)&c->red_out, sizeof(double) * kRedVals);
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->h_red, sizeof(double) * kRedVals, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc((void**)&c->d_flag, sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void**)&c->d_bar, 4 * sizeof(unsigned));
    if (e == hipSuccess) e = hipMemset(c->d_bar, 0, 4 * sizeof(unsigned));
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->h_bar_err, sizeof(int), hipHostMallocDefault);
    if (e == hipSuccess) *c->h_bar_err = 0;
    if (e != hipSuccess) {
        std::string m = std::string("HIP init: ") + hipGetErrorString(e);
        pts_destroy(c);
        return fail(nullptr, PTS_ERR_HIP, m);
    }
    *out = c;
    return PTS_OK;
}

int pts_destroy(pts_ctx* c) {
    if (!c) return PTS_OK;
    (void)hipSetDevice(c->g.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second.exec);
    for (int q = 0; q < PTS_NFIELDS; ++q) {
        field_free(c, c->cur[q]);
        field_free(c, c->alt[q]);
        field_free(c, c->cur32[q]);
        if (c->alt32[q] != c->cur32[q]) field_free(c, c->alt32[q]);
        field_free(c, c->adj[0][q]);
        field_free(c, c->adj[1][q]);
        field_free(c, c->adj_tmp[q]);
    }
    for (double* p : c->adj_src) field_free(c, p);
    field_free(c, c->adj_grad);
    for (auto& slot : c->ckpt)
        for (double* p : slot) field_free(c, p);
    if (c->red_partials) (void)hipFree(c->red_partials);
    if (c->red_out) (void)hipFree(c->red_out);
    if (c->d_flag) (void)hipFree(c->d_flag);
    if (c->d_lds_xb) (void)hipFree(c->d_lds_xb);
    if (c->d_lds_flags) (void)hipFree(c->d_lds_flags);
    if (c->d_bar) (void)hipFree(c->d_bar);
    if (c->h_bar_err) (void)hipHostFree(c->h_bar_err);
    if (c->h_red) (void)hipHostFree(c->h_red);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return PTS_OK;
}

const char* pts_last_error(const pts_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int pts_field_shape(const pts_ctx* c, int f, int64_t shape[3]) {
    if (!c || !shape) return PTS_ERR_ARG;
    if (!field_valid(c, f)) return fail(const_cast<pts_ctx*>(c), PTS_ERR_STATE, "field not valid for this grid");
    field_shape(c->nd, c->nx, c->ny, c->nz, f, shape);
    return PTS_OK;
}

int pts_get_stream(const pts_ctx* c, void** stream) {
    if (!c || !stream) return PTS_ERR_ARG;
    *stream = (void*)c->stream;
    return PTS_OK;
}

int pts_device_bytes(const pts_ctx* c, int64_t* bytes) {
    if (!c || !bytes) return PTS_ERR_ARG;
    *bytes = c->bytes;
    return PTS_OK;
}

// Dense host array (x fastest) <-> the strided device layout of field f at `base`.
static int copy_dense(pts_ctx* c, int f, double* base, double* host, const double* chost, int64_t count, bool up) {
    int64_t s[3];
    field_shape(c->nd, c->nx, c->ny, c->nz, f, s);
    if (count != s[0] * s[1] * s[2])
        return fail(c, PTS_ERR_ARG, "count " + std::to_string(count) + " != field size " + std::to_string(s[0] * s[1] * s[2]));
    if (count == 0) return PTS_OK;
    const size_t w = (size_t)s[0] * sizeof(double);
    for (int64_t k = 0; k < s[2]; ++k) {
        double* dev = base + k * (int64_t)c->sxy;
        if (up)
            HIPCHK(c, hipMemcpy2DAsync(dev, (size_t)c->sx * sizeof(double), chost + k * s[0] * s[1], w, w, (size_t)s[1],
                                       hipMemcpyHostToDevice, c->stream));
        else
            HIPCHK(c, hipMemcpy2DAsync(host + k * s[0] * s[1], w, dev, (size_t)c->sx * sizeof(double), w, (size_t)s[1],
                                       hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PTS_OK;
}

static int copy_field(pts_ctx* c, int f, double* host, const double* chost, int64_t count, bool up) {
    if (!c || (up ? !chost : !host)) return fail(c, PTS_ERR_ARG, "null argument");
    if (int rc = setdev(c)) return rc;
    if (int rc = ensure(c, f)) return rc;
    if (int rc = copy_dense(c, f, c->cur[f], host, chost, count, up)) return rc;
    if (up) {
        touched(c);
        wrote(c, f);
    }
    return PTS_OK;
}

int pts_upload(pts_ctx* c, int f, const double* host, int64_t count) { return copy_field(c, f, nullptr, host, count, true); }
int pts_download(pts_ctx* c, int f, double* host, int64_t count) { return copy_field(c, f, host, nullptr, count, false); }

int pts_fill(pts_ctx* c, int f, double v) {
    if (!c) return PTS_ERR_ARG;
    if (int rc = setdev(c)) return rc;
    if (int rc = ensure(c, f)) return rc;
    int64_t s[3];
    field_shape(c->nd, c->nx, c->ny, c->nz, f, s);
    HIPCHK(c, launch_fill(c->stream, c->cur[f], v, F(c), s));
    touched(c);
    wrote(c, f);
    return PTS_OK;
}

int pts_copy(pts_ctx* c, int dst, int src) {
    if (!c) return PTS_ERR_ARG;
    if (int rc = setdev(c)) return rc;
    if (int rc = ensure_list(c, {dst, src})) return rc;
    int64_t a[3], b[3];
    field_shape(c->nd, c->nx, c->ny, c->nz, dst, a);
    field_shape(c->nd, c->nx, c->ny, c->nz, src, b);
    if (a[0] != b[0] || a[1] != b[1] || a[2] != b[2]) return fail(c, PTS_ERR_ARG, "copy: shape mismatch");
    HIPCHK(c, launch_copy(c->stream, c->cur[dst], c->cur[src], F(c), a));
    touched(c);
    wrote(c, dst);
    return PTS_OK;
}

int pts_swap(pts_ctx* c, int a, int b) {
    if (!c) return PTS_ERR_ARG;
    if (int rc = ensure_list(c, {a, b})) return rc;
    int64_t x[3], y[3];
    field_shape(c->nd, c->nx, c->ny, c->nz, a, x);
    field_shape(c->nd, c->nx, c->ny, c->nz, b, y);
    if (x[0] != y[0] || x[1] != y[1] || x[2] != y[2]) return fail(c, PTS_ERR_ARG, "swap: shape mismatch");
    std::swap(c->cur[a], c->cur[b]);
    std::swap(c->alt[a], c->alt[b]);
    touched(c);
    wrote(c, a);
    wrote(c, b);
    return PTS_OK;
}

int pts_sync(pts_ctx* c) {
    if (!c) return PTS_ERR_ARG;
    if (int rc = setdev(c)) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PTS_OK;
}

// ---- setup ------------------------------------------------------------------
int pts_init_inclusion(pts_ctx* c, double lx, double ly, double lz, double mus0, double musi, double rad2, double ebg,
                       double rin, double rout, int32_t igg) {
    if (!c) return PTS_ERR_ARG;
    if (int rc = setdev(c)) return rc;
    int st[16];
    const int ns = state_fields(c, st);
    for (int q = 0; q < ns; ++q)
        if (int rc = ensure(c, st[q])) return rc;
    if (int rc = ensure(c, PTS_MUS)) return rc;
    if (c->g.body_force)
        if (int rc = ensure(c, PTS_RHOG)) return rc;
    if (c->g.ve) {
        for (int f : {PTS_TXX_O, PTS_TYY_O, PTS_TXY_O})
            if (int rc = ensure(c, f)) return rc;
        if (c->nd == 3)
            for (int f : {PTS_TZZ_O, PTS_TXZ_O, PTS_TYZ_O})
                if (int rc = ensure(c, f)) return rc;
    }
    // zero all state and stress-history arrays (the scripts' @zeros)
    for (int f = 0; f < PTS_NFIELDS; ++f)
        if (c->cur[f]) HIPCHK(c, hipMemsetAsync(c->cur[f], 0, (size_t)c->nalloc * sizeof(double), c->stream));
    HIPCHK(c, launch_init_inclusion(c->stream, F(c), c->nd, lx, ly, lz, mus0, musi, rad2, ebg, rin, rout, igg));
    touched(c);
    wrote(c, -1);
    return PTS_OK;
}

int pts_init_grf_viscosity(pts_ctx* c, int32_t nmodes, const double* modes, double sigma, double lo, double hi,
                           double lx, double ly, double lz, double musi, double rad2) {
    if (!c || !modes || nmodes < 1) return fail(c, PTS_ERR_ARG, "bad argument");
    if (int rc = setdev(c)) return rc;
    if (int rc = ensure(c, PTS_MUS)) return rc;
    double* dm = nullptr;
    const size_t bytes = sizeof(double) * 4 * (size_t)nmodes;
    HIPCHK(c, hipMalloc((void**)&dm, bytes));
    hipError_t e = hipMemcpyAsync(dm, modes, bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = launch_grf(c->stream, F(c), c->nd, dm, nmodes, sigma, lo, hi, lx, ly, lz, musi, rad2);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dm);
    if (e != hipSuccess) return fail(c, PTS_ERR_HIP, std::string("grf: ") + hipGetErrorString(e));
    wrote(c, PTS_MUS);
    return PTS_OK;
}

int pts_smooth(pts_ctx* c, int dst, int src, double fact) {
    if (!c) return PTS_ERR_ARG;
    if (int rc = setdev(c)) return rc;
    if (int rc = ensure_list(c, {dst, src})) return rc;
    int64_t s[3];
    field_shape(c->nd, c->nx, c->ny, c->nz, src, s);
    HIPCHK(c, launch_smooth(c->stream, F(c), c->nd, c->cur[dst], c->cur[src], s, fact));
    touched(c);
    wrote(c, dst);
    return PTS_OK;
}

int pts_maxloc(pts_ctx* c, int dst, int src) {
    if (!c) return PTS_ERR_ARG;
    if (int rc = setdev(c)) return rc;
    if (int rc = ensure_list(c, {dst, src})) return rc;
    int64_t s[3];
    field_shape(c->nd, c->nx, c->ny, c->nz, src, s);
    HIPCHK(c, launch_maxloc(c->stream, F(c), c->nd, c->cur[dst], c->cur[src], s));
    touched(c);
    wrote(c, dst);
    return PTS_OK;
}

static int bc_dim(pts_ctx* c, int f, int dim) {
    if (!c) return PTS_ERR_ARG;
    if (dim == 2 && c->nd != 3) return fail(c, PTS_ERR_STATE, "bc_z on a 2D grid");
    if (int rc = setdev(c)) return rc;
    if (int rc = ensure(c, f)) return rc;
    int64_t s[3];
    field_shape(c->nd, c->nx, c->ny, c->nz, f, s);
    HIPCHK(c, launch_bc(c->stream, F(c), c->cur[f], s, dim));
    touched(c);
    wrote(c, f);
    return PTS_OK;
}
int pts_bc_x(pts_ctx* c, int f) { return bc_dim(c, f, 0); }
int pts_bc_y(pts_ctx* c, int f) { return bc_dim(c, f, 1); }
int pts_bc_z(pts_ctx* c, int f) { return bc_dim(c, f, 2); }

int pts_iter_params(pts_ctx* c, double Vp, double G, double dt, double Re, double r, double ml) {
    if (!c) return PTS_ERR_ARG;
    if (int rc = setdev(c)) return rc;
    if (int rc = ensure_list(c, {PTS_MUSTAU, PTS_GDTAU, PTS_DTAU_RHO})) return rc;
    HIPCHK(c, launch_iter_params(c->stream, F(c), c->nd, c->g.ve != 0, Vp, G, dt, Re, r, ml));
    // the uniform sub-expressions of k_iter_params, evaluated the same way
    c->dp.dp_ve = c->g.ve ? 1 : 0;
    c->dp.ip_igdt = 1.0 / (G * dt);
    c->dp.ip_k1 = Vp * ml / Re;
    c->dp.ip_k2 = Vp * Vp;
    c->dp.ip_r2 = r + 2.0;
    c->dp_ok = true;
    return PTS_OK;
}

int pts_assign_tau(pts_ctx* c) {
    if (!c) return PTS_ERR_ARG;
    if (int rc = setdev(c)) return rc;
    const int src2[] = {PTS_TXX, PTS_TYY, PTS_TXY}, dst2[] = {PTS_TXX_O, PTS_TYY_O, PTS_TXY_O};
    const int src3[] = {PTS_TXX, PTS_TYY, PTS_TZZ, PTS_TXY, PTS_TXZ, PTS_TYZ};
    const int dst3[] = {PTS_TXX_O, PTS_TYY_O, PTS_TZZ_O, PTS_TXY_O, PTS_TXZ_O, PTS_TYZ_O};
    const int n = c->nd == 2 ? 3 : 6;
    for (int q = 0; q < n; ++q) {
        const int s = c->nd == 2 ? src2[q] : src3[q], d = c->nd == 2 ? dst2[q] : dst3[q];
        if (int rc = ensure_list(c, {s, d})) return rc;
        int64_t sh[3];
        field_shape(c->nd, c->nx, c->ny, c->nz, s, sh);
        HIPCHK(c, launch_copy(c->stream, c->cur[d], c->cur[s], F(c), sh));
    }
    return PTS_OK;
}

// ---- per-kernel reference API --------------------------------------------
static int need_iter_fields(pts_ctx* c, bool with_o) {
    int st[16];
    const int ns = state_fields(c, st);
    for (int q = 0; q < ns; ++q)
        if (int rc = ensure(c, st[q])) return rc;
    if (int rc = ensure_list(c, {PTS_MUS, PTS_GDTAU, PTS_DTAU_RHO})) return rc;
    if (c->g.body_force)
        if (int rc = ensure(c, PTS_RHOG)) return rc;
    if (with_o && c->g.ve) {
        for (int f : {PTS_TXX_O, PTS_TYY_O, PTS_TXY_O})
            if (int rc = ensure(c, f)) return rc;
        if (c->nd == 3)
            for (int f : {PTS_TZZ_O, PTS_TXZ_O, PTS_TYZ_O})
                if (int rc = ensure(c, f)) return rc;
    }
    return PTS_OK;
}

int pts_compute_P(pts_ctx* c, double r, double dx, double dy, double dz) {
    if (!c) return PTS_ERR_ARG;
    if (int rc = setdev(c)) return rc;
    if (int rc = need_iter_fields(c, false)) return rc;
    if (int rc = ensure(c, PTS_DIVV)) return rc;
    StepScalars s = scal(r, 1.0, 1.0, dx, dy, dz);
    HIPCHK(c, c->nd == 2 ? launch2_compute_P(c->stream, F(c), s) : launch3_compute_P(c->stream, F(c), s));
    return PTS_OK;
}

int pts_compute_tau(pts_ctx* c, double G, double dt, double dx, double dy, double dz) {
    if (!c) return PTS_ERR_ARG;
    if (int rc = setdev(c)) return rc;
    if (int rc = need_iter_fields(c, true)) return rc;
    StepScalars s = scal(0.0, G, dt, dx, dy, dz);
    const bool ve = c->g.ve != 0;
    HIPCHK(c, c->nd == 2 ? launch2_compute_tau(c->stream, F(c), s, ve) : launch3_compute_tau(c->stream, F(c), s, ve));
    return PTS_OK;
}

int pts_compute_dV(pts_ctx* c, double dx, double dy, double dz) {
    if (!c) return PTS_ERR_ARG;
    if (int rc = setdev(c)) return rc;
    if (int rc = need_iter_fields(c, false)) return rc;
    if (int rc = ensure_list(c, {PTS_DVX, PTS_DVY})) return rc;
    if (c->nd == 2) {
        if (int rc = ensure_list(c, {PTS_RX, PTS_RY})) return rc;
    } else {
        if (int rc = ensure(c, PTS_DVZ)) return rc;
    }
    StepScalars s = scal(0.0, 1.0, 1.0, dx, dy, dz);
    const bool bf = c->g.body_force != 0;
    HIPCHK(c, c->nd == 2 ? launch2_compute_dV(c->stream, F(c), s, bf) : launch3_compute_dV(c->stream, F(c), s, bf));
    return PTS_OK;
}

int pts_compute_V(pts_ctx* c) {
    if (!c) return PTS_ERR_ARG;
    if (int rc = setdev(c)) return rc;
    if (int rc = need_iter_fields(c, false)) return rc;
    if (int rc = ensure_list(c, {PTS_DVX, PTS_DVY})) return rc;
    if (c->nd == 3)
        if (int rc = ensure(c, PTS_DVZ)) return rc;
    HIPCHK(c, c->nd == 2 ? launch2_compute_V(c->stream, F(c)) : launch3_compute_V(c->stream, F(c)));
    touched(c);
    return PTS_OK;
}

static int res_common(pts_ctx* c, double dx, double dy, double dz, bool mul) {
    if (!c) return PTS_ERR_ARG;
    if (int rc = setdev(c)) return rc;
    if (int rc = need_iter_fields(c, false)) return rc;
    if (int rc = ensure_list(c, {PTS_RX, PTS_RY, PTS_DIVV})) return rc;
    if (c->nd == 3)
        if (int rc = ensure(c, PTS_RZ)) return rc;
    StepScalars s = mul ? scal(0.0, 1.0, 1.0, 1.0 / dx, 1.0 / dy, 1.0 / dz) : scal(0.0, 1.0, 1.0, dx, dy, dz);
    if (mul) {  // callers pass _dx; keep the given reciprocal exactly
        s._dx = dx; s._dy = dy; s._dz = dz;
    }
    const bool bf = c->g.body_force != 0;
    HIPCHK(c, c->nd == 2 ? launch2_compute_Res(c->stream, F(c), s, bf, mul)
                         : launch3_compute_Res(c->stream, F(c), s, bf, mul));
    return PTS_OK;
}
int pts_compute_Res(pts_ctx* c, double dx, double dy, double dz) { return res_common(c, dx, dy, dz, false); }
int pts_compute_Res_perf(pts_ctx* c, double _dx, double _dy, double _dz) { return res_common(c, _dx, _dy, _dz, true); }

int pts_compute_Pt_tau(pts_ctx* c, double r, double G, double dt, double _dx, double _dy, double _dz) {
    if (!c) return PTS_ERR_ARG;
    if (int rc = setdev(c)) return rc;
    if (int rc = need_iter_fields(c, true)) return rc;
    StepScalars s = scal(r, G, dt, 1.0, 1.0, 1.0);
    s._dx = _dx; s._dy = _dy; s._dz = _dz;
    const bool ve = c->g.ve != 0;
    HIPCHK(c, c->nd == 2 ? launch2_compute_Pt_tau(c->stream, F(c), s, ve)
                         : launch3_compute_Pt_tau(c->stream, F(c), s, ve, true));
    return PTS_OK;
}

int pts_compute_V_perf(pts_ctx* c, double _dx, double _dy, double _dz) {
    if (!c) return PTS_ERR_ARG;
    if (int rc = setdev(c)) return rc;
    if (int rc = need_iter_fields(c, false)) return rc;
    StepScalars s = scal(0.0, 1.0, 1.0, 1.0, 1.0, 1.0);
    s._dx = _dx; s._dy = _dy; s._dz = _dz;
    const bool bf = c->g.body_force != 0;
    HIPCHK(c, c->nd == 2 ? launch2_compute_V_perf(c->stream, F(c), s, bf)
                         : launch3_compute_V_perf(c->stream, F(c), s, bf, true, false));
    touched(c);
    return PTS_OK;
}

// ---- norms -----------------------------------------------------------------
static void unpack_norms(const pts_ctx* c, const double* v, pts_norms* o) {
    o->Vmin = v[0]; o->Vmax = v[1]; o->Pmin = v[2]; o->Pmax = v[3];
    for (int q = 0; q < 4; ++q) {
        o->sumsq[q] = v[4 + q];
        o->maxabs[q] = v[8 + q];
    }
    const int fs[4] = {PTS_RX, PTS_RY, PTS_RZ, PTS_DIVV};
    for (int q = 0; q < 4; ++q) {
        if (q == 2 && c->nd == 2) { o->len[q] = 0; continue; }
        int64_t s[3];
        field_shape(c->nd, c->nx, c->ny, c->nz, fs[q], s);
        o->len[q] = s[0] * s[1] * s[2];
    }
}

int pts_norms_stored(pts_ctx* c, pts_norms* out) {
    if (!c || !out) return PTS_ERR_ARG;
    if (int rc = setdev(c)) return rc;
    if (int rc = ensure_list(c, {PTS_VX, PTS_PT, PTS_RX, PTS_RY, PTS_DIVV})) return rc;
    if (c->nd == 3)
        if (int rc = ensure(c, PTS_RZ)) return rc;
    HIPCHK(c, launch_norms_stored(c->stream, F(c), c->nd, c->red_partials, c->red_out));
    HIPCHK(c, hipMemcpyAsync(c->h_red, c->red_out, sizeof(double) * kRedVals, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    unpack_norms(c, c->h_red, out);
    return PTS_OK;
}

double pts_err_from_norms(const pts_norms* n, int32_t ndim, double lx, int32_t kind, double comp[4]) {
    double err = -INFINITY;
    bool nan = false;
    for (int q = 0; q < 4; ++q) {
        double v = 0.0;
        if (!(q == 2 && ndim == 2)) {
            const double den = q == 3 ? (n->Vmax - n->Vmin) : (n->Pmax - n->Pmin);
            if (kind == 1) v = n->maxabs[q] / den * lx;
            else v = std::sqrt(n->sumsq[q]) / den * lx / std::sqrt((double)n->len[q]);
            if (v != v) nan = true;
            if (v > err) err = v;
        }
        if (comp) comp[q] = v;
    }
    return nan ? NAN : err;
}

// ---- device-resident PT iteration -------------------------------------------
static bool single_pass(const pts_ctx* c) { return c->nd == 2 || !c->two_kernel_3d; }

static bool use_dp(const pts_ctx* c);

static int prepare_iter(pts_ctx* c) {
    if (int rc = need_iter_fields(c, true)) return rc;
    if (c->nd == 3 && c->mx_state == 0 && (c->nx < 3 || c->ny < 3 || c->nz < 3)) c->mx_state = -1;
    if (use_dp(c) && c->mx_enabled && c->mx_state == 0) {
        if (int rc = ensure(c, PTS_MUSTAU)) return rc;
        HIPCHK(c, hipMemsetAsync(c->d_flag, 0, sizeof(int), c->stream));
        if (c->nd == 2) HIPCHK(c, launch_check_mustau2(c->stream, F(c), c->d_flag));
        else HIPCHK(c, launch_check_mustau3(c->stream, F(c), c->d_flag));
        int bad = 1;
        HIPCHK(c, hipMemcpyAsync(&bad, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->mx_state = bad ? -1 : 1;
    }
    if (single_pass(c)) {
        int st[16];
        const int ns = state_fields(c, st);
        for (int q = 0; q < ns; ++q)
            if (int rc = ensure_alt(c, st[q])) return rc;
        if (!c->alt_synced) {
            for (int f : kVel)
                if (field_valid(c, f))
                    HIPCHK(c, hipMemcpyAsync(c->alt[f], c->cur[f], (size_t)c->nalloc * sizeof(double),
                                             hipMemcpyDeviceToDevice, c->stream));
            c->alt_synced = true;
        }
    }
    return PTS_OK;
}

static void swap_state(pts_ctx* c) {
    int st[16];
    const int ns = state_fields(c, st);
    for (int q = 0; q < ns; ++q) std::swap(c->cur[st[q]], c->alt[st[q]]);
    c->parity ^= 1;
}

// Enqueue n iterations; for the 3D divide form, the last one stores ∇V of
// the pre-update velocities when store_last_dv is set.
static bool use_dp(const pts_ctx* c) { return c->dp_ok && c->dp_enabled; }
static bool use_mx(const pts_ctx* c) { return use_dp(c) && c->mx_enabled && c->mx_state == 1; }

static StepScalars with_dp(const pts_ctx* c, StepScalars s) {
    if (use_dp(c)) {
        s.dp_ve = c->dp.dp_ve;
        s.ip_igdt = c->dp.ip_igdt;
        s.ip_k1 = c->dp.ip_k1;
        s.ip_k2 = c->dp.ip_k2;
        s.ip_r2 = c->dp.ip_r2;
    }
    return s;
}

static bool same_scal(const StepScalars& a, const StepScalars& b) {
    return a.r == b.r && a.Gdt == b.Gdt && a.dx == b.dx && a.dy == b.dy && a.dz == b.dz && a._dx == b._dx &&
           a._dy == b._dy && a._dz == b._dz && a.dp_ve == b.dp_ve && a.ip_igdt == b.ip_igdt && a.ip_k1 == b.ip_k1 &&
           a.ip_k2 == b.ip_k2 && a.ip_r2 == b.ip_r2;
}

static bool same_fields(const Fields& a, const Fields& b) {
    for (int q = 0; q < PTS_NFIELDS; ++q)
        if (a.p[q] != b.p[q]) return false;
    return a.sx == b.sx && a.sxy == b.sxy && a.nx == b.nx && a.ny == b.ny && a.nz == b.nz;
}

static int enqueue_iters(pts_ctx* c, const StepScalars& s0, int64_t n, bool store_last_dv) {
    const bool ve = c->g.ve != 0, mul = c->g.fd_mode == PTS_FD_MUL, bf = c->g.body_force != 0;
    const bool dp = use_dp(c);
    const StepScalars s = with_dp(c, s0);
    for (int64_t q = 0; q < n; ++q) {
        if (single_pass(c)) {
            if (c->nd == 2)
                HIPCHK(c, launch2_iter(c->stream, F(c), fields_of(c, c->alt), s, ve, mul, bf, c->rows_per_block,
                                       c->variant_2d, c->persist_2d, dp, use_mx(c)));
            else
                HIPCHK(c, launch3_iter(c->stream, F(c), fields_of(c, c->alt), s, ve, mul, bf, c->planes_per_block,
                                       c->tile_rows_3d, dp, c->variant_3d, c->slot_bytes, use_mx(c),
                                       c->map_flat_3d, c->w64_3d, c->stream_3d));
            swap_state(c);
        } else {
            if (store_last_dv && q == n - 1 && !mul) {
                if (int rc = ensure(c, PTS_DIVV)) return rc;
                HIPCHK(c, launch3_compute_P(c->stream, F(c), s));
                HIPCHK(c, launch3_compute_tau(c->stream, F(c), s, ve));
            } else {
                HIPCHK(c, launch3_compute_Pt_tau(c->stream, F(c), s, ve, mul));
            }
            HIPCHK(c, launch3_compute_V_perf(c->stream, F(c), s, bf, mul, true));
        }
    }
    return PTS_OK;
}

// 2D small grids: n iterations of the default kernel in one cooperative
// launch (1: done; 0: not applicable, use the launches or the graph).
static int run_iters_multi(pts_ctx* c, const StepScalars& s0, int64_t n) {
    if (c->nd != 2 || !c->multi_2d || c->variant_2d != 7 || !use_mx(c) || n < 2 || n > (1 << 30)) return 0;
    if ((int64_t)c->nx * c->ny > c->multi_max_cells) return 0;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(c->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
        (void)hipGetLastError();
        return 0;
    }
    const bool ve = c->g.ve != 0, mul = c->g.fd_mode == PTS_FD_MUL, bf = c->g.body_force != 0;
    // The kernel reads every input from `in` and swaps in/out each iteration,
    // so both sets must be complete: the same read-only fields, the state
    // fields from cur and alt respectively.
    const Fields f0 = F(c);
    Fields f1 = f0;
    int st[16];
    const int ns = state_fields(c, st);
    for (int q = 0; q < ns; ++q) f1.p[st[q]] = c->alt[st[q]];
    std::vector<int> reads = {PTS_PT, PTS_VX, PTS_VY, PTS_TXX, PTS_TYY, PTS_TXY, PTS_MUS, PTS_MUSTAU};
    if (ve) reads.insert(reads.end(), {PTS_TXX_O, PTS_TYY_O, PTS_TXY_O});
    if (bf) reads.push_back(PTS_RHOG);
    for (int f : reads)
        if (!f0.p[f] || !f1.p[f]) return 0;
    hipError_t e = launch2_iter_multi(c->stream, f0, f1, with_dp(c, s0), ve, mul, bf, c->multi_rows_2d, (int)n,
                                      c->d_bar, (int*)(c->d_bar + 2));
    if (e == hipErrorNotSupported || e == hipErrorCooperativeLaunchTooLarge) {
        (void)hipGetLastError();
        return 0;
    }
    if (e != hipSuccess) return fail(c, PTS_ERR_HIP, std::string("cooperative launch: ") + hipGetErrorString(e));
    e = hipMemcpyAsync(c->h_bar_err, c->d_bar + 2, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e != hipSuccess) return fail(c, PTS_ERR_HIP, std::string("barrier flag: ") + hipGetErrorString(e));
    for (int64_t q = 0; q < n; ++q) swap_state(c);
    return 1;
}

// 2D small grids: n iterations with the state in LDS in one cooperative
// launch (k2_lds; 1: done; 0: not applicable).  Tiles of R rows and C
// columns, the last tile of a row or column of tiles owning what is left
// (R .. 2R-1, or more when it takes over the tile before it): R as asked (or
// the next fewer that works), C as asked (default: the whole width), each
// at least 2, within what the LDS holds (160 KiB per CU) with H halo rows
// and columns on each side (H may exceed R or C), and few enough blocks to
// be resident at once.
static int run_iters_lds(pts_ctx* c, const StepScalars& s0, int64_t n) {
    if (c->nd != 2 || !c->lds_2d || !single_pass(c) || n < 2 || n > (1 << 30)) return 0;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(c->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
        (void)hipGetLastError();
        return 0;
    }
    const bool ve = c->g.ve != 0, mul = c->g.fd_mode == PTS_FD_MUL, bf = c->g.body_force != 0;
    const bool dp = use_dp(c);
    const size_t lds_max = 160 * 1024;
    // tiles of r rows over the positions 0..m: the rows a block updates in the
    // steps of an exchange period start at r0 - (h-1) .. r0, and none may be
    // m-1 (bc_y!(Vx) copies row ny-2 into row ny-1, bc_x!(Vy) column nx-2
    // into column nx-1).  Only the last tile's start can be (it is at most
    // m+1-r): then the last tile takes over the one before.
    auto tiles = [](int m, int r, int h) {
        int b = std::max(1, (m + 1) / r);
        while (b > 1 && (b - 1) * r >= m - 1 && (b - 1) * r <= m + h - 2) --b;
        return b;
    };
    auto works = [&](int r, int w, int h) {
        if (r < 2 || r > c->ny + 1 || w < 2 || w > c->nx + 1) return false;
        const int rl = c->ny + 1 - (tiles(c->ny, r, h) - 1) * r, cl = c->nx + 1 - (tiles(c->nx, w, h) - 1) * w;
        return lds2_bytes(lds2_extent(c->ny + 1, r, rl, h), lds2_extent(c->nx + 1, w, cl, h), ve, bf) <= lds_max;
    };
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    // columns per tile: as asked, else two tiles across grids at least 64
    // wide (twice the blocks, half the arithmetic each: 127^2 4.75 -> 4.2 us
    // per iteration, 255^2 8.1 -> 7.5 us, 3 rows every 3 iterations), else
    // the whole width
    std::vector<int> widths;
    if (c->lds_cols_2d > 0) widths.push_back(std::min(c->lds_cols_2d, c->nx + 1));
    else if (c->nx + 1 >= 64) widths.push_back((c->nx + 2) / 2);
    widths.push_back(c->nx + 1);
    int R = 0, H = 0, C = 0, nby = 0, nbx = 0;
    for (int w : widths) {
        int r_ = 0, h_ = 0;
        if (c->lds_rows_2d > 0 || c->lds_halo_2d > 0) {  // as asked, or the next fewer rows that work
            h_ = c->lds_halo_2d > 0 ? c->lds_halo_2d : 2;
            for (int r = c->lds_rows_2d > 0 ? c->lds_rows_2d : 4; r >= 2 && !r_; --r)
                if (works(r, w, h_)) r_ = r;
        } else {
            // short blocks, exchanged every few iterations: many blocks in
            // parallel against the latency of an exchange through memory
            // (two tiles across; 127^2: 4 rows every 4 iterations 3.63 us per
            // iteration, 5 every 4 3.74, 3 every 4 3.75, 4 every 3 3.77, 3
            // every 3 3.82, 2 every 4 4.0; 255^2: 4 every 4 5.8, 4 every 3
            // 6.1, 3 every 3 6.7, 3 every 2 8.0)
            const int pairs[][2] = {{4, 4}, {3, 3}, {3, 2}, {4, 2}, {2, 2}, {5, 2}, {6, 2}, {7, 2}, {8, 2}};
            for (const auto& p : pairs)
                if (works(p[0], w, p[1])) {
                    r_ = p[0];
                    h_ = p[1];
                    break;
                }
        }
        if (r_ < 2) continue;
        const int ty = tiles(c->ny, r_, h_), tx = tiles(c->nx, w, h_);
        if (ty * tx > cus) continue;
        R = r_, H = h_, C = w, nby = ty, nbx = tx;
        break;
    }
    if (R < 2) return 0;
    const int nb = nby * nbx;
    const Fields f0 = F(c);
    Fields f1 = f0;
    int st[16];
    const int ns = state_fields(c, st);
    for (int q = 0; q < ns; ++q) f1.p[st[q]] = c->alt[st[q]];
    std::vector<int> reads = {PTS_PT, PTS_VX, PTS_VY, PTS_TXX, PTS_TYY, PTS_TXY, PTS_MUS};
    if (dp) reads.push_back(PTS_MUSTAU);
    else reads.insert(reads.end(), {PTS_GDTAU, PTS_DTAU_RHO});
    if (ve) reads.insert(reads.end(), {PTS_TXX_O, PTS_TYY_O, PTS_TXY_O});
    if (bf) reads.push_back(PTS_RHOG);
    for (int f : reads)
        if (!f0.p[f]) return 0;
    for (int q = 0; q < ns; ++q)
        if (!f1.p[st[q]]) return 0;
    // the exchange buffer: two parities of every block's tile (each at the
    // size of the last tiles' extents)
    const size_t xb_bytes = (size_t)2 * nb * (c->ny + 1 - (nby - 1) * R) * 6 * (c->nx + 1 - (nbx - 1) * C) *
                            sizeof(double);
    if (xb_bytes > c->lds_xb_bytes) {
        if (c->d_lds_xb) (void)hipFree(c->d_lds_xb);
        c->d_lds_xb = nullptr;
        c->lds_xb_bytes = 0;
        HIPCHK(c, hipMalloc((void**)&c->d_lds_xb, xb_bytes));
        c->lds_xb_bytes = xb_bytes;
    }
    if (nb > c->lds_blocks) {
        if (c->d_lds_flags) (void)hipFree(c->d_lds_flags);
        c->d_lds_flags = nullptr;
        c->lds_blocks = 0;
        HIPCHK(c, hipMalloc((void**)&c->d_lds_flags, (size_t)nb * sizeof(unsigned long long)));
        HIPCHK(c, hipMemsetAsync(c->d_lds_flags, 0, (size_t)nb * sizeof(unsigned long long), c->stream));
        c->lds_blocks = nb;
        c->lds_gen = 0;
        c->lds_nb = nb;
    }
    if (nb != c->lds_nb) {  // another block count (a knob changed): the counters start again at zero
        HIPCHK(c, hipMemsetAsync(c->d_lds_flags, 0, (size_t)c->lds_blocks * sizeof(unsigned long long), c->stream));
        c->lds_gen = 0;
        c->lds_nb = nb;
    }
    // (the fault hook of the tests: every wait targets a generation no
    // block reaches and gives up after its bounded spin; the kernel still
    // runs to its end, and the solve reports it)
    hipError_t e = launch2_iter_lds(c->stream, f0, f1, with_dp(c, s0), ve, mul, bf, dp, R, nby, C, nbx, H, (int)n,
                                    c->d_lds_xb, c->d_lds_flags, c->lds_gen, (int*)(c->d_bar + 2), c->lds_threads_2d,
                                    c->lds_wait_spins, c->lds_fault_wait ? (1ull << 40) : 0ull);
    if (e == hipErrorNotSupported || e == hipErrorCooperativeLaunchTooLarge || e == hipErrorInvalidValue) {
        (void)hipGetLastError();
        return 0;
    }
    if (e != hipSuccess) return fail(c, PTS_ERR_HIP, std::string("cooperative launch: ") + hipGetErrorString(e));
    // every block publishes (n - 1) / H times
    c->lds_gen += (unsigned long long)((n - 1) / H);
    if (c->lds_fault_wait) c->lds_nb = -1;  // the counters start again at zero next time
    e = hipMemcpyAsync(c->h_bar_err, c->d_bar + 2, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e != hipSuccess) return fail(c, PTS_ERR_HIP, std::string("exchange flag: ") + hipGetErrorString(e));
    for (int64_t q = 0; q < n; ++q) swap_state(c);
    return 1;
}

// After a synchronisation: a grid barrier of a cooperative launch that gave up.
static int barrier_error(pts_ctx* c) {
    if (c->h_bar_err && *c->h_bar_err)
        return fail(c, PTS_ERR_HIP, "a bounded wait of a cooperative launch gave up (results invalid)");
    return PTS_OK;
}

static int run_iters(pts_ctx* c, const StepScalars& s, int64_t n, bool store_last_dv, bool use_graph) {
    if (n <= 0) return PTS_OK;
    if (use_graph) {  // (2D only: store_last_dv concerns the 3D kernel pair)
        const int l = run_iters_lds(c, s, n);
        if (l < 0) return l;
        if (l == 1) return PTS_OK;
        const int m = run_iters_multi(c, s, n);
        if (m < 0) return m;
        if (m == 1) return PTS_OK;
    }
    if (!use_graph || !single_pass(c) || n < 2) return enqueue_iters(c, s, n, store_last_dv);
    // Capture n single-pass iterations once per (n, parity) and replay; a
    // graph captured with other buffers or scalars is captured again.
    auto key = std::make_pair(n, c->parity);
    auto it = c->graphs.find(key);
    const Fields fin = F(c), fout = fields_of(c, c->alt);
    const StepScalars sd = with_dp(c, s);
    if (it != c->graphs.end() && !(same_fields(it->second.in, fin) && same_fields(it->second.out, fout) &&
                                   same_scal(it->second.s, sd) && it->second.dp == use_dp(c) &&
                                   it->second.mx == use_mx(c))) {
        (void)hipGraphExecDestroy(it->second.exec);
        c->graphs.erase(it);
        it = c->graphs.end();
    }
    if (it == c->graphs.end()) {
        hipGraph_t graph;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_iters(c, s, n, store_last_dv);
        hipError_t e = hipStreamEndCapture(c->stream, &graph);
        if (rc) return rc;
        if (e != hipSuccess) return fail(c, PTS_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        hipGraphExec_t ex;
        e = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(c, PTS_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        // capture already swapped the buffers n times: restore, then replay
        for (int64_t q = 0; q < n; ++q) swap_state(c);
        c->graphs[key] = pts_ctx::CachedGraph{ex, fin, fout, sd, use_dp(c), use_mx(c)};
        it = c->graphs.find(key);
    }
    HIPCHK(c, hipGraphLaunch(it->second.exec, c->stream));
    for (int64_t q = 0; q < n; ++q) swap_state(c);
    return PTS_OK;
}

// ---- placement of the field blocks ----------------------------------------
// The single-pass kernels stream 16 (2D VE) to 28 (3D VE) arrays at once.
// How evenly they load the HBM channels depends on where the driver placed
// the allocations physically: for one set of allocations the iteration time
// is constant (to about 0.1 %), but across sets it differs by up to 20 %
// (3D 511³: ≈6.1 or ≈7.3 ms; 2D 8191²: 1.56–1.85 ms once the device has seen
// allocations come and go).  place_fields measures that and keeps a fast
// set: it copies every block that holds a field into a spare set, times real
// iterations (both ping-pong parities) on the context's blocks, then up to
// place_tries_3d (2D: place_tries_2d) times allocates a fresh set (the
// current ones still held, so it lands elsewhere), copies the spare into it,
// times real iterations there and keeps it if it is more than 1 % faster,
// within a time budget of place_budget_iters iterations.  The timed
// iterations advance whichever set they run on; the spare, copied into the
// kept set at the end, restores the state from before them.

struct Block {
    char* q;     // handed pointer (alloc_base key)
    char* p;     // hipMalloc'ed base
    size_t len;  // hipMalloc'ed bytes
};

static bool inside(const Block& b, const void* x) { return (const char*)x >= b.p && (const char*)x < b.p + b.len; }

static std::vector<Block> field_blocks(const pts_ctx* c) {
    std::vector<Block> out;
    for (const auto& kv : c->alloc_base) {
        auto it = c->alloc_len.find(kv.first);
        if (it == c->alloc_len.end()) continue;
        const Block b{(char*)kv.first, (char*)kv.second, it->second};
        bool holds = false;
        for (int q = 0; q < PTS_NFIELDS && !holds; ++q) holds = inside(b, c->cur[q]) || inside(b, c->alt[q]);
        if (holds) out.push_back(b);
    }
    return out;
}

// Every field pointer into from[i] moves to the same offset of to[i].
static void repoint(pts_ctx* c, const std::vector<Block>& from, const std::vector<Block>& to) {
    auto moved = [&](void* x) -> void* {
        for (size_t i = 0; i < from.size(); ++i)
            if (inside(from[i], x)) return to[i].p + ((char*)x - from[i].p);
        return x;
    };
    for (int q = 0; q < PTS_NFIELDS; ++q) {
        c->cur[q] = (double*)moved(c->cur[q]);
        c->alt[q] = (double*)moved(c->alt[q]);
    }
    std::set<void*> sp;
    for (void* x : c->slot_ptrs) sp.insert(moved(x));
    c->slot_ptrs.swap(sp);
    for (size_t i = 0; i < from.size(); ++i) {
        c->alloc_base.erase(from[i].q);
        c->alloc_len.erase(from[i].q);
        c->alloc_base[to[i].q] = to[i].p;
        c->alloc_len[to[i].q] = to[i].len;
    }
}

// A set of blocks shaped like `like` (same lengths and handed offsets), or
// an empty vector when the device cannot hold it.
static std::vector<Block> alloc_like(const std::vector<Block>& like) {
    std::vector<Block> out;
    for (const Block& b : like) {
        void* p = nullptr;
        if (hipMalloc(&p, b.len) != hipSuccess) {
            for (const Block& x : out) (void)hipFree(x.p);
            (void)hipGetLastError();
            return {};
        }
        out.push_back(Block{(char*)p + (b.q - b.p), (char*)p, b.len});
    }
    return out;
}

static hipError_t copy_set(pts_ctx* c, const std::vector<Block>& dst, const std::vector<Block>& src) {
    for (size_t i = 0; i < src.size(); ++i)
        if (hipError_t e = hipMemcpyAsync(dst[i].p, src[i].p, src[i].len, hipMemcpyDeviceToDevice, c->stream))
            return e;
    return hipSuccess;
}

static void free_set(pts_ctx* c, const std::vector<Block>& set) {
    (void)hipStreamSynchronize(c->stream);
    for (const Block& b : set) (void)hipFree(b.p);
}

// Mean time of an even number of real iterations (after two untimed ones)
// on the context's current blocks: 10 in 2D, 4 in 3D (a 2D iteration at
// 8191² takes 1.6–1.8 ms, a 3D one at 511³ 6–7 ms).  The state pointers end
// where they began; the fields have advanced.
static int time_iter(pts_ctx* c, const StepScalars& s, float* ms) {
    const int timed = c->nd == 2 ? 10 : 4;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipError_t e = hipEventCreate(&ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&ev[1]);
    int rc = e == hipSuccess ? enqueue_iters(c, s, 2, false) : PTS_OK;
    if (e == hipSuccess && rc == PTS_OK) e = hipEventRecord(ev[0], c->stream);
    if (e == hipSuccess && rc == PTS_OK) rc = enqueue_iters(c, s, timed, false);
    if (e == hipSuccess && rc == PTS_OK) e = hipEventRecord(ev[1], c->stream);
    if (e == hipSuccess && rc == PTS_OK) e = hipEventSynchronize(ev[1]);
    float t = 0.f;
    if (e == hipSuccess && rc == PTS_OK) e = hipEventElapsedTime(&t, ev[0], ev[1]);
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    if (rc) return rc;
    if (e != hipSuccess) return fail(c, PTS_ERR_HIP, std::string("placement timing: ") + hipGetErrorString(e));
    *ms = t / (float)timed;
    return PTS_OK;
}

static int place_fields(pts_ctx* c, const StepScalars& s) {
    if (c->placed) return PTS_OK;
    c->placed = true;
    const int tries = c->nd == 3 ? c->place_tries_3d : c->place_tries_2d;
    if (!single_pass(c) || tries <= 0) return PTS_OK;
    if (c->nd == 3 && (c->variant_3d < 1 || !c->group_3d)) return PTS_OK;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIPCHK(c, hipStreamIsCapturing(c->stream, &cs));
    if (cs != hipStreamCaptureStatusNone) return PTS_OK;
    std::vector<Block> best = field_blocks(c);
    size_t bytes = 0;
    for (const Block& b : best) bytes += b.len;
    if ((int64_t)bytes < c->place_min_bytes) return PTS_OK;
    size_t avail = 0, total = 0;
    if (hipMemGetInfo(&avail, &total) != hipSuccess || avail < 2 * bytes + ((size_t)1 << 30)) {
        (void)hipGetLastError();
        return PTS_OK;  // room for the spare and one candidate set is needed
    }
    const auto t_start = std::chrono::steady_clock::now();
    const std::vector<Block> spare = alloc_like(best);
    if (spare.empty()) return PTS_OK;
    if (hipError_t e = copy_set(c, spare, best)) {
        free_set(c, spare);
        return fail(c, PTS_ERR_HIP, std::string("placement copy: ") + hipGetErrorString(e));
    }
    float tb = 0.f;
    int rc = time_iter(c, s, &tb);
    c->place_ms[0] = tb;
    bool moved = false;
    // time budget: place_budget_iters iterations at the original blocks' speed
    // (at least 1 s); later tries take longer (fresh blocks from memory that
    // other allocations used are cleared by the driver first)
    const double budget_s = std::max(1.0, (double)c->place_budget_iters * tb * 1e-3);
    for (int t = 1; t <= tries && rc == PTS_OK; ++t) {
        const double spent = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
        if (spent > budget_s) break;
        if (hipMemGetInfo(&avail, &total) != hipSuccess || avail < bytes + ((size_t)4 << 30)) break;
        const std::vector<Block> cand = alloc_like(best);
        if (cand.empty()) break;
        hipError_t e = copy_set(c, cand, spare);
        if (e != hipSuccess) {
            free_set(c, cand);
            rc = fail(c, PTS_ERR_HIP, std::string("placement copy: ") + hipGetErrorString(e));
            break;
        }
        repoint(c, best, cand);
        float tc = 0.f;
        rc = time_iter(c, s, &tc);
        c->place_ms[t] = tc;
        const bool keep = rc == PTS_OK && (c->place_accept >= 0 ? c->place_accept == 1 : tc < 0.99f * tb);
        if (keep) {
            free_set(c, best);
            best = cand;
            tb = tc;
            moved = true;
        } else {
            repoint(c, cand, best);
            free_set(c, cand);
        }
    }
    // the state before the timed iterations, into the kept set
    hipError_t e = copy_set(c, best, spare);
    free_set(c, spare);
    if (e != hipSuccess && rc == PTS_OK) rc = fail(c, PTS_ERR_HIP, std::string("placement copy: ") + hipGetErrorString(e));
    if (moved) {
        for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second.exec);
        c->graphs.clear();
    }
    return rc;
}

int pts_set_knob(pts_ctx* c, const char* name, int64_t value) {
    if (!c || !name) return fail(c, PTS_ERR_ARG, "null argument");
    const std::string k(name);
    const int v = (int)std::max<int64_t>(INT32_MIN, std::min<int64_t>(INT32_MAX, value));
    if (k == "PTS_PLANES_PER_BLOCK") c->planes_per_block = v;
    else if (k == "PTS_3D_VARIANT") c->variant_3d = v;
    else if (k == "PTS_3D_MAP_FLAT") c->map_flat_3d = v != 0;
    else if (k == "PTS_3D_W64") c->w64_3d = v != 0;
    else if (k == "PTS_3D_STREAM") c->stream_3d = v != 0;
    else if (k == "PTS_ROWS_PER_BLOCK") c->rows_per_block = v;
    else if (k == "PTS_2D_VARIANT") c->variant_2d = v;
    else if (k == "PTS_2D_PERSIST") c->persist_2d = v != 0;
    else if (k == "PTS_2D_MULTI") c->multi_2d = v != 0;
    else if (k == "PTS_2D_LDS") c->lds_2d = v != 0;
    else if (k == "PTS_2D_LDS_ROWS") c->lds_rows_2d = std::max(0, v);
    else if (k == "PTS_2D_LDS_COLS") c->lds_cols_2d = std::max(0, v);
    else if (k == "PTS_2D_LDS_THREADS") c->lds_threads_2d = std::min(1024, std::max(64, v / 64 * 64));
    else if (k == "PTS_2D_LDS_HALO") c->lds_halo_2d = (int)std::min<int64_t>(4, std::max<int64_t>(0, v));
    else if (k == "PTS_2D_MULTI_ROWS") c->multi_rows_2d = std::max(1, v);
    else if (k == "PTS_DC_PLANES_PER_BLOCK") c->dc_planes_per_block = v;
    else if (k == "PTS_DC_DP") c->dc_dp = std::max(-1, std::min(1, v));
    else if (k == "PTS_DC_MAP_XCD") c->dc_map_xcd = std::max(-1, std::min(1, v));
    else return fail(c, PTS_ERR_ARG, "unknown knob " + k);
    // captured graphs hold the old launch configuration
    for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second.exec);
    c->graphs.clear();
    return PTS_OK;
}

int pts_placement_times(const pts_ctx* c, float* ms, int32_t n) {
    if (!c || !ms || n < 0) return PTS_ERR_ARG;
    for (int q = 0; q < n; ++q) ms[q] = q < 32 ? c->place_ms[q] : 0.f;
    return PTS_OK;
}

// Placement selection costs up to place_budget_iters iterations' worth of
// time, so it runs ahead of a call that plans at least that many iterations
// (pts_pt_solve: iter_max + 1) or when asked for (pts_pt_iterate with
// niter = 0); shorter calls leave it pending.
static bool place_now(const pts_ctx* c, int64_t planned) {
    return planned == 0 || planned >= c->place_budget_iters;
}

int pts_pt_iterate(pts_ctx* c, const pts_pt_params* p, int64_t niter) {
    if (!c || !p || niter < 0) return fail(c, PTS_ERR_ARG, "bad argument");
    if (int rc = setdev(c)) return rc;
    if (int rc = prepare_iter(c)) return rc;
    StepScalars s = scal(p->r, p->G, p->dt, p->dx, p->dy, c->nd == 3 ? p->dz : 1.0);
    if (place_now(c, niter))
        if (int rc = place_fields(c, s)) return rc;
    return run_iters(c, s, niter, false, false);
}

static int live_norms(pts_ctx* c, const StepScalars& s, pts_norms* out) {
    const bool mul = c->g.fd_mode == PTS_FD_MUL, bf = c->g.body_force != 0;
    if (c->nd == 3 && !mul && !single_pass(c)) {
        // R live, ∇V stored by the last iteration's compute_P
        if (int rc = ensure_list(c, {PTS_RX, PTS_RY, PTS_RZ})) return rc;
        HIPCHK(c, launch3_compute_Res(c->stream, F(c), s, bf, false));
        HIPCHK(c, launch_norms_stored(c->stream, F(c), c->nd, c->red_partials, c->red_out));
    } else {
        // divide form (4-kernel scripts): ∇V of the velocities before the last
        // update, which the ping-pong partner still holds
        Fields fv = (mul || !single_pass(c)) ? F(c) : fields_of(c, c->alt);
        HIPCHK(c, launch_norms_live(c->stream, F(c), fv, s, c->nd, bf, mul, c->red_partials, c->red_out));
    }
    HIPCHK(c, hipMemcpyAsync(c->h_red, c->red_out, sizeof(double) * kRedVals, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    unpack_norms(c, c->h_red, out);
    return PTS_OK;
}

int pts_pt_solve(pts_ctx* c, const pts_pt_params* p, const pts_solve_params* sp, pts_solve_result* res) {
    if (!c || !p || !sp || !res || sp->nout < 1) return fail(c, PTS_ERR_ARG, "bad argument");
    if (int rc = setdev(c)) return rc;
    if (int rc = prepare_iter(c)) return rc;
    StepScalars s = scal(p->r, p->G, p->dt, p->dx, p->dy, c->nd == 3 ? p->dz : 1.0);
    if (place_now(c, sp->iter_max + 1))
        if (int rc = place_fields(c, s)) return rc;
    std::memset(res, 0, sizeof(*res));
    int64_t iter = 0;
    double err = 2.0 * sp->eps;
    // while err > ε && iter <= iterMax (Stokes2D_ve.jl:153)
    while (err > sp->eps && iter <= sp->iter_max) {
        int64_t n = sp->nout - (iter % sp->nout);
        n = std::min<int64_t>(n, sp->iter_max + 1 - iter);
        const bool check = ((iter + n) % sp->nout) == 0;
        if (int rc = run_iters(c, s, n, check, sp->use_graph != 0)) return rc;
        iter += n;
        if (check) {
            pts_norms nr;
            if (int rc = live_norms(c, s, &nr)) return rc;
            if (int rc = barrier_error(c)) return rc;
            err = pts_err_from_norms(&nr, c->nd, p->lx, sp->norm_kind, res->comp);
            res->nchecks += 1;
            if (err != err) {
                res->iters = iter;
                res->err = err;
                return fail(c, PTS_ERR_NAN, "NaN in residual norm after " + std::to_string(iter) + " iterations");
            }
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (int rc = barrier_error(c)) return rc;
    res->iters = iter;
    res->err = err;
    res->converged = err <= sp->eps ? 1 : 0;
    if (!res->converged) return fail(c, PTS_ERR_NOCONV, "iter_max reached with err > eps");
    return PTS_OK;
}

// ---- mixed-precision defect correction ------------------------------------
static int ensure32(pts_ctx* c, float** slot) {
    if (*slot) return PTS_OK;
    void* p = nullptr;
    if (int rc = field_alloc(c, sizeof(float), &p)) return rc;
    *slot = (float*)p;
    return PTS_OK;
}

static Fields32 fields32_of(const pts_ctx* c, float* const* src) {
    Fields32 f;
    for (int q = 0; q < PTS_NFIELDS; ++q) f.p[q] = src[q];
    f.sx = c->sx;
    f.sxy = c->sxy;
    f.nx = c->nx;
    f.ny = c->ny;
    f.nz = c->nz;
    return f;
}

// An FP32 correction solve whose residual makes no 1 % progress over this many
// checks has reached its roundoff floor.
static const int kDcStallChecks = 40;

int pts_dc_solve(pts_ctx* c, const pts_pt_params* p, const pts_dc_params* dp, pts_dc_result* res) {
    if (!c || !p || !dp || !res || dp->nout_inner < 1 || dp->max_outer < 1) return fail(c, PTS_ERR_ARG, "bad argument");
    if (c->g.ve) return fail(c, PTS_ERR_STATE, "defect correction is implemented for viscous grids");
    if (int rc = setdev(c)) return rc;
    if (int rc = need_iter_fields(c, false)) return rc;
    std::memset(res, 0, sizeof(*res));
    int st[16];
    const int ns = state_fields(c, st);
    for (int q = 0; q < ns; ++q) {
        if (int rc = ensure32(c, &c->cur32[st[q]])) return rc;
        if (int rc = ensure32(c, &c->alt32[st[q]])) return rc;
    }
    // the 3D sweep derives Gdτ and dτ_Rho from Musτ when the context's are
    // derived from it (compute_iter_params!, no write since)
    // Where the sweep's 16 FP32 read arrays fit the 256 MB Infinity Cache it is
    // compute-bound: the flat tile order and stored Gdτ, dτ_Rho are faster
    // there (127^3: 0.066 against 0.078 ms per sweep); beyond, the XCD-wise
    // order and Gdτ, dτ_Rho from Musτ (191^3: 0.258 against 0.269 ms; 511^3:
    // 4.03 against 4.20 ms).
    const bool big = c->nd == 3 && (double)c->nalloc * sizeof(float) * 16 > 256.0 * (1 << 20);
    const bool dcdp = c->nd == 3 && (c->dc_dp < 0 ? big : c->dc_dp != 0) && use_dp(c);
    const bool dcmap = c->dc_map_xcd < 0 ? big : c->dc_map_xcd != 0;
    const std::vector<int> consts = dcdp ? std::vector<int>{PTS_MUS, PTS_MUSTAU}
                                         : std::vector<int>{PTS_MUS, PTS_GDTAU, PTS_DTAU_RHO};
    for (int f : consts) {
        if (int rc = ensure32(c, &c->cur32[f])) return rc;
        HIPCHK(c, launch_to_f32(c->stream, c->cur32[f], c->cur[f], c->nalloc));
        c->alt32[f] = c->cur32[f];  // read-only: both sides of the ping-pong see it
    }
    const int srcs[] = {PTS_DIVV, PTS_RX, PTS_RY, PTS_RZ};
    for (int f : srcs) {
        if (f == PTS_RZ && c->nd == 2) continue;
        if (int rc = ensure32(c, &c->cur32[f])) return rc;
        c->alt32[f] = c->cur32[f];
    }
    const bool bf = c->g.body_force != 0;
    const StepScalars s = scal(p->r, p->G, p->dt, p->dx, p->dy, c->nd == 3 ? p->dz : 1.0);
    StepScalars32 s32;
    s32.r = (float)s.r;
    s32._dx = (float)s._dx;
    s32._dy = (float)s._dy;
    s32._dz = (float)s._dz;
    s32.k1 = (float)c->dp.ip_k1;
    s32.kg = (float)(c->dp.ip_k2 / (c->dp.ip_k1 * c->dp.ip_r2));
    for (int outer = 0;; ++outer) {
        HIPCHK(c, launch_dc_project(c->stream, F(c), s, c->nd));
        pts_norms nr;
        HIPCHK(c, launch_norms_live(c->stream, F(c), F(c), s, c->nd, bf, true, c->red_partials, c->red_out));
        HIPCHK(c, hipMemcpyAsync(c->h_red, c->red_out, sizeof(double) * kRedVals, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        unpack_norms(c, c->h_red, &nr);
        double err = pts_err_from_norms(&nr, c->nd, p->lx, dp->norm_kind, nullptr);
        // a state with constant P or Vx (the zero initial guess) has no scale yet
        if (!(nr.Pmax > nr.Pmin) || !(nr.Vmax > nr.Vmin)) {
            bool finite = true;
            for (int q = 0; q < 4; ++q) finite = finite && (nr.sumsq[q] == nr.sumsq[q]);
            if (finite) err = INFINITY;
        }
        if (outer < 32) res->err_outer[outer] = err;
        res->err = err;
        if (err != err) return fail(c, PTS_ERR_NAN, "NaN in the FP64 residual");
        if (err <= dp->eps) {
            res->converged = 1;
            break;
        }
        if (outer >= dp->max_outer) break;
        // correction problem: sources, zero state, FP32 PT sweeps
        HIPCHK(c, launch_dc_sources(c->stream, F(c), fields32_of(c, c->cur32), s, c->nd, bf));
        for (int q = 0; q < ns; ++q) {
            HIPCHK(c, hipMemsetAsync(c->cur32[st[q]], 0, (size_t)c->nalloc * sizeof(float), c->stream));
            HIPCHK(c, hipMemsetAsync(c->alt32[st[q]], 0, (size_t)c->nalloc * sizeof(float), c->stream));
        }
        // Inner stop: the correction residual, weighted like err but with the
        // scales of this outer step frozen (and guarded: P starts at zero),
        // must drop by rtol_inner from its initial value (the outer residual).
        const double dP = nr.Pmax - nr.Pmin, dV = nr.Vmax - nr.Vmin;
        const double wR = p->lx / (dP > 0.0 ? dP : 1.0), wD = p->lx / (dV > 0.0 ? dV : 1.0);
        auto weighted = [&](const double* ss, const double* ma) {
            double e = 0.0;
            for (int q = 0; q < 4; ++q) {
                if (q == 2 && c->nd == 2) continue;
                const double w = q == 3 ? wD : wR;
                const double v = dp->norm_kind == 1 ? ma[q] * w : std::sqrt(ss[q]) * w / std::sqrt((double)nr.len[q]);
                if (v != v) return v;
                e = std::max(e, v);
            }
            return e;
        };
        const double err_in0 = weighted(nr.sumsq, nr.maxabs);
        // the last correction only needs to bring err below eps (with margin)
        const double rtol = std::isfinite(err) ? std::max(dp->rtol_inner, 0.25 * dp->eps / err) : dp->rtol_inner;
        int64_t it = 0;
        double err_in = err_in0, best = err_in0;
        int stalled = 0;  // checks since the residual last fell 1 % below its best
        while (err_in > rtol * err_in0 && it < dp->iter_max_inner && stalled < kDcStallChecks) {
            const int64_t n = std::min<int64_t>(dp->nout_inner, dp->iter_max_inner - it);
            for (int64_t q = 0; q < n; ++q) {
                HIPCHK(c, launch_dc_iter(c->stream, fields32_of(c, c->cur32), fields32_of(c, c->alt32), s32, c->nd,
                                         c->dc_planes_per_block, dcdp, dcmap));
                for (int m = 0; m < ns; ++m) std::swap(c->cur32[st[m]], c->alt32[st[m]]);
            }
            it += n;
            HIPCHK(c, launch_dc_norms(c->stream, fields32_of(c, c->cur32), s32, c->nd, c->red_partials, c->red_out));
            HIPCHK(c, hipMemcpyAsync(c->h_red, c->red_out, sizeof(double) * 8, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            err_in = weighted(c->h_red, c->h_red + 4);
            if (err_in != err_in) return fail(c, PTS_ERR_NAN, "NaN in the FP32 correction solve");
            if (err_in < 0.99 * best) {
                best = err_in;
                stalled = 0;
            } else {
                ++stalled;
            }
        }
        res->inner_iters += it;
        HIPCHK(c, launch_dc_accumulate(c->stream, F(c), fields32_of(c, c->cur32), c->nd));
        res->outer = outer + 1;
    }
    touched(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (!res->converged) return fail(c, PTS_ERR_NOCONV, "defect correction did not reach eps");
    return PTS_OK;
}


// ---- adjoint of the converged 2D PT iteration ---------------------------------
static const int kAdj2[] = {PTS_PT, PTS_TXX, PTS_TYY, PTS_TXY, PTS_VX, PTS_VY};
static const int kAdjSrc[] = {PTS_VX, PTS_VY, PTS_TXX, PTS_TYY, PTS_TXY};

static int alloc_zero(pts_ctx* c, double** slot) {
    if (*slot) return PTS_OK;
    void* p = nullptr;
    if (int rc = field_alloc(c, sizeof(double), &p)) return rc;
    *slot = (double*)p;
    return PTS_OK;
}

static bool adj_field(int f) {
    for (int q : kAdj2)
        if (q == f) return true;
    return false;
}

static int prepare_adj(pts_ctx* c) {
    if (c->nd != 2) return fail(c, PTS_ERR_STATE, "the adjoint is implemented for 2D grids");
    if (int rc = setdev(c)) return rc;
    if (int rc = ensure_list(c, {PTS_PT, PTS_VX, PTS_VY, PTS_TXX, PTS_TYY, PTS_TXY, PTS_MUS, PTS_GDTAU, PTS_DTAU_RHO}))
        return rc;
    for (int f : kAdj2) {
        if (int rc = alloc_zero(c, &c->adj[0][f])) return rc;
        if (int rc = alloc_zero(c, &c->adj[1][f])) return rc;
        if (int rc = alloc_zero(c, &c->adj_tmp[f])) return rc;
    }
    for (int f : kAdjSrc)
        if (int rc = alloc_zero(c, &c->adj_src[f])) return rc;
    return alloc_zero(c, &c->adj_grad);
}

static Fields adj_fields(const pts_ctx* c, double* const* src) {
    Fields f = F(c);
    for (int q = 0; q < PTS_NFIELDS; ++q) f.p[q] = nullptr;
    for (int q : kAdj2) f.p[q] = src[q];
    return f;
}

// sources of one adjoint solve: ∂J/∂V (with_v) and the coupling to the next
// physical step (with_tau)
static Fields src_fields(const pts_ctx* c, bool with_v, bool with_tau) {
    Fields g = F(c);
    for (int q = 0; q < PTS_NFIELDS; ++q) g.p[q] = nullptr;
    if (with_v) g.p[PTS_VX] = c->adj_src[PTS_VX], g.p[PTS_VY] = c->adj_src[PTS_VY];
    if (with_tau)
        for (int q : {PTS_TXX, PTS_TYY, PTS_TXY}) g.p[q] = c->adj_src[q];
    return g;
}

static StepScalars adj_scal(const pts_ctx* c, const pts_pt_params* p) {
    return scal(p->r, p->G, p->dt, p->dx, p->dy, 1.0);
}

// 1/(G·dt) of the visco-elastic τ update (0: viscous)
static double adj_igdt(const pts_ctx* c, const pts_pt_params* p) { return c->g.ve ? 1.0 / (p->G * p->dt) : 0.0; }

static int zero_lambda(pts_ctx* c) {
    for (int f : kAdj2)
        for (int b = 0; b < 2; ++b)
            HIPCHK(c, hipMemsetAsync(c->adj[b][f], 0, (size_t)c->nalloc * sizeof(double), c->stream));
    return PTS_OK;
}

int pts_adjoint_set_source(pts_ctx* c, const double* dJ_dVx, int64_t nvx, const double* dJ_dVy, int64_t nvy) {
    if (!c) return fail(c, PTS_ERR_ARG, "null argument");
    if (int rc = prepare_adj(c)) return rc;
    const int fs[2] = {PTS_VX, PTS_VY};
    const double* hs[2] = {dJ_dVx, dJ_dVy};
    const int64_t ns[2] = {nvx, nvy};
    for (int q = 0; q < 2; ++q) {
        HIPCHK(c, hipMemsetAsync(c->adj_src[fs[q]], 0, (size_t)c->nalloc * sizeof(double), c->stream));
        if (hs[q])
            if (int rc = copy_dense(c, fs[q], c->adj_src[fs[q]], nullptr, hs[q], ns[q], true)) return rc;
    }
    if (int rc = zero_lambda(c)) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PTS_OK;
}

int pts_adjoint_upload(pts_ctx* c, int f, const double* host, int64_t count) {
    if (!c || !host) return fail(c, PTS_ERR_ARG, "null argument");
    if (!adj_field(f)) return fail(c, PTS_ERR_ARG, "field has no adjoint");
    if (int rc = prepare_adj(c)) return rc;
    return copy_dense(c, f, c->adj[c->adj_cur][f], nullptr, host, count, true);
}

int pts_adjoint_download(pts_ctx* c, int f, double* host, int64_t count) {
    if (!c || !host) return fail(c, PTS_ERR_ARG, "null argument");
    if (!adj_field(f)) return fail(c, PTS_ERR_ARG, "field has no adjoint");
    if (int rc = prepare_adj(c)) return rc;
    return copy_dense(c, f, c->adj[c->adj_cur][f], host, nullptr, count, false);
}

static int adj_steps(pts_ctx* c, const StepScalars& s, double igdt, const Fields& g, int64_t n) {
    const Fields f = F(c), t = adj_fields(c, c->adj_tmp);
    for (int64_t q = 0; q < n; ++q) {
        const int a = c->adj_cur;
        HIPCHK(c, launch2_adj_step(c->stream, f, adj_fields(c, c->adj[a]), t, adj_fields(c, c->adj[a ^ 1]), s, igdt,
                                   g, c->adj_fused));
        c->adj_cur = a ^ 1;
    }
    return PTS_OK;
}

int pts_adjoint_iterate(pts_ctx* c, const pts_pt_params* p, int64_t niter) {
    if (!c || !p || niter < 0) return fail(c, PTS_ERR_ARG, "bad argument");
    if (int rc = prepare_adj(c)) return rc;
    if (int rc = adj_steps(c, adj_scal(c, p), adj_igdt(c, p), src_fields(c, true, false), niter)) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PTS_OK;
}

// λ <- Φᵀλ + g until max|Δλ_V| <= eps·max|λ_V|, checked every nout iterations
static int adj_solve_loop(pts_ctx* c, const StepScalars& s, double igdt, const Fields& g, const pts_solve_params* sp,
                          pts_solve_result* res) {
    int64_t iter = 0;
    double err = INFINITY;
    while (err > sp->eps && iter < sp->iter_max) {
        const int64_t n = std::min<int64_t>(sp->nout, sp->iter_max - iter);
        if (int rc = adj_steps(c, s, igdt, g, n)) return rc;
        iter += n;
        HIPCHK(c, launch2_adj_change(c->stream, F(c), adj_fields(c, c->adj[c->adj_cur]),
                                     adj_fields(c, c->adj[c->adj_cur ^ 1]), c->red_partials, c->red_out));
        HIPCHK(c, hipMemcpyAsync(c->h_red, c->red_out, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        const double d = c->h_red[0], m = c->h_red[1];
        res->comp[0] = d;
        res->comp[1] = m;
        res->nchecks += 1;
        if (d != d || m != m) {
            res->iters += iter;
            res->err = NAN;
            return fail(c, PTS_ERR_NAN, "NaN in the adjoint state after " + std::to_string(iter) + " iterations");
        }
        err = m > 0.0 ? d / m : 0.0;
    }
    res->iters += iter;
    res->err = std::max(res->err, err);
    if (err > sp->eps) res->converged = 0;
    return PTS_OK;
}

int pts_adjoint_solve(pts_ctx* c, const pts_pt_params* p, const pts_solve_params* sp, pts_solve_result* res) {
    if (!c || !p || !sp || !res || sp->nout < 1) return fail(c, PTS_ERR_ARG, "bad argument");
    if (int rc = prepare_adj(c)) return rc;
    std::memset(res, 0, sizeof(*res));
    res->converged = 1;
    if (int rc = adj_solve_loop(c, adj_scal(c, p), adj_igdt(c, p), src_fields(c, true, false), sp, res)) return rc;
    if (!res->converged) return fail(c, PTS_ERR_NOCONV, "adjoint: iter_max reached with err > eps");
    return PTS_OK;
}

int pts_grad_logmu(pts_ctx* c, const pts_pt_params* p, double* grad, int64_t count) {
    if (!c || !p || !grad) return fail(c, PTS_ERR_ARG, "null argument");
    if (int rc = prepare_adj(c)) return rc;
    if (count != (int64_t)c->nx * c->ny)
        return fail(c, PTS_ERR_ARG, "count " + std::to_string(count) + " != nx*ny");
    const Fields t = adj_fields(c, c->adj_tmp);
    HIPCHK(c, launch2_adj_pass1(c->stream, F(c), adj_fields(c, c->adj[c->adj_cur]), t, adj_scal(c, p)));
    HIPCHK(c, launch2_adj_grad(c->stream, F(c), t, adj_igdt(c, p), c->adj_grad, false));
    HIPCHK(c, hipMemcpyAsync(grad, c->adj_grad, (size_t)count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PTS_OK;
}

int pts_checkpoint(pts_ctx* c, int32_t slot) {
    if (!c || slot < 0 || slot > 100000) return fail(c, PTS_ERR_ARG, "bad checkpoint slot");
    if (c->nd != 2) return fail(c, PTS_ERR_STATE, "checkpoints are implemented for 2D grids");
    if (int rc = setdev(c)) return rc;
    if (int rc = ensure_list(c, {PTS_TXX, PTS_TYY, PTS_TXY})) return rc;
    if ((int)c->ckpt.size() <= slot) c->ckpt.resize(slot + 1, std::array<double*, 3>{nullptr, nullptr, nullptr});
    const int fs[3] = {PTS_TXX, PTS_TYY, PTS_TXY};
    for (int q = 0; q < 3; ++q) {
        if (int rc = alloc_zero(c, &c->ckpt[slot][q])) return rc;
        HIPCHK(c, hipMemcpyAsync(c->ckpt[slot][q], c->cur[fs[q]], (size_t)c->nalloc * sizeof(double),
                                 hipMemcpyDeviceToDevice, c->stream));
    }
    return PTS_OK;
}

int pts_adjoint_ve_solve(pts_ctx* c, const pts_pt_params* p, int32_t nsteps, const pts_solve_params* sp,
                         pts_solve_result* res, double* grad, int64_t count) {
    if (!c || !p || !sp || !res || !grad || sp->nout < 1 || nsteps < 1) return fail(c, PTS_ERR_ARG, "bad argument");
    if (!c->g.ve) return fail(c, PTS_ERR_STATE, "pts_adjoint_ve_solve needs a visco-elastic grid");
    if (count != (int64_t)c->nx * c->ny) return fail(c, PTS_ERR_ARG, "count " + std::to_string(count) + " != nx*ny");
    if ((int)c->ckpt.size() < nsteps) return fail(c, PTS_ERR_STATE, "missing checkpoints");
    for (int n = 0; n < nsteps; ++n)
        if (!c->ckpt[n][0]) return fail(c, PTS_ERR_STATE, "missing checkpoint " + std::to_string(n));
    if (int rc = prepare_adj(c)) return rc;
    std::memset(res, 0, sizeof(*res));
    res->converged = 1;
    const StepScalars s = adj_scal(c, p);
    const double igdt = adj_igdt(c, p);
    const Fields t = adj_fields(c, c->adj_tmp);
    for (int n = nsteps - 1; n >= 0; --n) {
        // step n: λ = Φᵀλ + ∂J/∂V (last step) + coupling from step n+1
        if (int rc = zero_lambda(c)) return rc;
        const Fields g = src_fields(c, n == nsteps - 1, n < nsteps - 1);
        if (int rc = adj_solve_loop(c, s, igdt, g, sp, res)) return rc;
        HIPCHK(c, launch2_adj_pass1(c->stream, F(c), adj_fields(c, c->adj[c->adj_cur]), t, s));
        Fields fk = F(c);  // the converged stresses of step n
        fk.p[PTS_TXX] = c->ckpt[n][0];
        fk.p[PTS_TYY] = c->ckpt[n][1];
        fk.p[PTS_TXY] = c->ckpt[n][2];
        HIPCHK(c, launch2_adj_grad(c->stream, fk, t, igdt, c->adj_grad, n < nsteps - 1));
        if (n > 0) HIPCHK(c, launch2_adj_couple(c->stream, F(c), t, igdt, src_fields(c, false, true)));
    }
    HIPCHK(c, hipMemcpyAsync(grad, c->adj_grad, (size_t)count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (!res->converged) return fail(c, PTS_ERR_NOCONV, "adjoint: iter_max reached with err > eps");
    return PTS_OK;
}

}  // extern "C"
