// Host side of the multigrid projection solver and its C ABI (include/stzmg.h).
//
// The hierarchy, cycle schedule and solve loop follow the reference
// multigrid<vec3,M> (mg_3d_template/multigrid.cc:136-297) and the projection
// loop of shear_sim_3d::projection_step (src/shear_sim_3d.cpp:529-600); all
// arithmetic runs in the kernels of mg_kernels.hpp on one HIP stream.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <mutex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <algorithm>
#include <tuple>
#include <vector>

#include "../../include/stzmg.h"
#include "mg_coarse.hpp"
#include "mg_internal.hpp"
#include "mg_fused.hpp"
#include "mg_kernels.hpp"
#include "mg_march.hpp"
#include "mg_resid.hpp"
#include "mg_krylov.hpp"

using namespace stzmg;

namespace {

struct Level {
    Dims d{};
    int mode = 0;
    long ncls = 0;
    double *ops64 = nullptr, *inv64 = nullptr;
    float *ops32 = nullptr, *inv32 = nullptr;
    double *x64 = nullptr, *r64 = nullptr, *t64 = nullptr;
    float *x32 = nullptr, *r32 = nullptr, *t32 = nullptr;
    // Second solution buffers of the fused smoother (mg_fused.hpp); x64/x32
    // and these swap after every fused sweep.
    double *x64b = nullptr;
    float *x32b = nullptr;
    bool fused = false;   // fused two-launch sweeps on this level
    bool tiled = false;   // LDS plane-marching residual on this level
    bool tiled_small = false;  // the same with the small tiles (no norm partials)
    int ftile = 0;        // fused-smoother tile class of FP32 sweeps (fused_tile)
    int ftile64 = 0;      // and of FP64 sweeps
    int pat = 0;          // operator zero pattern (PAT_*)
    bool sym = false;     // every plane's stencil centrosymmetric (k_gs_march pairs neighbours)
    // packed per-plane coefficients of the z-marching sweep (k_pack_march)
    double *pk64 = nullptr;
    float *pk32 = nullptr;
    // and of the marching residual (the centre block in place of its inverse)
    double *pkr64 = nullptr;
    float *pkr32 = nullptr;
};

inline unsigned nblk(long n, int t = kThreads) { return (unsigned)((n + t - 1) / t); }

// Device memory of stzmg_norm_noalloc: the per-workgroup norm partials
// (pairs sum |r|^4, sum |r|^2) followed by the reduced pair.  Each GPU the
// library runs on holds its own copy (1 MiB) until the process ends; host
// code reaches it only through the address hipGetSymbolAddress returns for
// the current device (scratch_on_device).
constexpr long kNoallocParts = 65536;
static __device__ __align__(256) double scratch[2 * kNoallocParts + 2];

constexpr int kMaxDevices = 64;

// Address of this device's scratch, looked up once per device with that
// device current; null on failure (the HIP error in *err).
double *scratch_on_device(int dev, hipError_t *err) {
    static std::mutex mu;
    static double *addr[kMaxDevices] = {};
    *err = hipSuccess;
    if (dev < 0 || dev >= kMaxDevices) {
        *err = hipErrorInvalidDevice;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        *err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (*err != hipSuccess) return nullptr;
        addr[dev] = static_cast<double *>(p);
    }
    return addr[dev];
}

}  // namespace

struct stzmg_s {
    stzmg_desc desc{};
    std::vector<Level> lv;
    hipStream_t stream = nullptr;
    std::string err;
    double *part = nullptr;   // per-block norm partials
    long part_cap = 0;
    double *dres = nullptr;   // reduced norms (2 doubles)
    int *dflag = nullptr;     // singular-block flag
    unsigned *dcfg = nullptr; // k_level_config words of every level
    double *hres = nullptr;   // pinned host copy of dres (reference loop)
    hipEvent_t evres = nullptr;  // recorded after that copy
    bool ops_set = false;
    int std_iters = 16;       // per-handle copy of the reference's static heuristic
    long coarse_max = kCoarseMaxPoints;  // levels this small run in one workgroup (0: off)
    int fused_smoother = 1;   // two-launch fused sweeps on eligible levels
    int tiled_resid = 1;      // plane-marching residual on eligible levels
    int centro_galerkin = 1;  // k_centro_sym after each per-plane Galerkin product
    int bottom_map = 1;       // coarsest solve as x = Q r (mg_coarse.hpp)
    long chunk_cap = 0;       // cap on the workgroups the z-chunked launches assume resident (0: none)
    // stzmg_time_kernel kinds 4, 5: event pairs recorded around each fused
    // sweep launch (null: no recording)
    std::vector<hipEvent_t> *launch_ev = nullptr;
    size_t launch_ev_next = 0;
    int restrict_planes = 0;  // coarse planes per restriction thread (0: chosen from the grid)
    double *qmap = nullptr;   // Q, row-major, qmap_n x qmap_n
    long qmap_n = 0;
    bool qmap_ok = false;     // Q belongs to the current operator
    // flexible GMRES workspace (allocated on first use)
    double *kV = nullptr;     // (restart + 1) FP64 basis vectors
    float *kZ = nullptr;      // restart FP32 preconditioned vectors
    double *ktmp = nullptr;   // one FP64 vector
    double *kpart = nullptr;  // dot partials
    double *kdev = nullptr;   // reduced dots / coefficients
    int kcap = 0;             // restart length the workspace holds
    float *kV32 = nullptr;    // FP32 Krylov basis of the GMRES-IR solve
    double *kZ64 = nullptr;   // FP64 preconditioned vectors of the FP64 GMRES solve
    double *kX64 = nullptr, *kB64 = nullptr;  // its solution and right-hand side
    int kcapz64 = 0;
    int kcap64 = 0, kcap32 = 0;  // restart lengths the two bases hold
    hipGraphExec_t g64 = nullptr, g32 = nullptr;
    hipGraphExec_t g64z = nullptr, g32z = nullptr;  // V-cycles that start from x = 0
    double *scratch = nullptr;  // this device's scratch of stzmg_norm_noalloc
    int *pkmap = nullptr;       // k_pack_march source maps, kPackMax ints per pattern
    int *pkrmap = nullptr;      // and those of the residual tables
};

namespace {

int fail(stzmg_t h, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (h) h->err = buf;
    return code;
}

#define HIPCHK(h, call)                                                                \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail((h), STZMG_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

#define LAUNCHCHK(h)                                                                   \
    do {                                                                               \
        hipError_t e_ = hipGetLastError();                                             \
        if (e_ != hipSuccess)                                                          \
            return fail((h), STZMG_ERR_HIP, "kernel launch failed: %s", hipGetErrorString(e_)); \
    } while (0)

int coarse_size(int nf, int prd) { return (nf + (prd ? 1 : 2)) >> 1; }

void free_level(Level &L) {
    hipFree(L.ops64); hipFree(L.inv64); hipFree(L.ops32); hipFree(L.inv32);
    hipFree(L.x64); hipFree(L.r64); hipFree(L.t64);
    hipFree(L.x32); hipFree(L.r32); hipFree(L.t32);
    hipFree(L.x64b); hipFree(L.x32b);
    hipFree(L.pk64); hipFree(L.pk32);
    hipFree(L.pkr64); hipFree(L.pkr32);
    L = Level{};
}

template <typename T>
hipError_t alloc(T **p, long n) {
    return hipMalloc((void **)p, sizeof(T) * (size_t)(n > 0 ? n : 1));
}

// ---- kernel launch wrappers ------------------------------------------------

template <typename T>
const T *ops_of(const Level &L);
template <>
const double *ops_of<double>(const Level &L) { return L.ops64; }
template <>
const float *ops_of<float>(const Level &L) { return L.ops32; }
template <typename T>
const T *inv_of(const Level &L);
template <>
const double *inv_of<double>(const Level &L) { return L.inv64; }
template <>
const float *inv_of<float>(const Level &L) { return L.inv32; }
template <typename T>
T *x_of(Level &L);
template <>
double *x_of<double>(Level &L) { return L.x64; }
template <>
float *x_of<float>(Level &L) { return L.x32; }
template <typename T>
T *r_of(Level &L);
template <>
double *r_of<double>(Level &L) { return L.r64; }
template <>
float *r_of<float>(Level &L) { return L.r32; }
template <typename T>
T *t_of(Level &L);
template <>
double *t_of<double>(Level &L) { return L.t64; }
template <>
float *t_of<float>(Level &L) { return L.t32; }

template <typename T>
T *pkr_of(Level &L);
template <>
double *pkr_of<double>(Level &L) { return L.pkr64; }
template <>
float *pkr_of<float>(Level &L) { return L.pkr32; }

// Tiles of the plane-marching residual (k_resid_march): 64 x 16 in FP32,
// 32 x 16 in FP64, 256 threads.
template <typename T>
struct ResidTile;
template <>
struct ResidTile<float> {
    static constexpr int TX = 64, TY = 16;
};
template <>
struct ResidTile<double> {
    static constexpr int TX = 32, TY = 16;
};

bool tiled_ok(const Level &L) {
    return L.mode == STZMG_OP_PLANE && L.d.px && L.d.py && L.d.m % ResidTile<float>::TX == 0 &&
           L.d.n % ResidTile<float>::TY == 0 && L.d.m % ResidTile<double>::TX == 0 &&
           L.d.n % ResidTile<double>::TY == 0;
}

// Smaller tiles of the marching residual (16 x 16, one point per thread) for
// plane levels periodic in x and y that the tiles above do not divide (the
// 96^2 and 48^2 levels of a 192^2 grid, 32^3 and 16^3 below 256^3); used for
// the V-cycle's residuals, which take no norm partials.
constexpr int kResidSmall = 16;
bool tiled_small_ok(const Level &L) {
    return L.mode == STZMG_OP_PLANE && L.d.px && L.d.py && !tiled_ok(L) && L.d.m % kResidSmall == 0 &&
           L.d.n % kResidSmall == 0;
}

// Workgroups of the plane-marching residual in type T resident on the
// device at once (CUs x the occupancy its LDS and registers allow; every
// variant stages the same planes), queried once.
template <typename T>
long resid_resident() {
    static const long r = [] {
        int dev = 0, ncu = 0, per = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
            ncu = 1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per,
                reinterpret_cast<const void *>(
                    &k_resid_march<T, PAT_DENSE, true, false, ResidTile<T>::TX, ResidTile<T>::TY, kThreads>),
                kThreads, 0) != hipSuccess ||
            per < 1)
            per = 1;
        return (long)per * ncu;
    }();
    return r;
}

// Workgroups of the residual with norm partials: a fixed count (768 FP32,
// 1536 FP64; k_resid_march keeps three workgroups per CU of a 256-CU
// MI355X in either type, so the FP32 grid is one round and the FP64 grid
// two), so that the partial sums are grouped the same way on every device
// and a stopping decision taken on the norm does not depend on the
// device's CU count.
template <typename T>
constexpr long kNormResident = sizeof(T) == 4 ? 768 : 1536;

// Grid of the plane-marching residual: the z range is cut into as many
// chunks as let every tile's chunks be resident at once, so a launch runs
// as one round of workgroups, each staging its first planes once.  With
// norm partials the chunking follows kNormResident instead of the device.
template <typename T>
dim3 tiled_grid(const Level &L, long cap, int *kchunk, bool norm) {
    const long tiles = (long)(L.d.m / ResidTile<T>::TX) * (L.d.n / ResidTile<T>::TY);
    const long dev = norm ? kNormResident<T> : resid_resident<T>();
    const long resident = cap > 0 ? std::min(cap, dev) : dev;
    const long nz = std::min<long>(L.d.o, std::max<long>(1, resident / tiles));
    const int kc = (int)((L.d.o + nz - 1) / nz);
    *kchunk = kc;
    return dim3(L.d.m / ResidTile<T>::TX, L.d.n / ResidTile<T>::TY, (L.d.o + kc - 1) / kc);
}

// Number of norm partials the residual of level L writes.
template <typename T>
long resid_parts(stzmg_t h, const Level &L) {
    if (L.tiled) {
        int kc;
        const dim3 g = tiled_grid<T>(L, h->chunk_cap, &kc, true);
        return (long)g.x * g.y * g.z;
    }
    return (long)nblk((long)L.d.m * L.d.n) * L.d.o;
}

template <typename T, bool APPLY, int PAT>
void launch_resid_small(stzmg_t h, Level &L, const T *x, const T *r, T *out, T *xcopy) {
    // z chunks: about 1024 workgroups
    const long tiles = (long)(L.d.m / kResidSmall) * (L.d.n / kResidSmall);
    const long nz = std::min<long>(L.d.o, std::max<long>(1, 1024 / tiles));
    const int kc = (int)((L.d.o + nz - 1) / nz);
    dim3 g(L.d.m / kResidSmall, L.d.n / kResidSmall, (L.d.o + kc - 1) / kc);
    constexpr int S = kResidSmall;
    if (L.sym)
        hipLaunchKernelGGL((k_resid_march<T, PAT + kMarchSym, false, APPLY, S, S, kThreads>), g, dim3(kThreads), 0,
                           h->stream, L.d, x, r, (const T *)pkr_of<T>(L), out, (double *)nullptr, kc, xcopy);
    else
        hipLaunchKernelGGL((k_resid_march<T, PAT, false, APPLY, S, S, kThreads>), g, dim3(kThreads), 0, h->stream,
                           L.d, x, r, (const T *)pkr_of<T>(L), out, (double *)nullptr, kc, xcopy);
}

template <typename T, bool NORM, bool APPLY, int PAT>
void launch_resid_tiled(stzmg_t h, Level &L, const T *x, const T *r, T *out, T *xcopy, double *part) {
    int kc;
    const dim3 g = tiled_grid<T>(L, h->chunk_cap, &kc, NORM);
    constexpr int TX = ResidTile<T>::TX, TY = ResidTile<T>::TY;
    if (L.sym)
        hipLaunchKernelGGL((k_resid_march<T, PAT + kMarchSym, NORM, APPLY, TX, TY, kThreads>), g, dim3(kThreads), 0,
                           h->stream, L.d, x, r, (const T *)pkr_of<T>(L), out, part, kc, xcopy);
    else
        hipLaunchKernelGGL((k_resid_march<T, PAT, NORM, APPLY, TX, TY, kThreads>), g, dim3(kThreads), 0, h->stream,
                           L.d, x, r, (const T *)pkr_of<T>(L), out, part, kc, xcopy);
}

// out = r - A x (APPLY: out = A x) on level L, optionally with norm partials
// into part (h->part when null).  With APPLY and xcopy given, x is also
// copied to xcopy.
template <typename T, bool NORM, bool APPLY>
void launch_resid_any(stzmg_t h, Level &L, const T *x, const T *r, T *out, T *xcopy = nullptr,
                      double *part = nullptr) {
    if (!part) part = h->part;
    if (L.tiled) {
        switch (L.pat) {
            case PAT_SHEAR: launch_resid_tiled<T, NORM, APPLY, PAT_SHEAR>(h, L, x, r, out, xcopy, part); break;
            case PAT_TRANS_XZ: launch_resid_tiled<T, NORM, APPLY, PAT_TRANS_XZ>(h, L, x, r, out, xcopy, part); break;
            case PAT_NOCORNER: launch_resid_tiled<T, NORM, APPLY, PAT_NOCORNER>(h, L, x, r, out, xcopy, part); break;
            default: launch_resid_tiled<T, NORM, APPLY, PAT_DENSE>(h, L, x, r, out, xcopy, part); break;
        }
        return;
    }
    if (!NORM && L.tiled_small) {
        switch (L.pat) {
            case PAT_SHEAR: launch_resid_small<T, APPLY, PAT_SHEAR>(h, L, x, r, out, xcopy); break;
            case PAT_TRANS_XZ: launch_resid_small<T, APPLY, PAT_TRANS_XZ>(h, L, x, r, out, xcopy); break;
            case PAT_NOCORNER: launch_resid_small<T, APPLY, PAT_NOCORNER>(h, L, x, r, out, xcopy); break;
            default: launch_resid_small<T, APPLY, PAT_DENSE>(h, L, x, r, out, xcopy); break;
        }
        return;
    }
    if (APPLY && xcopy)
        (void)hipMemcpyAsync(xcopy, x, sizeof(T) * 3 * L.d.N, hipMemcpyDeviceToDevice, h->stream);
    dim3 g(nblk((long)L.d.m * L.d.n), L.d.o);
    if (L.mode == STZMG_OP_PLANE)
        hipLaunchKernelGGL((k_resid<T, 0, NORM, APPLY>), g, dim3(kThreads), 0, h->stream, L.d, x, r,
                           ops_of<T>(L), out, part);
    else
        hipLaunchKernelGGL((k_resid<T, 1, NORM, APPLY>), g, dim3(kThreads), 0, h->stream, L.d, x, r,
                           ops_of<T>(L), out, part);
}

template <typename T, bool NORM>
void launch_resid(stzmg_t h, Level &L, const T *x, const T *r, T *out, double *part = nullptr) {
    launch_resid_any<T, NORM, false>(h, L, x, r, out, nullptr, part);
}

template <typename T>
void launch_apply(stzmg_t h, Level &L, const T *x, T *out, T *xcopy = nullptr) {
    launch_resid_any<T, false, true>(h, L, x, (const T *)nullptr, out, xcopy);
}

int ncol(int n, int prd) { return (prd && (n & 1)) ? 3 : 2; }

// Tile and workgroup size of the large fused-smoother class (k_gs_march):
// 64 x 16 with 512 threads in FP32 (three staged planes of 71 x 24 points,
// 61,344 B of LDS: two workgroups per CU) and 32 x 16 with 256 threads in
// FP64 (39 x 24 points, 67,392 B: two per CU).  Measured on the 256^3
// transformed-operator sweep (MI355X): FP32 64 x 16 / 512 threads 0.256 ms,
// 64 x 8 / 256 0.272 ms, 32 x 16 / 256 0.291 ms; FP64 32 x 16 0.519 ms,
// 32 x 8 0.666 ms.
#ifndef STZ_F32_TX
#define STZ_F32_TX 64
#endif
#ifndef STZ_F32_TY
#define STZ_F32_TY 16
#endif
#ifndef STZ_F32_NT
#define STZ_F32_NT 512
#endif
#ifndef STZ_F64_TX
#define STZ_F64_TX 32
#endif
#ifndef STZ_F64_TY
#define STZ_F64_TY 16
#endif
#ifndef STZ_F64_NT
#define STZ_F64_NT 256
#endif
template <typename T>
struct FusedTile;
template <>
struct FusedTile<float> {
    static constexpr int TX = STZ_F32_TX, TY = STZ_F32_TY, NT = STZ_F32_NT;
};
template <>
struct FusedTile<double> {
    static constexpr int TX = STZ_F64_TX, TY = STZ_F64_TY, NT = STZ_F64_NT;
};

template <typename T>
T *&xb_of(Level &L);
template <>
double *&xb_of<double>(Level &L) { return L.x64b; }
template <>
float *&xb_of<float>(Level &L) { return L.x32b; }
template <typename T>
T *&xref_of(Level &L);
template <>
double *&xref_of<double>(Level &L) { return L.x64; }
template <>
float *&xref_of<float>(Level &L) { return L.x32; }

// Small levels (sides multiples of 16, e.g. 32^3 and 16^3) use 16 x 16
// tiles with 128 threads: one thread per point of the widest sub-pass.
// Sides that are multiples of 8 only (e.g. 24 x 24, a level of the
// 192 x 192 grids) use 8 x 8 tiles with one wave.
constexpr int kSmallTile = 16, kSmallThreads = 128;
constexpr int kTinyTile = 8, kTinyThreads = 64;

// Fused-smoother tile class of a level for sweeps in type T: 0 none, 1 the
// large tiles of T (64 x 8 FP32, 32 x 16 FP64), 2 the 16 x 16 tiles, 3 the
// 8 x 8 tiles.  The largest tile that divides the
// plane is taken unless one launch would then have fewer than
// STZ_FUSED_MIN_GROUPS workgroups (tiles times planes of one parity) and a
// smaller tile divides the plane too.
#ifndef STZ_FUSED_MIN_GROUPS
#define STZ_FUSED_MIN_GROUPS 0
#endif
template <typename T>
int fused_tile(const Level &L) {
    if (!(L.mode == STZMG_OP_PLANE && L.d.px && L.d.py && L.d.o >= 4 && L.d.N > 1024)) return 0;
    const long nP = (L.d.o + 1) / 2;
    auto groups = [&](int tx, int ty) { return (long)(L.d.m / tx) * (L.d.n / ty) * nP; };
    int best = 0;
    if (L.d.m % FusedTile<T>::TX == 0 && L.d.n % FusedTile<T>::TY == 0) {
        best = 1;
        if (groups(FusedTile<T>::TX, FusedTile<T>::TY) >= STZ_FUSED_MIN_GROUPS) return best;
    }
    if (L.d.m % kSmallTile == 0 && L.d.n % kSmallTile == 0) {
        best = 2;
        if (groups(kSmallTile, kSmallTile) >= STZ_FUSED_MIN_GROUPS) return best;
    }
    if (L.d.m % kTinyTile == 0 && L.d.n % kTinyTile == 0) best = 3;
    return best;
}

// Every tile class sweeps with the z-marching kernel k_gs_march
// (mg_march.hpp); STZ_MARCH=0 builds the staged-broadcast one-plane kernel
// k_gs_planes instead (for comparison; without the zero-start sweeps).
#ifndef STZ_MARCH
#define STZ_MARCH 1
#endif
template <typename T, int PAT, int TX, int TY, int NT, bool MARCH>
const void *sweep_kernel() {
    if constexpr (MARCH)
        return reinterpret_cast<const void *>(&k_gs_march<T, PAT, TX, TY, NT>);
    else
        return reinterpret_cast<const void *>(&k_gs_planes<T, PAT, TX, TY, NT>);
}

template <typename T>
T *pk_of(Level &L);
template <>
double *pk_of<double>(Level &L) { return L.pk64; }
template <>
float *pk_of<float>(Level &L) { return L.pk32; }

// Workgroups of a sweep kernel resident on the device at once (CUs x the
// occupancy its LDS and registers allow), queried once per kernel.
template <typename T, int PAT, int TX, int TY, int NT, bool MARCH>
long fused_resident() {
    static const long r = [] {
        int dev = 0, ncu = 0, per = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
            ncu = 1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, sweep_kernel<T, PAT, TX, TY, NT, MARCH>(), NT, 0) !=
                hipSuccess ||
            per < 1)
            per = 1;
        return (long)per * ncu;
    }();
    return r;
}

// zero_start: the sweep starts from x = 0 (its buffer need not hold the
// zeros; k_gs_march only, see zero_start_ok).
template <typename T, int PAT, int TX, int TY, int NT, bool MARCH = false>
void launch_fused_tile(stzmg_t h, Level &L, bool zero_start = false) {
    const Dims &d = L.d;
    T *x = xref_of<T>(L), *xb = xb_of<T>(L);
    const long tiles = (long)(d.m / TX) * (d.n / TY);
    const long dev_resident = fused_resident<T, PAT, TX, TY, NT, MARCH>();
    const long resident = h->chunk_cap > 0 ? std::min(h->chunk_cap, dev_resident) : dev_resident;
    for (int P = 0; P < 2; P++) {
        const int nP = (d.o - P + 1) / 2;
        // Planes per workgroup: the z range is cut into as many chunks as
        // let every tile's chunks be resident at once, so the launch runs as
        // one round of workgroups and each stages its first planes once
        // (same-parity planes do not interact: the chunking leaves the
        // results unchanged).
        const long nz = std::min<long>(nP, std::max<long>(1, resident / tiles));
        const int kchunk = (int)((nP + nz - 1) / nz);
        dim3 g(d.m / TX, d.n / TY, (nP + kchunk - 1) / kchunk);
        const T *xo = P == 0 ? x : xb;
        const bool rec = h->launch_ev && h->launch_ev_next + 2 <= h->launch_ev->size();
        if (rec) (void)hipEventRecord((*h->launch_ev)[h->launch_ev_next], h->stream);
        if constexpr (MARCH)
            hipLaunchKernelGGL((k_gs_march<T, PAT, TX, TY, NT>), g, dim3(NT), 0, h->stream, d, (const T *)x, xo, xb,
                               (const T *)r_of<T>(L), (const T *)pk_of<T>(L), P, kchunk,
                               zero_start ? (P == 0 ? 3 : 1) : 0);
        else
            hipLaunchKernelGGL((k_gs_planes<T, PAT, TX, TY, NT>), g, dim3(NT), 0, h->stream, d, (const T *)x, xo,
                               xb, (const T *)r_of<T>(L), ops_of<T>(L), inv_of<T>(L), P, kchunk);
        if (rec) {
            (void)hipEventRecord((*h->launch_ev)[h->launch_ev_next + 1], h->stream);
            h->launch_ev_next += 2;
        }
    }
    std::swap(xref_of<T>(L), xb_of<T>(L));
}

template <typename T, int PAT>
void launch_fused_pat(stzmg_t h, Level &L, bool zero_start) {
    const int ft = sizeof(T) == 4 ? L.ftile : L.ftile64;
    if (ft == 1) {
        constexpr int TX = FusedTile<T>::TX, TY = FusedTile<T>::TY, NT = FusedTile<T>::NT;
        if constexpr (STZ_MARCH != 0) {
            if (L.sym)
                launch_fused_tile<T, PAT + kMarchSym, TX, TY, NT, true>(h, L, zero_start);
            else
                launch_fused_tile<T, PAT, TX, TY, NT, true>(h, L, zero_start);
        } else {
            launch_fused_tile<T, PAT, TX, TY, NT, false>(h, L);
        }
    }
    else if (ft == 2) {
        if constexpr (STZ_MARCH != 0) {
            if (L.sym)
                launch_fused_tile<T, PAT + kMarchSym, kSmallTile, kSmallTile, kSmallThreads, true>(h, L, zero_start);
            else
                launch_fused_tile<T, PAT, kSmallTile, kSmallTile, kSmallThreads, true>(h, L, zero_start);
        } else {
            launch_fused_tile<T, PAT, kSmallTile, kSmallTile, kSmallThreads>(h, L);
        }
    }
    else {
        if constexpr (STZ_MARCH != 0) {
            if (L.sym)
                launch_fused_tile<T, PAT + kMarchSym, kTinyTile, kTinyTile, kTinyThreads, true>(h, L, zero_start);
            else
                launch_fused_tile<T, PAT, kTinyTile, kTinyTile, kTinyThreads, true>(h, L, zero_start);
        } else {
            launch_fused_tile<T, PAT, kTinyTile, kTinyTile, kTinyThreads>(h, L);
        }
    }
}

template <typename T>
void launch_gs_sweep_colours(stzmg_t h, Level &L);

// Whether a sweep of level L can start from x = 0 without the zeros in
// memory (the fused sweep k_gs_march skips the planes it knows are zero).
bool zero_start_ok(const Level &L) { return STZ_MARCH != 0 && L.fused; }

// One Gauss-Seidel sweep of a level: two fused launches where the level
// allows, eight colour launches otherwise.  zero_start (only where
// zero_start_ok): the sweep starts from x = 0 whatever x holds.
template <typename T>
void launch_gs_sweep(stzmg_t h, Level &L, bool zero_start = false) {
    if (!L.fused) {
        launch_gs_sweep_colours<T>(h, L);
        return;
    }
    switch (L.pat) {
        case PAT_SHEAR: launch_fused_pat<T, PAT_SHEAR>(h, L, zero_start); break;
        case PAT_TRANS_XZ: launch_fused_pat<T, PAT_TRANS_XZ>(h, L, zero_start); break;
        case PAT_NOCORNER: launch_fused_pat<T, PAT_NOCORNER>(h, L, zero_start); break;
        default: launch_fused_pat<T, PAT_DENSE>(h, L, zero_start); break;
    }
}

// Put a level's solution back into the buffer it started in (a level swept
// an odd number of times by the fused smoother ends in the other one).  A
// captured V-cycle must leave every pointer as it found it.
template <typename T>
void restore_level(stzmg_t h, Level &L, T *home) {
    if (xref_of<T>(L) == home) return;
    const long n3 = 3 * L.d.N;
    hipLaunchKernelGGL((k_convert<T, T>), dim3(nblk_ew(n3)), dim3(kThreads), 0, h->stream, (const T *)xref_of<T>(L),
                       home, n3);
    xb_of<T>(L) = xref_of<T>(L);
    xref_of<T>(L) = home;
}

template <typename T>
void launch_gs_sweep_colours(stzmg_t h, Level &L) {
    const Dims &d = L.d;
    const int ncx = ncol(d.m, d.px), ncy = ncol(d.n, d.py), ncz = ncol(d.o, d.pz);
    for (int cz = 0; cz < ncz; cz++)
        for (int cy = 0; cy < ncy; cy++)
            for (int cx = 0; cx < ncx; cx++) {
                const int cntx = col_count(d.m, cx, d.px), cnty = col_count(d.n, cy, d.py),
                          cntz = col_count(d.o, cz, d.pz);
                if (cntx == 0 || cnty == 0 || cntz == 0) continue;
                dim3 g(nblk((long)cntx * cnty), cntz);
                if (L.mode == STZMG_OP_PLANE)
                    hipLaunchKernelGGL((k_gs<T, 0>), g, dim3(kThreads), 0, h->stream, d, x_of<T>(L),
                                       (const T *)r_of<T>(L), ops_of<T>(L), inv_of<T>(L), cx, cy, cz,
                                       cntx, cnty);
                else
                    hipLaunchKernelGGL((k_gs<T, 1>), g, dim3(kThreads), 0, h->stream, d, x_of<T>(L),
                                       (const T *)r_of<T>(L), ops_of<T>(L), inv_of<T>(L), cx, cy, cz,
                                       cntx, cnty);
            }
}

// Level l (coarse) receives the restricted residual of level l-1.
template <typename T>
void launch_restrict(stzmg_t h, int l) {
    Level &F = h->lv[l - 1], &C = h->lv[l];
    launch_resid<T, false>(h, F, x_of<T>(F), r_of<T>(F), t_of<T>(F));
    if (F.d.px && F.d.py && C.d.m % kRestrictTX == 0 && C.d.n % kRestrictTY == 0) {
        // LDS-staged tiles; coarse planes per workgroup: as many as keep
        // about 1024 workgroups
        const long tiles = (long)(C.d.m / kRestrictTX) * (C.d.n / kRestrictTY);
        int kc = 1;
        while (kc < 16 && tiles * ((C.d.o + 2 * kc - 1) / (2 * kc)) >= 1024) kc *= 2;
        if (h->restrict_planes > 0) kc = h->restrict_planes;
        dim3 g(C.d.m / kRestrictTX, C.d.n / kRestrictTY, (C.d.o + kc - 1) / kc);
        hipLaunchKernelGGL((k_restrict_tile<T>), g, dim3(kThreads), 0, h->stream, F.d, C.d, (const T *)t_of<T>(F),
                           r_of<T>(C), x_of<T>(C), kc);
        return;
    }
    // coarse planes per thread: as many as keep about 1024 workgroups
    const unsigned bx = nblk((long)C.d.m * C.d.n);
    int kc = 1;
    while (kc < 16 && (long)bx * ((C.d.o + 2 * kc - 1) / (2 * kc)) >= 1024) kc *= 2;
    if (h->restrict_planes > 0) kc = h->restrict_planes;
    dim3 g(bx, (C.d.o + kc - 1) / kc);
    hipLaunchKernelGGL((k_restrict_march<T>), g, dim3(kThreads), 0, h->stream, F.d, C.d,
                       (const T *)t_of<T>(F), r_of<T>(C), x_of<T>(C), kc);
}

template <typename T>
void launch_interp(stzmg_t h, int l) {
    Level &F = h->lv[l - 1], &C = h->lv[l];
    dim3 g(nblk((long)C.d.m * C.d.n), C.d.o);
    hipLaunchKernelGGL((k_interp_cells<T>), g, dim3(kThreads), 0, h->stream, F.d, C.d,
                       (const T *)x_of<T>(C), x_of<T>(F));
}

int sweeps_down(stzmg_t h, int l) { return h->desc.v_down + h->desc.sweep_growth * l; }
int sweeps_up(stzmg_t h, int l) { return h->desc.v_up + h->desc.sweep_growth * l; }

// First level handled by the single-workgroup coarse kernel (lv.size() when
// that kernel is switched off).
int first_coarse_level(stzmg_t h) {
    const int L = (int)h->lv.size();
    if (h->coarse_max <= 0) return L;
    int l = 0;
    while (l < L && h->lv[l].d.N > h->coarse_max) l++;
    if (L - l > kMaxCoarse) l = L - kMaxCoarse;
    // the staged per-plane tables must fit in LDS
    while (l < L && h->lv[l].mode == STZMG_OP_PLANE && coarse_lds_bytes(h->lv[l].d.o) > 64 * 1024) l++;
    return l;
}

// Dynamic LDS limit of the coarse kernel (allowed once per process by
// coarse_kernel_attrs, outside any stream capture).
constexpr size_t kCoarseLdsMax = 128 * 1024;

int coarse_kernel_attrs(stzmg_t h) {
    static hipError_t done = [] {
        hipError_t e = hipSuccess;
        const void *fns[] = {(const void *)k_coarse_cycle<float, 0>, (const void *)k_coarse_cycle<float, 1>,
                             (const void *)k_coarse_cycle<double, 0>, (const void *)k_coarse_cycle<double, 1>,
                             (const void *)k_bottom_map<0>, (const void *)k_bottom_map<1>};
        for (const void *f : fns) {
            const hipError_t e1 = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kCoarseLdsMax);
            if (e1 != hipSuccess) e = e1;
        }
        return e;
    }();
    HIPCHK(h, done);
    return 0;
}

template <typename T>
void launch_coarse(stzmg_t h, int first) {
    const int L = (int)h->lv.size();
    CoarseArgs<T> a{};
    a.nl = L - first;
    a.bottom = h->desc.bottom_sweeps;
    for (int i = 0; i < a.nl; i++) {
        Level &V = h->lv[first + i];
        CoarseLevel<T> &c = a.L[i];
        c.d = V.d;
        c.x = x_of<T>(V);
        c.r = r_of<T>(V);
        c.t = t_of<T>(V);
        c.ops = ops_of<T>(V);
        c.inv = inv_of<T>(V);
        c.down = sweeps_down(h, first + i);
        c.up = sweeps_up(h, first + i);
        const bool fp64_bottom = sizeof(T) == 4 && L > 1 && i == a.nl - 1;
        c.x64 = fp64_bottom ? V.x64 : nullptr;
        c.r64 = fp64_bottom ? V.r64 : nullptr;
        c.ops64 = V.ops64;
        c.inv64 = V.inv64;
    }
    const bool plane = h->lv[first].mode == STZMG_OP_PLANE;
    int omax = 0;
    if (plane)
        for (int i = 0; i < a.nl; i++) omax = a.L[i].d.o > omax ? a.L[i].d.o : omax;
    size_t bytes = coarse_lds_bytes(omax);
    const size_t fields = bottom_lds_bytes(a.L[a.nl - 1].d.N);
    a.table_vals = -1;
    if (bytes + fields <= kCoarseLdsMax) {
        a.table_vals = (int)(bytes / sizeof(double));
        bytes += fields;
    }
    if (plane)
        hipLaunchKernelGGL((k_coarse_cycle<T, 0>), dim3(1), dim3(kCoarseThreads), bytes, h->stream, a);
    else
        hipLaunchKernelGGL((k_coarse_cycle<T, 1>), dim3(1), dim3(kCoarseThreads), bytes, h->stream, a);
}

// One V-cycle (multigrid::v_cycle, multigrid.cc:253-297) in type T.  In FP32
// the coarsest level is solved in FP64: r32 -> r64, bottom sweeps, x64 -> x32.
// Levels from first_coarse_level() down run in one workgroup (mg_coarse.hpp).
// Whether a V-cycle's first operation on the fine level is a sweep that can
// start from x = 0 without the zeros in memory (k_gs_march): the fine level
// is smoothed on the way down, at least once, by the fused sweep.
bool fine_zero_start(stzmg_t h) {
    const int L = (int)h->lv.size();
    const int Lc = first_coarse_level(h);
    const int top = Lc < L - 1 ? Lc : L - 1;
    return top > 0 && sweeps_down(h, 0) > 0 && zero_start_ok(h->lv[0]);
}

// zero_top: the fine level enters with x = 0, which need not be in memory
// where fine_zero_start.  Every coarser level enters with x = 0 from the
// restriction, so the first down-sweep of every level starts from zero.
template <typename T>
void enqueue_vcycle(stzmg_t h, bool zero_top = false) {
    const int L = (int)h->lv.size();
    const int Lc = first_coarse_level(h);
    const int top = Lc < L - 1 ? Lc : L - 1;
    T *home[16];
    for (int i = 0; i < L; i++) home[i] = xref_of<T>(h->lv[i]);
    double *home_b64 = h->lv[L - 1].x64;
    for (int i = 0; i < top; i++) {
        for (int q = 0; q < sweeps_down(h, i); q++)
            launch_gs_sweep<T>(h, h->lv[i],
                               q == 0 && (i > 0 ? zero_start_ok(h->lv[i]) : zero_top && fine_zero_start(h)));
        launch_restrict<T>(h, i + 1);
    }
    if (h->qmap_ok && L > 1 && Lc >= L - 1) {
        // the coarsest level, entered with x = 0: x = Q r
        Level &B = h->lv[L - 1];
        const int n3 = (int)(3 * B.d.N);
        hipLaunchKernelGGL((k_bottom_apply<T>), dim3((n3 + kApplyThreads / 64 - 1) / (kApplyThreads / 64)),
                           dim3(kApplyThreads), 0, h->stream, (const double *)h->qmap, (const T *)r_of<T>(B),
                           x_of<T>(B), n3);
    } else if (Lc < L) {
        launch_coarse<T>(h, Lc);
    } else {
        Level &B = h->lv[L - 1];
        const long n3 = 3 * B.d.N;
        if (sizeof(T) == 4 && L > 1) {
            hipLaunchKernelGGL((k_convert<float, double>), dim3(nblk_ew(n3)), dim3(kThreads), 0, h->stream,
                               (const float *)B.r32, B.r64, n3);
            hipLaunchKernelGGL((k_fill<double>), dim3(nblk(n3)), dim3(kThreads), 0, h->stream, B.x64, 0.0, n3);
            for (int q = 0; q < h->desc.bottom_sweeps; q++) launch_gs_sweep<double>(h, B);
            hipLaunchKernelGGL((k_convert<double, float>), dim3(nblk_ew(n3)), dim3(kThreads), 0, h->stream,
                               (const double *)B.x64, B.x32, n3);
        } else {
            for (int q = 0; q < h->desc.bottom_sweeps; q++) launch_gs_sweep<T>(h, B);
        }
    }
    for (int i = top; i > 0; i--) {
        launch_interp<T>(h, i);
        for (int q = 0; q < sweeps_up(h, i - 1); q++) launch_gs_sweep<T>(h, h->lv[i - 1]);
    }
    for (int i = 0; i < L; i++) restore_level<T>(h, h->lv[i], home[i]);
    if (sizeof(T) == 4) restore_level<double>(h, h->lv[L - 1], home_b64);
}

int ensure_part(stzmg_t h, long nb) {
    if (nb <= h->part_cap) return 0;
    hipFree(h->part);
    h->part = nullptr;
    HIPCHK(h, alloc(&h->part, 2 * nb));
    h->part_cap = nb;
    return 0;
}

// Residual norms of level l in FP64: out[0] = sum |r|^4, out[1] = sum |r|^2.
// The residual is left in t64 of that level.
int norms64(stzmg_t h, int l, double out[2]) {
    Level &L = h->lv[l];
    const long nb = resid_parts<double>(h, L);
    if (ensure_part(h, nb)) return STZMG_ERR_NOMEM;
    launch_resid<double, true>(h, L, L.x64, L.r64, L.t64);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, h->stream, (const double *)h->part, nb,
                       h->dres);
    LAUNCHCHK(h);
    HIPCHK(h, hipMemcpyAsync(out, h->dres, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

// The same norms, enqueued: the reduced pair is copied to the pinned h->hres
// and h->evres recorded after the copy; the caller may enqueue more work
// and then wait on the event.
int norms64_enqueue(stzmg_t h, int l) {
    Level &L = h->lv[l];
    const long nb = resid_parts<double>(h, L);
    if (ensure_part(h, nb)) return STZMG_ERR_NOMEM;
    if (!h->hres) HIPCHK(h, hipHostMalloc((void **)&h->hres, 2 * sizeof(double), hipHostMallocDefault));
    if (!h->evres) HIPCHK(h, hipEventCreateWithFlags(&h->evres, hipEventDisableTiming));
    launch_resid<double, true>(h, L, L.x64, L.r64, L.t64);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, h->stream, (const double *)h->part, nb,
                       h->dres);
    LAUNCHCHK(h);
    HIPCHK(h, hipMemcpyAsync(h->hres, h->dres, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipEventRecord(h->evres, h->stream));
    return 0;
}

// Norms of r - A x on level l for explicit fields x, r (the residual left in
// t64 of that level).
int norms64_of(stzmg_t h, int l, const double *x, const double *r, double out[2]) {
    Level &L = h->lv[l];
    const long nb = resid_parts<double>(h, L);
    if (ensure_part(h, nb)) return STZMG_ERR_NOMEM;
    launch_resid<double, true>(h, L, x, r, L.t64);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, h->stream, (const double *)h->part, nb,
                       h->dres);
    LAUNCHCHK(h);
    HIPCHK(h, hipMemcpyAsync(out, h->dres, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

// Workspace of the FP64 GMRES solve: m FP64 preconditioned vectors and the
// solution and right-hand side while the level-0 fields serve the
// preconditioner.
int ensure_krylov64(stzmg_t h, int m, long n3) {
    if (h->kcapz64 < m) {
        hipFree(h->kZ64); hipFree(h->kX64); hipFree(h->kB64);
        h->kZ64 = h->kX64 = h->kB64 = nullptr;
        h->kcapz64 = 0;
        HIPCHK(h, alloc(&h->kZ64, (long)m * n3));
        HIPCHK(h, alloc(&h->kX64, n3));
        HIPCHK(h, alloc(&h->kB64, n3));
        h->kcapz64 = m;
    }
    return 0;
}

// Workspace of the GMRES solves for restart length m: the FP32
// preconditioned vectors and the reduction scratch, plus the FP64 Krylov
// basis (flexible GMRES) or the FP32 one (GMRES-IR).
int ensure_krylov(stzmg_t h, int m, long n3, bool basis64) {
    if (h->kcap < m) {
        hipFree(h->kZ); hipFree(h->kpart); hipFree(h->kdev);
        h->kZ = nullptr; h->kpart = nullptr; h->kdev = nullptr;
        h->kcap = 0;
        HIPCHK(h, alloc(&h->kZ, (long)m * n3));
        HIPCHK(h, alloc(&h->kpart, (long)kDotBlocks * (kKrylovMax + 1)));
        HIPCHK(h, alloc(&h->kdev, 2L * (kKrylovMax + 1)));
        h->kcap = m;
    }
    if (basis64 && h->kcap64 < m) {
        hipFree(h->kV); hipFree(h->ktmp);
        h->kV = nullptr; h->ktmp = nullptr;
        h->kcap64 = 0;
        HIPCHK(h, alloc(&h->kV, (long)(m + 1) * n3));
        HIPCHK(h, alloc(&h->ktmp, n3));
        h->kcap64 = m;
    }
    if (!basis64 && h->kcap32 < m) {
        hipFree(h->kV32);
        h->kV32 = nullptr;
        h->kcap32 = 0;
        HIPCHK(h, alloc(&h->kV32, (long)(m + 1) * n3));
        h->kcap32 = m;
    }
    return 0;
}

int check_level(stzmg_t h, int l) {
    if (!h) return STZMG_ERR_ARG;
    if (l < 0 || l >= (int)h->lv.size()) return fail(h, STZMG_ERR_ARG, "level %d out of range", l);
    return 0;
}

int require_ops(stzmg_t h) {
    if (!h->ops_set) return fail(h, STZMG_ERR_STATE, "operator not set");
    return 0;
}

// Q of the coarsest level for the current operator (k_bottom_map): only for
// a hierarchy of several levels whose coarsest level fits one workgroup.
// Called before the first V-cycle after the operator changed; a level that
// does not qualify leaves qmap_ok false and is checked again (no GPU work).
int build_bottom_map(stzmg_t h) {
    h->qmap_ok = false;
    if (!h->bottom_map || h->lv.size() < 2) return 0;
    Level &B = h->lv.back();
    if (B.d.N > kCoarseMaxPoints) return 0;
    const bool plane = B.mode == STZMG_OP_PLANE;
    const size_t tab = plane ? coarse_lds_bytes(B.d.o) : 0;
    const size_t bytes = tab + bottom_lds_bytes(B.d.N);
    if (bytes > kCoarseLdsMax) return 0;
    const long n3 = 3 * B.d.N;
    if (h->qmap_n != n3) {
        hipFree(h->qmap);
        h->qmap = nullptr;
        h->qmap_n = 0;
        HIPCHK(h, alloc(&h->qmap, n3 * n3));
        h->qmap_n = n3;
    }
    const int tv = (int)(tab / sizeof(double));
    if (plane)
        hipLaunchKernelGGL(k_bottom_map<0>, dim3((unsigned)n3), dim3(kCoarseThreads), bytes, h->stream, B.d,
                           (const double *)B.ops64, (const double *)B.inv64, h->desc.bottom_sweeps, tv, h->qmap);
    else
        hipLaunchKernelGGL(k_bottom_map<1>, dim3((unsigned)n3), dim3(kCoarseThreads), bytes, h->stream, B.d,
                           (const double *)B.ops64, (const double *)B.inv64, h->desc.bottom_sweeps, tv, h->qmap);
    LAUNCHCHK(h);
    h->qmap_ok = true;
    return 0;
}

// Packed coefficient tables of the levels that sweep with k_gs_march (the
// large tile class): the nonzero entries of each plane's stencil in the
// kernel's load order, then the inverse centre block.
constexpr int kPackMax = kBlk + 9;
constexpr int kPats = 2 * kMarchSym;  // zero patterns, plain and centrosymmetric

int pack_march(stzmg_t h) {
    if (!h->pkmap) {
        // the sweep's tables end with the inverse centre block (source 243 +
        // e), the residual's with the centre block itself (9 * 13 + e)
        std::vector<int> src((size_t)kPats * kPackMax, 0), srcr(src.size(), 0);
        for (int pat = 0; pat < kPats; pat++)
            for (int j = 0; j < march_stride(pat); j++) {
                const int q = march_src(pat, j);
                src[(size_t)pat * kPackMax + j] = q;
                srcr[(size_t)pat * kPackMax + j] = q < kBlk ? q : 9 * 13 + (q - kBlk);
            }
        HIPCHK(h, alloc(&h->pkmap, (long)src.size()));
        HIPCHK(h, hipMemcpy(h->pkmap, src.data(), sizeof(int) * src.size(), hipMemcpyHostToDevice));
        HIPCHK(h, alloc(&h->pkrmap, (long)srcr.size()));
        HIPCHK(h, hipMemcpy(h->pkrmap, srcr.data(), sizeof(int) * srcr.size(), hipMemcpyHostToDevice));
    }
    for (auto &L : h->lv) {
        if (L.mode != STZMG_OP_PLANE) continue;
        const int key = L.pat + (L.sym ? kMarchSym : 0);
        const int stride = march_stride(key);
        const long n = L.ncls * stride;
        // the marching residual's tables, wherever it may run (the tiled
        // residual can be switched on after the operator is set)
        if (tiled_ok(L) || tiled_small_ok(L)) {
            const int *mapr = h->pkrmap + (long)key * kPackMax;
            if (!L.pkr32) HIPCHK(h, alloc(&L.pkr32, L.ncls * kPackMax));
            if (!L.pkr64) HIPCHK(h, alloc(&L.pkr64, L.ncls * kPackMax));
            hipLaunchKernelGGL(k_pack_march<float>, dim3(nblk(n)), dim3(kThreads), 0, h->stream,
                               (const float *)L.ops32, (const float *)L.inv32, mapr, stride, L.ncls, L.pkr32);
            hipLaunchKernelGGL(k_pack_march<double>, dim3(nblk(n)), dim3(kThreads), 0, h->stream,
                               (const double *)L.ops64, (const double *)L.inv64, mapr, stride, L.ncls, L.pkr64);
            LAUNCHCHK(h);
        }
        if (!STZ_MARCH || !L.fused) continue;
        const int *map = h->pkmap + (long)key * kPackMax;
        // every fused tile class sweeps with k_gs_march, which reads the
        // packed table (a level without one would hand it a null pointer)
        if (L.ftile >= 1) {
            if (!L.pk32) HIPCHK(h, alloc(&L.pk32, L.ncls * kPackMax));
            hipLaunchKernelGGL(k_pack_march<float>, dim3(nblk(n)), dim3(kThreads), 0, h->stream,
                               (const float *)L.ops32, (const float *)L.inv32, map, stride, L.ncls, L.pk32);
        }
        if (L.ftile64 >= 1) {
            if (!L.pk64) HIPCHK(h, alloc(&L.pk64, L.ncls * kPackMax));
            hipLaunchKernelGGL(k_pack_march<double>, dim3(nblk(n)), dim3(kThreads), 0, h->stream,
                               (const double *)L.ops64, (const double *)L.inv64, map, stride, L.ncls, L.pk64);
        }
        LAUNCHCHK(h);
        if ((L.ftile >= 1 && !L.pk32) || (L.ftile64 >= 1 && !L.pk64))
            return fail(h, STZMG_ERR_STATE, "packed coefficient table missing on a fused level");
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

// Forget the captured V-cycles (their kernels or arguments changed).
void drop_graphs(stzmg_t h) {
    for (hipGraphExec_t *g : {&h->g64, &h->g32, &h->g64z, &h->g32z})
        if (*g) {
            hipGraphExecDestroy(*g);
            *g = nullptr;
        }
}

// Galerkin hierarchy and inverse tables after the fine operator changed.
int build_hierarchy_ops(stzmg_t h) {
    if (int rc = coarse_kernel_attrs(h)) return rc;
    HIPCHK(h, hipMemsetAsync(h->dflag, 0, sizeof(int), h->stream));
    for (size_t l = 0; l < h->lv.size(); l++) {
        Level &L = h->lv[l];
        if (l > 0) {
            Level &F = h->lv[l - 1];
            const long nthr = L.ncls * kBlk;
            if (F.mode == STZMG_OP_PLANE && L.mode == STZMG_OP_PLANE) {
                hipLaunchKernelGGL((k_rat<0, 0>), dim3(nblk(nthr, 64)), dim3(64), 0, h->stream, F.d, L.d,
                                   (const double *)F.ops64, L.ops64, L.ncls);
                if (h->centro_galerkin)
                    hipLaunchKernelGGL(k_centro_sym, dim3(1), dim3(kSymThreads), 0, h->stream, L.ops64, L.ncls,
                                       kCentroTol);
            } else if (F.mode == STZMG_OP_POINT && L.mode == STZMG_OP_POINT)
                hipLaunchKernelGGL((k_rat<1, 1>), dim3(nblk(nthr, 64)), dim3(64), 0, h->stream, F.d, L.d,
                                   (const double *)F.ops64, L.ops64, L.ncls);
            else
                return fail(h, STZMG_ERR_STATE, "mixed operator modes");
            LAUNCHCHK(h);
        }
        const long ne = L.ncls * kBlk;
        hipLaunchKernelGGL(k_cast_d2f, dim3(nblk(ne)), dim3(kThreads), 0, h->stream, (const double *)L.ops64,
                           L.ops32, ne);
        hipLaunchKernelGGL(k_inverse, dim3(nblk(L.ncls, 64)), dim3(64), 0, h->stream, (const double *)L.ops64,
                           L.ncls, L.inv64, L.inv32, h->dflag);
        LAUNCHCHK(h);
    }
    // zero pattern and centrosymmetry of the per-plane levels, read back with
    // the singular-block flag in one synchronization
    const size_t ncfg = h->lv.size() * kCfgWords;
    if (!h->dcfg) HIPCHK(h, alloc(&h->dcfg, (long)ncfg));
    for (size_t l = 0; l < h->lv.size(); l++)
        if (h->lv[l].mode == STZMG_OP_PLANE)
            hipLaunchKernelGGL(k_level_config, dim3(1), dim3(kThreads), 0, h->stream, (const double *)h->lv[l].ops64,
                               h->lv[l].ncls, h->dcfg + l * kCfgWords);
    LAUNCHCHK(h);
    std::vector<unsigned> cfg(ncfg, 0);
    HIPCHK(h, hipMemcpyAsync(cfg.data(), h->dcfg, sizeof(unsigned) * ncfg, hipMemcpyDeviceToHost, h->stream));
    int flag = 0;
    HIPCHK(h, hipMemcpyAsync(&flag, h->dflag, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    // The captured V-cycles hold each level's kernel instances and buffer
    // pointers, not its tables: they stay valid while every level keeps its
    // kernel configuration (an operator that changes every step, as in the
    // transformed step, then keeps its graphs).
    auto config = [](const Level &L) {
        return std::make_tuple(L.fused, L.tiled, L.tiled_small, L.ftile, L.ftile64, L.pat, L.sym);
    };
    std::vector<decltype(config(h->lv[0]))> before;
    for (auto &L : h->lv) before.push_back(config(L));
    // Zero pattern of every per-plane level: the tightest compiled pattern
    // that holds all nonzero entries selects the fused smoother and tiled
    // residual instances.
    for (size_t l = 0; l < h->lv.size(); l++) {
        Level &L = h->lv[l];
        L.fused = false;
        L.tiled = false;
        L.tiled_small = false;
        if (L.mode != STZMG_OP_PLANE) continue;
        const unsigned *used = cfg.data() + l * kCfgWords;  // per block: entries nonzero on some plane
        int pat = PAT_DENSE;
        for (int cand : {PAT_SHEAR, PAT_TRANS_XZ, PAT_NOCORNER}) {
            bool fits = true;
            for (int b = 0; b < 27; b++)
                if (b != 13 && (used[b] & ~pat_mask(cand, b))) fits = false;
            if (fits) {
                pat = cand;
                break;
            }
        }
        L.pat = pat;
        // centrosymmetric: block(-o) equal to block(o) in every plane
        L.sym = used[27] != 0;
        L.ftile = fused_tile<float>(L);
        L.ftile64 = fused_tile<double>(L);
        L.fused = h->fused_smoother && L.ftile && L.ftile64 && L.x64b;
        L.tiled = h->tiled_resid && tiled_ok(L);
        L.tiled_small = h->tiled_resid && tiled_small_ok(L);
    }
    if (int rc = pack_march(h)) return rc;
    bool same = h->ops_set;
    for (size_t l = 0; l < h->lv.size() && same; l++) same = config(h->lv[l]) == before[l];
    if (!same) drop_graphs(h);
    h->ops_set = true;
    h->qmap_ok = false;  // tabulated at the first V-cycle with this operator
    if (flag) return fail(h, STZMG_ERR_SINGULAR, "singular or non-finite centre block on some level");
    return 0;
}

template <typename T>
int capture_vcycle(stzmg_t h, hipGraphExec_t *exec, bool zero_top) {
    if (*exec) return 0;
    hipGraph_t g = nullptr;
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    enqueue_vcycle<T>(h, zero_top);
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (e != hipSuccess) return fail(h, STZMG_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return fail(h, STZMG_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
    return 0;
}

// n V-cycles; zero_first: the first starts from x = 0 on the fine level
// (which then need not hold the zeros where zero_start_ok).
template <typename T>
int run_vcycles(stzmg_t h, int n, bool graph, bool zero_first = false) {
    // Q of the current operator, before any capture that would use it
    if (h->bottom_map && !h->qmap_ok)
        if (int rc = build_bottom_map(h)) return rc;
    if (graph) {
        hipGraphExec_t *ex = sizeof(T) == 8 ? &h->g64 : &h->g32;
        hipGraphExec_t *exz = sizeof(T) == 8 ? &h->g64z : &h->g32z;
        for (int q = 0; q < n; q++) {
            hipGraphExec_t *e = q == 0 && zero_first ? exz : ex;
            if (int rc = capture_vcycle<T>(h, e, e == exz)) return rc;
            HIPCHK(h, hipGraphLaunch(*e, h->stream));
        }
    } else {
        for (int q = 0; q < n; q++) enqueue_vcycle<T>(h, q == 0 && zero_first);
        LAUNCHCHK(h);
    }
    return 0;
}

bool size_supported(int n, int prd) { return prd ? (n % 2 == 0) : (n % 2 == 1); }

}  // namespace

namespace stzmg_internal {
double *x0(stzmg_t h) { return h->lv[0].x64; }
double *r0(stzmg_t h) { return h->lv[0].r64; }
}  // namespace stzmg_internal

// ---------------------------------------------------------------------------
extern "C" {

int stzmg_desc_default(stzmg_desc *d, int m, int n, int o, int px, int py, int pz) {
    if (!d) return STZMG_ERR_ARG;
    d->m = m; d->n = n; d->o = o;
    d->prd[0] = px; d->prd[1] = py; d->prd[2] = pz;
    d->coarsest_total = 1024;
    d->v_down = 2; d->v_up = 2;
    d->bottom_sweeps = 24;
    d->sweep_growth = 1;
    d->op_mode = (px && py) ? STZMG_OP_PLANE : STZMG_OP_POINT;
    return 0;
}

static thread_local std::string g_create_err;

int stzmg_create(const stzmg_desc *d, stzmg_t *out) {
    if (!d || !out) return STZMG_ERR_ARG;
    *out = nullptr;
    stzmg_t h = new stzmg_s;
    h->desc = *d;
    auto bail = [&](int code) {
        g_create_err = h->err;
        for (auto &L : h->lv) free_level(L);
        if (h->stream) hipStreamDestroy(h->stream);
        delete h;
        return code;
    };
    if (d->m < 3 || d->n < 3 || d->o < 3 || d->coarsest_total < 1 || d->v_down < 0 || d->v_up < 0 ||
        d->bottom_sweeps < 0 || d->sweep_growth < 0)
        return bail(fail(h, STZMG_ERR_ARG, "invalid descriptor"));
    if (d->op_mode == STZMG_OP_PLANE && !(d->prd[0] && d->prd[1]))
        return bail(fail(h, STZMG_ERR_UNSUPPORTED, "PLANE operators need periodic x and y"));
    // Level sizes: m' = (m + (prd?1:2)) >> 1 while m*n*o > coarsest_total
    // (region.cc:856, multigrid.cc:157).
    std::vector<Dims> dims;
    Dims cur{d->m, d->n, d->o, d->prd[0] != 0, d->prd[1] != 0, d->prd[2] != 0, 0};
    for (;;) {
        cur.N = (long)cur.m * cur.n * cur.o;
        dims.push_back(cur);
        if (cur.N <= d->coarsest_total) break;
        if ((int)dims.size() == 16) return bail(fail(h, STZMG_ERR_UNSUPPORTED, "more than 16 levels"));
        // Coarsened dimensions must be periodic-even or non-periodic-odd: for
        // other sizes the reference's transfer routines deviate from P^T A P
        // and from R = P^T/8 (odd periodic wrap in the RAT; the even
        // non-periodic restriction of region.cc:2230-2240).
        if (!size_supported(cur.m, cur.px) || !size_supported(cur.n, cur.py) ||
            !size_supported(cur.o, cur.pz))
            return bail(fail(h, STZMG_ERR_UNSUPPORTED,
                             "level %zu size %dx%dx%d: coarsened dimensions must be even when periodic "
                             "and odd when walled",
                             dims.size() - 1, cur.m, cur.n, cur.o));
        Dims nx = cur;
        nx.m = coarse_size(cur.m, cur.px);
        nx.n = coarse_size(cur.n, cur.py);
        nx.o = coarse_size(cur.o, cur.pz);
        if ((nx.px && nx.m < 3) || (nx.py && nx.n < 3) || (nx.pz && nx.o < 3) || nx.m < 2 || nx.n < 2 ||
            nx.o < 2)
            return bail(fail(h, STZMG_ERR_UNSUPPORTED, "grid too small to coarsen to %d points",
                             (int)d->coarsest_total));
        cur = nx;
    }
    if (hipStreamCreate(&h->stream) != hipSuccess) return bail(fail(h, STZMG_ERR_HIP, "hipStreamCreate failed"));
    {
        int dev = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) h->scratch = scratch_on_device(dev, &e);
        if (e != hipSuccess)
            return bail(fail(h, STZMG_ERR_HIP, "scratch address lookup failed: %s", hipGetErrorString(e)));
    }
    h->lv.resize(dims.size());
    for (size_t l = 0; l < dims.size(); l++) {
        Level &L = h->lv[l];
        L.d = dims[l];
        L.mode = d->op_mode;
        L.ncls = L.mode == STZMG_OP_PLANE ? L.d.o : L.d.N;
        const long n3 = 3 * L.d.N;
        if (alloc(&L.ops64, L.ncls * kBlk) || alloc(&L.ops32, L.ncls * kBlk) || alloc(&L.inv64, L.ncls * 9) ||
            alloc(&L.inv32, L.ncls * 9) || alloc(&L.x64, n3) || alloc(&L.r64, n3) || alloc(&L.t64, n3) ||
            alloc(&L.x32, n3) || alloc(&L.r32, n3) || alloc(&L.t32, n3))
            return bail(fail(h, STZMG_ERR_NOMEM, "device allocation failed at level %zu", l));
        if (fused_tile<float>(L) && fused_tile<double>(L) && (alloc(&L.x64b, n3) || alloc(&L.x32b, n3)))
            return bail(fail(h, STZMG_ERR_NOMEM, "device allocation failed at level %zu", l));
        hipMemsetAsync(L.ops64, 0, sizeof(double) * L.ncls * kBlk, h->stream);
        hipMemsetAsync(L.x64, 0, sizeof(double) * n3, h->stream);
        hipMemsetAsync(L.r64, 0, sizeof(double) * n3, h->stream);
        hipMemsetAsync(L.x32, 0, sizeof(float) * n3, h->stream);
        hipMemsetAsync(L.r32, 0, sizeof(float) * n3, h->stream);
    }
    if (alloc(&h->dres, 2) || alloc(&h->dflag, 1)) return bail(fail(h, STZMG_ERR_NOMEM, "allocation failed"));
    if (hipStreamSynchronize(h->stream) != hipSuccess) return bail(fail(h, STZMG_ERR_HIP, "init sync failed"));
    *out = h;
    return 0;
}

int stzmg_destroy(stzmg_t h) {
    if (!h) return STZMG_ERR_ARG;
    hipStreamSynchronize(h->stream);
    drop_graphs(h);
    for (auto &L : h->lv) free_level(L);
    hipFree(h->part);
    hipFree(h->dres);
    hipFree(h->dflag);
    hipFree(h->dcfg);
    if (h->hres) hipHostFree(h->hres);
    if (h->evres) hipEventDestroy(h->evres);
    hipFree(h->kV); hipFree(h->kZ); hipFree(h->ktmp); hipFree(h->kpart); hipFree(h->kdev);
    hipFree(h->kV32);
    hipFree(h->kZ64); hipFree(h->kX64); hipFree(h->kB64);
    hipFree(h->qmap);
    hipFree(h->pkmap);
    hipFree(h->pkrmap);
    hipStreamDestroy(h->stream);
    delete h;
    return 0;
}

const char *stzmg_last_error(stzmg_t h) { return h ? h->err.c_str() : g_create_err.c_str(); }

int stzmg_nlevels(stzmg_t h) { return h ? (int)h->lv.size() : STZMG_ERR_ARG; }

int stzmg_level_dims(stzmg_t h, int level, int dims[3]) {
    if (int rc = check_level(h, level)) return rc;
    dims[0] = h->lv[level].d.m;
    dims[1] = h->lv[level].d.n;
    dims[2] = h->lv[level].d.o;
    return 0;
}

void *stzmg_stream(stzmg_t h) { return h ? (void *)h->stream : nullptr; }

int stzmg_set_option(stzmg_t h, int option, int value) {
    if (!h) return STZMG_ERR_ARG;
    if (option == STZMG_OPT_COARSE_KERNEL) {
        h->coarse_max = value == 1 ? kCoarseMaxPoints : (value < 0 ? 0 : value);
    } else if (option == STZMG_OPT_FUSED_SMOOTHER) {
        h->fused_smoother = value != 0;
        if (h->ops_set) {
            for (auto &L : h->lv) {
                restore_level<double>(h, L, L.x64);  // pointers are home between calls
                if (!h->fused_smoother) L.fused = false;
            }
            if (h->fused_smoother) {
                int rc = build_hierarchy_ops(h);
                if (rc) return rc;
            }
        }
    } else if (option == STZMG_OPT_CENTRO_GALERKIN) {
        h->centro_galerkin = value != 0;
        if (h->ops_set) {
            for (auto &L : h->lv) restore_level<double>(h, L, L.x64);  // pointers are home between calls
            if (int rc = build_hierarchy_ops(h)) return rc;
        }
    } else if (option == STZMG_OPT_TILED_RESIDUAL) {
        h->tiled_resid = value != 0;
        for (auto &L : h->lv) {
            L.tiled = h->ops_set && h->tiled_resid && tiled_ok(L);
            L.tiled_small = h->ops_set && h->tiled_resid && tiled_small_ok(L);
        }
    } else if (option == STZMG_OPT_BOTTOM_MAP) {
        h->bottom_map = value != 0;
        h->qmap_ok = false;  // (re)built at the next V-cycle when on
    } else if (option == STZMG_OPT_CHUNK_CAP) {
        if (value < 0) return fail(h, STZMG_ERR_ARG, "chunk cap must be >= 0");
        h->chunk_cap = value;
    } else if (option == STZMG_OPT_RESTRICT_PLANES) {
        if (value < 0) return fail(h, STZMG_ERR_ARG, "restriction planes must be >= 0");
        h->restrict_planes = value;
    } else {
        return fail(h, STZMG_ERR_ARG, "unknown option %d", option);
    }
    drop_graphs(h);
    return 0;
}

int stzmg_set_operator_planes(stzmg_t h, const double *planes) {
    if (!h || !planes) return STZMG_ERR_ARG;
    Level &L = h->lv[0];
    if (L.mode != STZMG_OP_PLANE) return fail(h, STZMG_ERR_STATE, "handle holds per-point operators");
    HIPCHK(h, hipMemcpyAsync(L.ops64, planes, sizeof(double) * L.ncls * kBlk, hipMemcpyHostToDevice, h->stream));
    return build_hierarchy_ops(h);
}

int stzmg_set_operator_points(stzmg_t h, const double *points) {
    if (!h || !points) return STZMG_ERR_ARG;
    Level &L = h->lv[0];
    if (L.mode != STZMG_OP_POINT) return fail(h, STZMG_ERR_STATE, "handle holds per-plane operators");
    HIPCHK(h, hipMemcpyAsync(L.ops64, points, sizeof(double) * L.ncls * kBlk, hipMemcpyHostToDevice, h->stream));
    return build_hierarchy_ops(h);
}

int stzmg_set_operator_const(stzmg_t h, const double *blocks, const double *fix) {
    if (!h || !blocks) return STZMG_ERR_ARG;
    Level &L = h->lv[0];
    const Dims &d = L.d;
    const bool walls = !d.px || !d.py || !d.pz;
    if (walls && !fix) return fail(h, STZMG_ERR_ARG, "fix block required with non-periodic faces");
    std::vector<double> tab((size_t)L.ncls * kBlk, 0.0);
    auto put = [&](long c, bool bnd) {
        double *o = tab.data() + c * kBlk;
        if (bnd) {
            for (int e = 0; e < 9; e++) o[9 * 13 + e] = fix[e];
        } else {
            memcpy(o, blocks, sizeof(double) * kBlk);
        }
    };
    if (L.mode == STZMG_OP_PLANE) {
        for (int k = 0; k < d.o; k++) put(k, !d.pz && (k == 0 || k == d.o - 1));
    } else {
        for (int k = 0; k < d.o; k++)
            for (int j = 0; j < d.n; j++)
                for (int i = 0; i < d.m; i++) {
                    const bool bnd = (!d.px && (i == 0 || i == d.m - 1)) || (!d.py && (j == 0 || j == d.n - 1)) ||
                                     (!d.pz && (k == 0 || k == d.o - 1));
                    put(i + (long)d.m * (j + (long)d.n * k), bnd);
                }
    }
    HIPCHK(h, hipMemcpyAsync(L.ops64, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return build_hierarchy_ops(h);
}

int stzmg_get_level_operator(stzmg_t h, int level, double *out, long *ncls) {
    if (int rc = check_level(h, level)) return rc;
    Level &L = h->lv[level];
    if (ncls) *ncls = L.ncls;
    if (out) {
        HIPCHK(h, hipMemcpyAsync(out, L.ops64, sizeof(double) * L.ncls * kBlk, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    return 0;
}

int stzmg_set_level_fields(stzmg_t h, int level, const double *x, const double *r) {
    if (int rc = check_level(h, level)) return rc;
    Level &L = h->lv[level];
    const long N = L.d.N;
    // Stage through the temporary field (AoS upload, SoA conversion).
    for (int q = 0; q < 2; q++) {
        const double *src = q == 0 ? x : r;
        double *dst = q == 0 ? L.x64 : L.r64;
        if (!src) continue;
        HIPCHK(h, hipMemcpyAsync(L.t64, src, sizeof(double) * 3 * N, hipMemcpyHostToDevice, h->stream));
        hipLaunchKernelGGL(k_aos_to_soa, dim3(nblk(N)), dim3(kThreads), 0, h->stream, (const double *)L.t64, dst, N);
        LAUNCHCHK(h);
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

int stzmg_get_level_fields(stzmg_t h, int level, double *x, double *r) {
    if (int rc = check_level(h, level)) return rc;
    Level &L = h->lv[level];
    const long N = L.d.N;
    for (int q = 0; q < 2; q++) {
        double *dst = q == 0 ? x : r;
        const double *src = q == 0 ? L.x64 : L.r64;
        if (!dst) continue;
        hipLaunchKernelGGL(k_soa_to_aos, dim3(nblk(N)), dim3(kThreads), 0, h->stream, src, L.t64, N);
        LAUNCHCHK(h);
        HIPCHK(h, hipMemcpyAsync(dst, L.t64, sizeof(double) * 3 * N, hipMemcpyDeviceToHost, h->stream));
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

int stzmg_set_fields_device(stzmg_t h, const double *x_dev, const double *r_dev) {
    if (!h) return STZMG_ERR_ARG;
    Level &L = h->lv[0];
    const size_t bytes = sizeof(double) * 3 * L.d.N;
    if (x_dev) HIPCHK(h, hipMemcpyAsync(L.x64, x_dev, bytes, hipMemcpyDeviceToDevice, h->stream));
    if (r_dev) HIPCHK(h, hipMemcpyAsync(L.r64, r_dev, bytes, hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

int stzmg_get_x_device(stzmg_t h, double *x_dev) {
    if (!h || !x_dev) return STZMG_ERR_ARG;
    Level &L = h->lv[0];
    HIPCHK(h, hipMemcpyAsync(x_dev, L.x64, sizeof(double) * 3 * L.d.N, hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

int stzmg_gauss_seidel(stzmg_t h, int level, int nsweeps) {
    if (int rc = check_level(h, level)) return rc;
    if (int rc = require_ops(h)) return rc;
    double *home = h->lv[level].x64;
    for (int q = 0; q < nsweeps; q++) launch_gs_sweep<double>(h, h->lv[level]);
    restore_level<double>(h, h->lv[level], home);
    LAUNCHCHK(h);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

namespace {
// Level fields rounded to FP32 (x32 <- x64, r32 <- r64).
void level_to_fp32(stzmg_t h, Level &L) {
    const long n3 = 3 * L.d.N;
    hipLaunchKernelGGL((k_convert<double, float>), dim3(nblk_ew(n3)), dim3(kThreads), 0, h->stream,
                       (const double *)L.x64, L.x32, n3);
    hipLaunchKernelGGL((k_convert<double, float>), dim3(nblk_ew(n3)), dim3(kThreads), 0, h->stream,
                       (const double *)L.r64, L.r32, n3);
}
}  // namespace

int stzmg_gauss_seidel_fp32(stzmg_t h, int level, int nsweeps) {
    if (int rc = check_level(h, level)) return rc;
    if (int rc = require_ops(h)) return rc;
    if (nsweeps < 0) return fail(h, STZMG_ERR_ARG, "nsweeps must be >= 0");
    Level &L = h->lv[level];
    float *home = L.x32;
    level_to_fp32(h, L);
    for (int q = 0; q < nsweeps; q++) launch_gs_sweep<float>(h, L);
    restore_level<float>(h, L, home);
    const long n3 = 3 * L.d.N;
    hipLaunchKernelGGL((k_convert<float, double>), dim3(nblk_ew(n3)), dim3(kThreads), 0, h->stream,
                       (const float *)L.x32, L.x64, n3);
    LAUNCHCHK(h);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

int stzmg_residual_fp32(stzmg_t h, int level, double *out) {
    if (int rc = check_level(h, level)) return rc;
    if (int rc = require_ops(h)) return rc;
    if (!out) return STZMG_ERR_ARG;
    Level &L = h->lv[level];
    const long N = L.d.N;
    level_to_fp32(h, L);
    launch_resid<float, false>(h, L, L.x32, L.r32, L.t32);
    hipLaunchKernelGGL((k_convert<float, double>), dim3(nblk_ew(3 * N)), dim3(kThreads), 0, h->stream,
                       (const float *)L.t32, L.t64, 3 * N);
    LAUNCHCHK(h);
    std::vector<double> soa((size_t)3 * N);
    HIPCHK(h, hipMemcpyAsync(soa.data(), L.t64, sizeof(double) * soa.size(), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (long p = 0; p < N; p++) {
        out[3 * p] = soa[p];
        out[3 * p + 1] = soa[p + N];
        out[3 * p + 2] = soa[p + 2 * N];
    }
    return 0;
}

int stzmg_restriction(stzmg_t h, int level) {
    if (int rc = check_level(h, level)) return rc;
    if (level == 0) return fail(h, STZMG_ERR_ARG, "restriction targets a coarse level (>= 1)");
    if (int rc = require_ops(h)) return rc;
    launch_restrict<double>(h, level);
    LAUNCHCHK(h);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

int stzmg_interpolation(stzmg_t h, int level) {
    if (int rc = check_level(h, level)) return rc;
    if (level == 0) return fail(h, STZMG_ERR_ARG, "interpolation starts from a coarse level (>= 1)");
    launch_interp<double>(h, level);
    LAUNCHCHK(h);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

int stzmg_residual(stzmg_t h, int level, double *out) {
    if (int rc = check_level(h, level)) return rc;
    if (int rc = require_ops(h)) return rc;
    Level &L = h->lv[level];
    launch_resid<double, false>(h, L, L.x64, L.r64, L.t64);
    LAUNCHCHK(h);
    if (out) {
        // t64 holds the residual; convert through x32/r32-free staging: use
        // the FP32 temporary's storage is too small, so convert in place on
        // the host side of a scratch copy.
        std::vector<double> soa((size_t)3 * L.d.N);
        HIPCHK(h, hipMemcpyAsync(soa.data(), L.t64, sizeof(double) * soa.size(), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        const long N = L.d.N;
        for (long p = 0; p < N; p++) {
            out[3 * p] = soa[p];
            out[3 * p + 1] = soa[p + N];
            out[3 * p + 2] = soa[p + 2 * N];
        }
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

int stzmg_norm(stzmg_t h, int level, int kind, double *out) {
    if (int rc = check_level(h, level)) return rc;
    if (int rc = require_ops(h)) return rc;
    if (!out || (kind != STZMG_NORM_QUARTIC && kind != STZMG_NORM_L2SQ)) return STZMG_ERR_ARG;
    double v[2];
    if (int rc = norms64(h, level, v)) return rc;
    *out = kind == STZMG_NORM_QUARTIC ? v[0] : v[1];
    return 0;
}

int stzmg_norm_noalloc(stzmg_t h, int level, int kind, double *out) {
    if (int rc = check_level(h, level)) return rc;
    if (int rc = require_ops(h)) return rc;
    if (!out || (kind != STZMG_NORM_QUARTIC && kind != STZMG_NORM_L2SQ)) return STZMG_ERR_ARG;
    Level &L = h->lv[level];
    const long nb = resid_parts<double>(h, L);
    if (nb > kNoallocParts)
        return fail(h, STZMG_ERR_UNSUPPORTED,
                    "stzmg_norm_noalloc: level %d needs %ld norm partials, the scratch holds %ld "
                    "(use stzmg_norm)",
                    level, nb, kNoallocParts);
    // every workgroup writes its own pair and the reduction overwrites the
    // result, so nothing needs clearing between calls
    double *part = h->scratch, *res = h->scratch + 2 * kNoallocParts;
    launch_resid<double, true>(h, L, L.x64, L.r64, L.t64, part);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, h->stream, (const double *)part, nb, res);
    LAUNCHCHK(h);
    double v[2];
    HIPCHK(h, hipMemcpyAsync(v, res, sizeof(v), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    *out = kind == STZMG_NORM_QUARTIC ? v[0] : v[1];
    return 0;
}

int stzmg_vcycle(stzmg_t h, int n) {
    if (!h) return STZMG_ERR_ARG;
    if (int rc = require_ops(h)) return rc;
    if (int rc = run_vcycles<double>(h, n, false)) return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

int stzmg_vcycle_fp32(stzmg_t h, int n) {
    if (!h) return STZMG_ERR_ARG;
    if (int rc = require_ops(h)) return rc;
    if (int rc = run_vcycles<float>(h, n, false)) return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

void stzmg_solve_opts_default(stzmg_solve_opts *o) {
    if (!o) return;
    o->mode = STZMG_SOLVE_RELRES;
    o->tol = 0;
    o->extra_cycles = 2;
    o->ref_norm_schedule = 0;
    o->rtol = 1e-10;
    o->max_cycles = 1000;
    o->inner_cycles = 3;
    o->use_graph = 1;
    o->restart = 10;
    o->inner_rtol = 0;
}

int stzmg_solve(stzmg_t h, const stzmg_solve_opts *opts, stzmg_solve_stats *st) {
    if (!h || !opts) return STZMG_ERR_ARG;
    if (int rc = require_ops(h)) return rc;
    stzmg_solve_stats s{};
    const bool graph = opts->use_graph != 0;
    auto t0 = std::chrono::steady_clock::now();
    double v[2];
    int rc = 0;
    if (opts->mode == STZMG_SOLVE_REF) {
        // shear_sim_3d::projection_step, src/shear_sim_3d.cpp:555-600.
        int num_iters = 1;
        double err = 1e16;
        if ((rc = norms64(h, 0, v))) return rc;
        s.res0 = v[0];
        s.norms++;
        if (!std::isfinite(v[0])) rc = STZMG_ERR_NONFINITE;
        // When extra cycles follow convergence, the V-cycle after a norm is
        // enqueued before the host waits for that norm: if the loop goes on,
        // it is the next iteration's cycle, if not, the first extra cycle --
        // the same cycles in the same order either way, without the GPU
        // idling while the host reads the norm.
        int ahead = 0;
        while (rc == 0 && err > opts->tol) {
            if (!ahead) {
                if (s.cycles >= opts->max_cycles) { rc = STZMG_ERR_NOT_CONVERGED; break; }
                if ((rc = run_vcycles<double>(h, 1, graph))) return rc;
                s.cycles++;
            }
            ahead = 0;
            const bool eval = opts->ref_norm_schedule == 0 || num_iters >= h->std_iters / 16;
            if (eval) {
                if ((rc = norms64_enqueue(h, 0))) return rc;
                if (opts->extra_cycles > 0 && s.cycles < opts->max_cycles) {
                    if ((rc = run_vcycles<double>(h, 1, graph))) return rc;
                    s.cycles++;
                    ahead = 1;
                }
                HIPCHK(h, hipEventSynchronize(h->evres));
                s.norms++;
                err = h->hres[0];
                if (!std::isfinite(err)) { rc = STZMG_ERR_NONFINITE; break; }
            }
            num_iters++;
        }
        if (num_iters >= h->std_iters / 16) h->std_iters = num_iters * 16;
        h->std_iters--;
        if (rc == 0) {
            if ((rc = run_vcycles<double>(h, opts->extra_cycles - ahead, graph))) return rc;
            s.cycles += opts->extra_cycles - ahead;
        }
        HIPCHK(h, hipStreamSynchronize(h->stream));
        if (rc == 0 || rc == STZMG_ERR_NOT_CONVERGED) {
            int r2 = norms64(h, 0, v);
            if (r2) return r2;
            s.res = v[0];
        }
    } else if (opts->mode == STZMG_SOLVE_RELRES) {
        if ((rc = norms64(h, 0, v))) return rc;
        s.norms++;
        const double r0 = std::sqrt(v[1]);
        s.res0 = s.res = r0;
        const double target = opts->rtol * r0;
        if (!std::isfinite(r0)) rc = STZMG_ERR_NONFINITE;
        while (rc == 0 && s.res > target) {
            if (s.cycles >= opts->max_cycles) { rc = STZMG_ERR_NOT_CONVERGED; break; }
            if ((rc = run_vcycles<double>(h, 1, graph))) return rc;
            s.cycles++;
            if ((rc = norms64(h, 0, v))) return rc;
            s.norms++;
            s.res = std::sqrt(v[1]);
            if (!std::isfinite(s.res)) { rc = STZMG_ERR_NONFINITE; break; }
        }
    } else if (opts->mode == STZMG_SOLVE_FP32_IR) {
        // FP64 defect -> scaled FP32 defect -> FP32 V-cycles from zero ->
        // FP64 update, repeated until the FP64 residual meets rtol.
        Level &L = h->lv[0];
        const long n3 = 3 * L.d.N;
        if ((rc = norms64(h, 0, v))) return rc;  // leaves r64 - A x64 in t64
        s.norms++;
        const double r0 = std::sqrt(v[1]);
        s.res0 = s.res = r0;
        const double target = opts->rtol * r0;
        const int inner = opts->inner_cycles > 0 ? opts->inner_cycles : 1;
        if (!std::isfinite(r0)) rc = STZMG_ERR_NONFINITE;
        while (rc == 0 && s.res > target) {
            if (s.cycles >= opts->max_cycles) { rc = STZMG_ERR_NOT_CONVERGED; break; }
            const double scale = s.res / std::sqrt((double)n3);  // RMS of the defect
            hipLaunchKernelGGL((k_defect_f32<double>), dim3(nblk_ew(n3)), dim3(kThreads), 0, h->stream, (const double *)L.t64,
                               1.0 / scale, L.r32, fine_zero_start(h) ? nullptr : L.x32, n3);
            if ((rc = run_vcycles<float>(h, inner, graph, true))) return rc;
            s.cycles += inner;
            hipLaunchKernelGGL(k_update_f64, dim3(nblk(n3)), dim3(kThreads), 0, h->stream, L.x64,
                               (const float *)L.x32, scale, n3);
            LAUNCHCHK(h);
            s.outer++;
            if ((rc = norms64(h, 0, v))) return rc;
            s.norms++;
            s.res = std::sqrt(v[1]);
            if (!std::isfinite(s.res)) { rc = STZMG_ERR_NONFINITE; break; }
        }
    } else if (opts->mode == STZMG_SOLVE_FP32_GMRES) {
        // Right-preconditioned flexible GMRES in FP64 (Saad's FGMRES):
        // z_j = M^-1 v_j is inner_cycles FP32 V-cycles from zero on
        // sigma v_j (sigma = sqrt(3N) gives the FP32 right-hand side entries
        // of order one), kept in FP32; w = A z_j in FP64; classical
        // Gram-Schmidt with one fused dot pass and one fused update+norm pass;
        // Givens rotations on the host.  Each restart updates x with the FP32
        // vectors and recomputes the true FP64 residual, which decides.
        Level &L = h->lv[0];
        const long n3 = 3 * L.d.N;
        const int m = std::min(std::max(opts->restart, 1), kKrylovMax);
        if ((rc = ensure_krylov(h, m, n3, true))) return rc;
        if ((rc = norms64(h, 0, v))) return rc;  // t64 = r64 - A x64
        s.norms++;
        double beta = std::sqrt(v[1]);
        s.res0 = s.res = beta;
        const double target = opts->rtol * beta;
        const int inner = opts->inner_cycles > 0 ? opts->inner_cycles : 1;
        const double sigma = std::sqrt((double)n3);
        if (!std::isfinite(beta)) rc = STZMG_ERR_NONFINITE;
        std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m), hc(m + 2);
        auto Hc = [&](int i, int j) -> double & { return H[(size_t)i * m + j]; };
        const unsigned nb3 = nblk(n3);
        while (rc == 0 && s.res > target && beta > 0) {
            hipLaunchKernelGGL((k_scale<double, double>), dim3(nblk_ew(n3)), dim3(kThreads), 0, h->stream, h->kV, (const double *)L.t64,
                               1.0 / beta, n3);
            std::fill(g.begin(), g.end(), 0.0);
            g[0] = beta;
            int nj = 0;
            for (int j = 0; j < m; j++) {
                if (s.cycles >= opts->max_cycles) { rc = STZMG_ERR_NOT_CONVERGED; break; }
                double *vj = h->kV + (long)j * n3, *w = h->kV + (long)(j + 1) * n3;
                float *zj = h->kZ + (long)j * n3;
                // z_j = M^-1 (sigma v_j)
                hipLaunchKernelGGL((k_defect_f32<double>), dim3(nblk_ew(n3)), dim3(kThreads), 0, h->stream, (const double *)vj, sigma,
                                   L.r32, fine_zero_start(h) ? nullptr : L.x32, n3);
                if ((rc = run_vcycles<float>(h, inner, graph, true))) return rc;
                s.cycles += inner;
                HIPCHK(h, hipMemcpyAsync(zj, L.x32, sizeof(float) * n3, hipMemcpyDeviceToDevice, h->stream));
                // w' = A z_j (the true w is w' / sigma)
                hipLaunchKernelGGL((k_convert<float, double>), dim3(nblk_ew(n3)), dim3(kThreads), 0, h->stream,
                                   (const float *)zj, h->ktmp, n3);
                launch_apply<double>(h, L, (const double *)h->ktmp, w);
                // d_i = (w', v_i); w' -= sum d_i v_i; |w'|^2
                launch_mdot_n<1>(h->stream, (const double *)w, (const double *)h->kV, j + 1, n3, h->kpart);
                hipLaunchKernelGGL(k_reduce_cols, dim3(j + 1), dim3(kThreads), 0, h->stream,
                                   (const double *)h->kpart, kDotBlocks, j + 1, h->kdev, (const double *)nullptr);
                launch_maxpy_n<1>(h->stream, w, (const double *)h->kV, (const double *)h->kdev, j + 1, n3,
                                  h->kpart);
                hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(kThreads), 0, h->stream, (const double *)h->kpart,
                                   kDotBlocks, 1, h->kdev + (j + 1), (const double *)nullptr);
                LAUNCHCHK(h);
                HIPCHK(h, hipMemcpyAsync(hc.data(), h->kdev, sizeof(double) * (j + 2), hipMemcpyDeviceToHost,
                                         h->stream));
                HIPCHK(h, hipStreamSynchronize(h->stream));
                const double wn = std::sqrt(hc[j + 1]);
                for (int i = 0; i <= j; i++) Hc(i, j) = hc[i] / sigma;
                Hc(j + 1, j) = wn / sigma;
                if (!std::isfinite(wn)) { rc = STZMG_ERR_NONFINITE; break; }
                if (wn > 0)
                    hipLaunchKernelGGL((k_scale<double, double>), dim3(nblk_ew(n3)), dim3(kThreads), 0, h->stream, w, (const double *)w,
                                       1.0 / wn, n3);
                // Givens rotations
                for (int i = 0; i < j; i++) {
                    const double a = Hc(i, j), b = Hc(i + 1, j);
                    Hc(i, j) = cs[i] * a + sn[i] * b;
                    Hc(i + 1, j) = -sn[i] * a + cs[i] * b;
                }
                const double a = Hc(j, j), b = Hc(j + 1, j), r = std::hypot(a, b);
                cs[j] = r > 0 ? a / r : 1.0;
                sn[j] = r > 0 ? b / r : 0.0;
                Hc(j, j) = r;
                Hc(j + 1, j) = 0;
                g[j + 1] = -sn[j] * g[j];
                g[j] = cs[j] * g[j];
                nj = j + 1;
                if (std::fabs(g[j + 1]) <= target || wn == 0) break;
            }
            if (nj == 0) break;
            // y = H^-1 g, x += sum (y_i / sigma) z_i
            for (int i = nj - 1; i >= 0; i--) {
                double t = g[i];
                for (int k2 = i + 1; k2 < nj; k2++) t -= Hc(i, k2) * y[k2];
                y[i] = t / Hc(i, i);
            }
            for (int i = 0; i < nj; i++) hc[i] = y[i] / sigma;
            HIPCHK(h, hipMemcpyAsync(h->kdev, hc.data(), sizeof(double) * nj, hipMemcpyHostToDevice, h->stream));
            hipLaunchKernelGGL(k_xupdate<float>, dim3(nblk_ew(n3)), dim3(kThreads), 0, h->stream, L.x64, (const float *)h->kZ,
                               (const double *)h->kdev, nj, n3);
            LAUNCHCHK(h);
            s.outer++;
            if ((rc = norms64(h, 0, v))) return rc;
            s.norms++;
            beta = std::sqrt(v[1]);
            s.res = beta;
            if (!std::isfinite(s.res)) { rc = STZMG_ERR_NONFINITE; break; }
        }
    } else if (opts->mode == STZMG_SOLVE_FP32_GMRES_IR) {
        // GMRES-based iterative refinement with an FP32 inner solve: each
        // refinement step runs right-preconditioned GMRES(restart) on
        // A e = r64 (the true FP64 residual) with an FP32 Krylov basis, FP32
        // operator applies and inner_cycles FP32 V-cycles from zero as the
        // preconditioner (dot products and the Hessenberg in FP64), until
        // its residual estimate has fallen by inner_rtol; x64 += e in FP64,
        // and the new FP64 residual decides.  FP32 rounding limits what one
        // step can gain to about 1e-6; FP64 stays on the residual, the update,
        // the coarse operators and the coarsest level.
        //
        // Basis vectors are kept unnormalised: u_j = nrm_j v_j.  The
        // reduction of the dot products scales (w, u_i) by 1/nrm_i^2, which
        // gives the Gram-Schmidt coefficient of u_i directly, and the next
        // defect scales u_j by sigma/nrm_j, so no separate normalisation pass
        // runs.
        Level &L = h->lv[0];
        const long n3 = 3 * L.d.N;
        const int m = std::min(std::max(opts->restart, 1), kKrylovMax);
        if ((rc = ensure_krylov(h, m, n3, false))) return rc;
        if ((rc = norms64(h, 0, v))) return rc;  // t64 = r64 - A x64
        s.norms++;
        double beta = std::sqrt(v[1]);
        s.res0 = s.res = beta;
        const double target = opts->rtol * beta;
        const double inner_rtol = opts->inner_rtol > 0 ? opts->inner_rtol : 1e-4;
        const int inner = opts->inner_cycles > 0 ? opts->inner_cycles : 1;
        const double sigma = std::sqrt((double)n3);
        if (!std::isfinite(beta)) rc = STZMG_ERR_NONFINITE;
        std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m), hc(m + 2), nrm(m + 1),
            isq(kKrylovMax + 1);
        auto Hc = [&](int i, int j) -> double & { return H[(size_t)i * m + j]; };
        const unsigned nb3 = nblk(n3);
        double *dscale = h->kdev + (kKrylovMax + 1);  // 1/nrm_i^2 on the device
        while (rc == 0 && s.res > target && beta > 0) {
            hipLaunchKernelGGL((k_scale<float, double>), dim3(nblk_ew(n3)), dim3(kThreads), 0, h->stream, h->kV32,
                               (const double *)L.t64, 1.0 / beta, n3);
            std::fill(g.begin(), g.end(), 0.0);
            g[0] = beta;
            nrm[0] = 1.0;
            isq[0] = 1.0;
            const double stop = std::max(target, inner_rtol * beta);
            int nj = 0;
            for (int j = 0; j < m; j++) {
                if (s.cycles >= opts->max_cycles) { rc = STZMG_ERR_NOT_CONVERGED; break; }
                const float *uj = h->kV32 + (long)j * n3;
                float *w = h->kV32 + (long)(j + 1) * n3;
                float *zj = h->kZ + (long)j * n3;
                // z_j = M^-1 (sigma v_j) = M^-1 (sigma / nrm_j u_j)
                hipLaunchKernelGGL((k_defect_f32<float>), dim3(nblk_ew(n3)), dim3(kThreads), 0, h->stream, uj,
                                   sigma / nrm[j], L.r32, fine_zero_start(h) ? nullptr : L.x32, n3);
                if ((rc = run_vcycles<float>(h, inner, graph, true))) return rc;
                s.cycles += inner;
                // w' = A z_j in FP32 (the true w is w' / sigma); the same
                // pass keeps z_j (x32 is overwritten by the next V-cycles)
                launch_apply<float>(h, L, (const float *)L.x32, w, zj);
                // c_i = (w', u_i) / nrm_i^2; w' -= sum c_i u_i; |w'|^2
                HIPCHK(h, hipMemcpyAsync(dscale + j, &isq[j], sizeof(double), hipMemcpyHostToDevice, h->stream));
                launch_mdot_n<1, float>(h->stream, (const float *)w, (const float *)h->kV32, j + 1, n3, h->kpart);
                hipLaunchKernelGGL(k_reduce_cols, dim3(j + 1), dim3(kThreads), 0, h->stream,
                                   (const double *)h->kpart, kDotBlocks, j + 1, h->kdev, (const double *)dscale);
                launch_maxpy_n<1, float>(h->stream, w, (const float *)h->kV32, (const double *)h->kdev, j + 1, n3,
                                         h->kpart);
                hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(kThreads), 0, h->stream, (const double *)h->kpart,
                                   kDotBlocks, 1, h->kdev + (j + 1), (const double *)nullptr);
                LAUNCHCHK(h);
                HIPCHK(h, hipMemcpyAsync(hc.data(), h->kdev, sizeof(double) * (j + 2), hipMemcpyDeviceToHost,
                                         h->stream));
                HIPCHK(h, hipStreamSynchronize(h->stream));
                const double wn = std::sqrt(hc[j + 1]);
                for (int i = 0; i <= j; i++) Hc(i, j) = hc[i] * nrm[i] / sigma;
                Hc(j + 1, j) = wn / sigma;
                if (!std::isfinite(wn)) { rc = STZMG_ERR_NONFINITE; break; }
                nrm[j + 1] = wn;
                isq[j + 1] = wn > 0 ? 1.0 / (wn * wn) : 0.0;
                for (int i = 0; i < j; i++) {
                    const double a = Hc(i, j), b = Hc(i + 1, j);
                    Hc(i, j) = cs[i] * a + sn[i] * b;
                    Hc(i + 1, j) = -sn[i] * a + cs[i] * b;
                }
                const double a = Hc(j, j), b = Hc(j + 1, j), r = std::hypot(a, b);
                cs[j] = r > 0 ? a / r : 1.0;
                sn[j] = r > 0 ? b / r : 0.0;
                Hc(j, j) = r;
                Hc(j + 1, j) = 0;
                g[j + 1] = -sn[j] * g[j];
                g[j] = cs[j] * g[j];
                nj = j + 1;
                if (std::fabs(g[j + 1]) <= stop || wn == 0) break;
            }
            if (nj == 0) break;
            for (int i = nj - 1; i >= 0; i--) {
                double t = g[i];
                for (int k2 = i + 1; k2 < nj; k2++) t -= Hc(i, k2) * y[k2];
                y[i] = t / Hc(i, i);
            }
            for (int i = 0; i < nj; i++) hc[i] = y[i] / sigma;
            HIPCHK(h, hipMemcpyAsync(h->kdev, hc.data(), sizeof(double) * nj, hipMemcpyHostToDevice, h->stream));
            hipLaunchKernelGGL(k_xupdate<float>, dim3(nblk_ew(n3)), dim3(kThreads), 0, h->stream, L.x64, (const float *)h->kZ,
                               (const double *)h->kdev, nj, n3);
            LAUNCHCHK(h);
            s.outer++;
            if ((rc = norms64(h, 0, v))) return rc;
            s.norms++;
            beta = std::sqrt(v[1]);
            s.res = beta;
            if (!std::isfinite(s.res)) { rc = STZMG_ERR_NONFINITE; break; }
        }
    } else if (opts->mode == STZMG_SOLVE_GMRES64) {
        // Flexible GMRES(m) in FP64 right-preconditioned by inner_cycles FP64
        // V-cycles from zero: z_j = M^-1 v_j with the level-0 FP64 fields
        // as the V-cycle's workspace, so the solution and right-hand side
        // live in kX64 / kB64 during the solve.  The FP64 counterpart of
        // STZMG_SOLVE_FP32_GMRES (same Gram-Schmidt, Givens rotations and
        // stopping on the true residual at each restart).
        Level &L = h->lv[0];
        const long n3 = 3 * L.d.N;
        const int m = std::min(std::max(opts->restart, 1), kKrylovMax);
        if ((rc = ensure_krylov(h, m, n3, true))) return rc;
        if ((rc = ensure_krylov64(h, m, n3))) return rc;
        HIPCHK(h, hipMemcpyAsync(h->kX64, L.x64, sizeof(double) * n3, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(h->kB64, L.r64, sizeof(double) * n3, hipMemcpyDeviceToDevice, h->stream));
        if ((rc = norms64_of(h, 0, h->kX64, h->kB64, v))) return rc;  // t64 = b - A x
        s.norms++;
        double beta = std::sqrt(v[1]);
        s.res0 = s.res = beta;
        const double target = opts->rtol * beta;
        const int inner = opts->inner_cycles > 0 ? opts->inner_cycles : 1;
        if (!std::isfinite(beta)) rc = STZMG_ERR_NONFINITE;
        std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m), hc(m + 2);
        auto Hc = [&](int i, int j) -> double & { return H[(size_t)i * m + j]; };
        const unsigned nb3 = nblk(n3);
        while (rc == 0 && s.res > target && beta > 0) {
            hipLaunchKernelGGL((k_scale<double, double>), dim3(nblk_ew(n3)), dim3(kThreads), 0, h->stream, h->kV,
                               (const double *)L.t64, 1.0 / beta, n3);
            std::fill(g.begin(), g.end(), 0.0);
            g[0] = beta;
            int nj = 0;
            for (int j = 0; j < m; j++) {
                if (s.cycles >= opts->max_cycles) { rc = STZMG_ERR_NOT_CONVERGED; break; }
                double *vj = h->kV + (long)j * n3, *w = h->kV + (long)(j + 1) * n3;
                double *zj = h->kZ64 + (long)j * n3;
                // z_j = M^-1 v_j
                HIPCHK(h, hipMemcpyAsync(L.r64, vj, sizeof(double) * n3, hipMemcpyDeviceToDevice, h->stream));
                if (!fine_zero_start(h))
                    hipLaunchKernelGGL((k_fill<double>), dim3(nb3), dim3(kThreads), 0, h->stream, L.x64, 0.0, n3);
                if ((rc = run_vcycles<double>(h, inner, graph, true))) return rc;
                s.cycles += inner;
                // w = A z_j; the same pass keeps z_j
                launch_apply<double>(h, L, (const double *)L.x64, w, zj);
                launch_mdot_n<1>(h->stream, (const double *)w, (const double *)h->kV, j + 1, n3, h->kpart);
                hipLaunchKernelGGL(k_reduce_cols, dim3(j + 1), dim3(kThreads), 0, h->stream,
                                   (const double *)h->kpart, kDotBlocks, j + 1, h->kdev, (const double *)nullptr);
                launch_maxpy_n<1>(h->stream, w, (const double *)h->kV, (const double *)h->kdev, j + 1, n3,
                                  h->kpart);
                hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(kThreads), 0, h->stream, (const double *)h->kpart,
                                   kDotBlocks, 1, h->kdev + (j + 1), (const double *)nullptr);
                LAUNCHCHK(h);
                HIPCHK(h, hipMemcpyAsync(hc.data(), h->kdev, sizeof(double) * (j + 2), hipMemcpyDeviceToHost,
                                         h->stream));
                HIPCHK(h, hipStreamSynchronize(h->stream));
                const double wn = std::sqrt(hc[j + 1]);
                for (int i = 0; i <= j; i++) Hc(i, j) = hc[i];
                Hc(j + 1, j) = wn;
                if (!std::isfinite(wn)) { rc = STZMG_ERR_NONFINITE; break; }
                if (wn > 0)
                    hipLaunchKernelGGL((k_scale<double, double>), dim3(nblk_ew(n3)), dim3(kThreads), 0, h->stream, w,
                                       (const double *)w, 1.0 / wn, n3);
                for (int i = 0; i < j; i++) {
                    const double a = Hc(i, j), b = Hc(i + 1, j);
                    Hc(i, j) = cs[i] * a + sn[i] * b;
                    Hc(i + 1, j) = -sn[i] * a + cs[i] * b;
                }
                const double a = Hc(j, j), b = Hc(j + 1, j), r = std::hypot(a, b);
                cs[j] = r > 0 ? a / r : 1.0;
                sn[j] = r > 0 ? b / r : 0.0;
                Hc(j, j) = r;
                Hc(j + 1, j) = 0;
                g[j + 1] = -sn[j] * g[j];
                g[j] = cs[j] * g[j];
                nj = j + 1;
                if (std::fabs(g[j + 1]) <= target || wn == 0) break;
            }
            if (nj == 0) break;
            for (int i = nj - 1; i >= 0; i--) {
                double t = g[i];
                for (int k2 = i + 1; k2 < nj; k2++) t -= Hc(i, k2) * y[k2];
                y[i] = t / Hc(i, i);
            }
            HIPCHK(h, hipMemcpyAsync(h->kdev, y.data(), sizeof(double) * nj, hipMemcpyHostToDevice, h->stream));
            hipLaunchKernelGGL(k_xupdate<double>, dim3(nblk_ew(n3)), dim3(kThreads), 0, h->stream, h->kX64,
                               (const double *)h->kZ64, (const double *)h->kdev, nj, n3);
            LAUNCHCHK(h);
            s.outer++;
            if ((rc = norms64_of(h, 0, h->kX64, h->kB64, v))) return rc;
            s.norms++;
            beta = std::sqrt(v[1]);
            s.res = beta;
            if (!std::isfinite(s.res)) { rc = STZMG_ERR_NONFINITE; break; }
        }
        HIPCHK(h, hipMemcpyAsync(L.x64, h->kX64, sizeof(double) * n3, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(L.r64, h->kB64, sizeof(double) * n3, hipMemcpyDeviceToDevice, h->stream));
    } else {
        return fail(h, STZMG_ERR_ARG, "unknown solve mode %d", opts->mode);
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    s.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (st) *st = s;
    if (rc == STZMG_ERR_NOT_CONVERGED) return fail(h, rc, "not converged after %d cycles (res %.3e)", s.cycles, s.res);
    if (rc == STZMG_ERR_NONFINITE) return fail(h, rc, "non-finite residual after %d cycles", s.cycles);
    return rc;
}

int stzmg_time_kernel(stzmg_t h, int kind, int level, int reps, double *ms_per_unit) {
    if (int rc = check_level(h, level)) return rc;
    if (int rc = require_ops(h)) return rc;
    if (reps < 1 || !ms_per_unit || kind < 0 || kind > 5) return STZMG_ERR_ARG;
    Level &L = h->lv[level];
    if (kind >= 4 && !(kind == 4 ? L.fused : L.fused && L.ftile >= 1))
        return fail(h, STZMG_ERR_UNSUPPORTED, "time_kernel: level %d has no fused sweep", level);
    double *home64 = L.x64;
    float *home32 = L.x32;
    hipEvent_t a, b;
    HIPCHK(h, hipEventCreate(&a));
    HIPCHK(h, hipEventCreate(&b));
    auto body = [&]() {
        if (kind == 0 || kind == 4) launch_gs_sweep<double>(h, L);
        else if (kind == 1) launch_resid<double, false>(h, L, L.x64, L.r64, L.t64);
        else if (kind == 2 || kind == 5) launch_gs_sweep<float>(h, L);
        else launch_resid<float, false>(h, L, L.x32, L.r32, L.t32);
    };
    body();
    float ms = 0;
    if (kind < 4) {
        HIPCHK(h, hipEventRecord(a, h->stream));
        for (int q = 0; q < reps; q++) body();
        HIPCHK(h, hipEventRecord(b, h->stream));
        HIPCHK(h, hipEventSynchronize(b));
        HIPCHK(h, hipEventElapsedTime(&ms, a, b));
        ms /= reps;
    } else {
        // one event pair around each launch of the fused sweep (two per
        // sweep): the mean kernel duration, without the gaps between launches
        std::vector<hipEvent_t> ev(4 * (size_t)reps);
        hipError_t e = hipSuccess;
        size_t made = 0;
        for (; made < ev.size() && e == hipSuccess; made++) e = hipEventCreate(&ev[made]);
        if (e == hipSuccess) {
            h->launch_ev = &ev;
            h->launch_ev_next = 0;
            for (int q = 0; q < reps; q++) body();
            h->launch_ev = nullptr;
            e = hipStreamSynchronize(h->stream);
        }
        double tot = 0;
        const size_t used = h->launch_ev_next;
        for (size_t q = 0; e == hipSuccess && q + 1 < used; q += 2) {
            float d = 0;
            e = hipEventElapsedTime(&d, ev[q], ev[q + 1]);
            tot += d;
        }
        h->launch_ev = nullptr;
        for (size_t q = 0; q < made; q++) (void)hipEventDestroy(ev[q]);
        if (e != hipSuccess) {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            return fail(h, STZMG_ERR_HIP, "time_kernel: %s", hipGetErrorString(e));
        }
        if (used < 2) {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            return fail(h, STZMG_ERR_UNSUPPORTED, "time_kernel: level %d has no fused sweep", level);
        }
        ms = (float)(tot / (double)(used / 2));
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    restore_level<double>(h, L, home64);
    restore_level<float>(h, L, home32);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    *ms_per_unit = ms;
    return 0;
}

// ---- operator builders -------------------------------------------------------

int stzmg_shear_stencil(double mu, double lamb, double dx, double dy, double dz, double dt, double kappa,
                        double *B, double *fix) {
    if (!B) return STZMG_ERR_ARG;
    // stencil_base (src/shear_sim_3d.cpp:216-331), then -dt scaling and the
    // Dirichlet fix block (src/shear_sim_3d.hh:45-55).
    double base[27][9];
    memset(base, 0, sizeof(base));
    const double L2 = lamb + 2 * mu, kd = kappa / dt, c = (lamb + mu) / 4;
    const double ax[3] = {L2 / (dx * dx) + kd / (dx * dx), mu / (dx * dx) + kd / (dx * dx), mu / (dx * dx) + kd / (dx * dx)};
    const double ay[3] = {mu / (dy * dy) + kd / (dy * dy), L2 / (dy * dy) + kd / (dy * dy), mu / (dy * dy) + kd / (dy * dy)};
    const double az[3] = {mu / (dz * dz) + kd / (dz * dz), mu / (dz * dz) + kd / (dz * dz), L2 / (dz * dz) + kd / (dz * dz)};
    auto bi = [](int di, int dj, int dk) { return 13 + di + 3 * dj + 9 * dk; };
    for (int s = -1; s <= 1; s += 2) {
        for (int e = 0; e < 3; e++) {
            base[bi(s, 0, 0)][4 * e] = ax[e];
            base[bi(0, s, 0)][4 * e] = ay[e];
            base[bi(0, 0, s)][4 * e] = az[e];
        }
    }
    for (int si = -1; si <= 1; si += 2)
        for (int sj = -1; sj <= 1; sj += 2) {
            const double sg = si * sj;
            double *b = base[bi(si, sj, 0)];
            b[1] = b[3] = sg * c / (dx * dy);
            b = base[bi(si, 0, sj)];
            b[2] = b[6] = sg * c / (dx * dz);
            b = base[bi(0, si, sj)];
            b[5] = b[7] = sg * c / (dy * dz);
        }
    const double ks = kd * (1 / (dx * dx) + 1 / (dy * dy) + 1 / (dz * dz));
    base[13][0] = -2 * (L2 / (dx * dx) + mu / (dy * dy) + mu / (dz * dz) + ks);
    base[13][4] = -2 * (mu / (dx * dx) + L2 / (dy * dy) + mu / (dz * dz) + ks);
    // The reference adds kappa/dt*(...) to a33 outside the -2*(...) factor
    // (src/shear_sim_3d.cpp:278); inert for kappa = 0 as in every QS run.
    base[13][8] = -2 * (mu / (dx * dx) + mu / (dy * dy) + L2 / (dz * dz)) + ks;
    for (int b = 0; b < 27; b++)
        for (int e = 0; e < 9; e++) B[9 * b + e] = -dt * base[b][e];
    if (fix) {
        const double f = -dt * 2 * (L2 / (dx * dx) + mu / (dy * dy) + mu / (dz * dz));
        for (int e = 0; e < 9; e++) fix[e] = (e % 4 == 0) ? f : 0.0;
    }
    return 0;
}

int stzmg_trans_stencil(const double *T9, double mu, double lamb, double dx, double dy, double dz, double dt,
                        double *B) {
    if (!T9 || !B) return STZMG_ERR_ARG;
    // Closed form of the stenc_funcs blocks: S = symmetric matrix from the
    // upper triangle of T, W = adj(S)/det(T);
    // K[i][m][a][b] = mu S_im (W W^T)_ab + (lamb+mu) W_ai delta_bm.
    double S[3][3];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) S[a][b] = a <= b ? T9[3 * a + b] : T9[3 * b + a];
    const double D = T9[0] * (T9[4] * T9[8] - T9[5] * T9[7]) - T9[1] * (T9[3] * T9[8] - T9[5] * T9[6]) +
                     T9[2] * (T9[3] * T9[7] - T9[4] * T9[6]);
    if (!(D != 0.0)) return STZMG_ERR_SINGULAR;
    double W[3][3];
    W[0][0] = (S[1][1] * S[2][2] - S[1][2] * S[2][1]) / D;
    W[0][1] = -(S[0][1] * S[2][2] - S[0][2] * S[2][1]) / D;
    W[0][2] = (S[0][1] * S[1][2] - S[0][2] * S[1][1]) / D;
    W[1][0] = -(S[1][0] * S[2][2] - S[1][2] * S[2][0]) / D;
    W[1][1] = (S[0][0] * S[2][2] - S[0][2] * S[2][0]) / D;
    W[1][2] = -(S[0][0] * S[1][2] - S[0][2] * S[1][0]) / D;
    W[2][0] = (S[1][0] * S[2][1] - S[1][1] * S[2][0]) / D;
    W[2][1] = -(S[0][0] * S[2][1] - S[0][1] * S[2][0]) / D;
    W[2][2] = (S[0][0] * S[1][1] - S[0][1] * S[1][0]) / D;
    double WW[3][3];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) WW[a][b] = W[a][0] * W[b][0] + W[a][1] * W[b][1] + W[a][2] * W[b][2];
    auto K = [&](int i, int m, int a, int b) {
        return mu * S[i][m] * WW[a][b] + (lamb + mu) * W[a][i] * (b == m ? 1.0 : 0.0);
    };
    const double h[3] = {dx, dy, dz};
    double out[27][9];
    memset(out, 0, sizeof(out));
    for (int a = 0; a < 3; a++) {
        int off[3] = {0, 0, 0};
        for (int s = -1; s <= 1; s += 2) {
            off[a] = s;
            const int b = 13 + off[0] + 3 * off[1] + 9 * off[2];
            for (int i = 0; i < 3; i++)
                for (int m = 0; m < 3; m++) out[b][3 * i + m] += K(i, m, a, a) / (h[a] * h[a]);
        }
        for (int i = 0; i < 3; i++)
            for (int m = 0; m < 3; m++) out[13][3 * i + m] -= 2 * K(i, m, a, a) / (h[a] * h[a]);
    }
    for (int a = 0; a < 3; a++)
        for (int c = a + 1; c < 3; c++)
            for (int sa = -1; sa <= 1; sa += 2)
                for (int sc = -1; sc <= 1; sc += 2) {
                    int off[3] = {0, 0, 0};
                    off[a] = sa;
                    off[c] = sc;
                    const int b = 13 + off[0] + 3 * off[1] + 9 * off[2];
                    for (int i = 0; i < 3; i++)
                        for (int m = 0; m < 3; m++)
                            out[b][3 * i + m] += sa * sc * (K(i, m, a, c) + K(i, m, c, a)) / (4 * h[a] * h[c]);
                }
    for (int b = 0; b < 27; b++)
        for (int e = 0; e < 9; e++) B[9 * b + e] = -dt * out[b][e];
    return 0;
}

}  // extern "C"
