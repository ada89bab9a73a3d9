// Random initial chi field on the GPU (include/stz.h, stz_grf_*,
// stz_set_random_initial_conditions).
//
// The reference (shear_sim_3d::set_random_initial_conditions,
// src/shear_sim_3d.cpp:2793-2934) convolves a noise grid, padded by `cut`
// points on each side of every non-periodic axis, with the weights
// exp(-llinv (i^2 + j^2 + k^2)) over the cube [-cut, cut]^3 by brute force
// ((2 cut + 1)^3 taps per cell, convolve :2997-3041) and caches the result
// in cg_*.dat.  The cube weight is a product of three 1-D weights, so here
// the convolution runs as three 1-D passes of 2 cut + 1 taps:
//   x pass  noise (nxt, nyt, nzt) -> t1 (Nx, nyt, nzt)
//   y pass  t1                    -> t2 (Nx, Ny, nzt)
//   z pass  t2                    -> chi (Nx, Ny, Nz), chi = chi_avg + chi_1 v
// each staging its window in LDS.  Results equal the reference's to
// rounding (the summation order differs).
//
// Noise is either given by the caller (for instance the reference's own
// grid) or drawn on the GPU: Philox4x64-10 (Salmon et al., SC 2011) with
// counter (pair, attempt, 0, 0) and key (seed, fixed word), fed to the polar
// Box-Muller method of the reference (box_muller, :3057-3067, with log(r)
// taken as log1p(r - 1) for r >= 1/2); pair p fills
// elements 2p and 2p+1 in the order of fill_noise_grid (:2978-2992).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/stz.h"
#include "stz_internal.hpp"

namespace {

constexpr uint64_t kGrfKey1 = 0x53545A2B2B475246ULL;  // second key word, fixed
constexpr int kGrfAttempts = 64;                      // polar-method attempts per pair
constexpr int kRowOut = 256;                          // x pass: outputs per workgroup
constexpr int kColX = 64, kColA = 64;                 // y/z pass tile: 64 x-columns, 64 outputs along the axis
constexpr int kLdsCut = 24;                           // y/z passes stage their window in LDS up to this cut (57 KB)

__device__ __forceinline__ void philox4x64(uint64_t c[4], uint64_t k0, uint64_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B97F4A7C15ULL;
            k1 += 0xBB67AE8584CAA73BULL;
        }
        const uint64_t a = 0xD2E7470EE14C6C93ULL, b = 0xCA5A826395121157ULL;
        const uint64_t hi0 = __umul64hi(a, c[0]), lo0 = a * c[0];
        const uint64_t hi1 = __umul64hi(b, c[2]), lo1 = b * c[2];
        const uint64_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
    }
}

// r = x^2 + y^2 with each product and the sum rounded separately, as the
// host restatement computes it: a fused multiply-add would change r in its
// last bit, and with it the accepted attempt near r = 1 and log(r) by up to
// 1e-13 relative near r = 1.
__device__ __forceinline__ double radius2(double x, double y) {
#pragma clang fp contract(off)
    const double xx = x * x;
    const double yy = y * y;
    return xx + yy;
}

// One thread per noise pair.
__global__ __launch_bounds__(256) void k_grf_noise(uint64_t seed, long n, double *__restrict__ out) {
    const long np = (n + 1) / 2;
    for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < np; p += (long)gridDim.x * 256) {
        double r0 = 0, r1 = 0;
        for (int a = 0; a < kGrfAttempts; a++) {
            uint64_t c[4] = {(uint64_t)p, (uint64_t)a, 0, 0};
            philox4x64(c, seed, kGrfKey1);
            const double x = 2.0 * ((double)(c[0] >> 11) * 0x1.0p-53) - 1.0;
            const double y = 2.0 * ((double)(c[1] >> 11) * 0x1.0p-53) - 1.0;
            const double r = radius2(x, y);
            if (r == 0.0 || r > 1.0) continue;
            // log(r) through log1p(r - 1) for r >= 1/2, where r - 1 is exact:
            // relative-accurate near r = 1, where log(r) itself is small
            const double lr = r >= 0.5 ? log1p(r - 1.0) : log(r);
            const double d = sqrt(-2.0 * lr / r);
            r0 = x * d;
            r1 = y * d;
            break;
        }
        out[2 * p] = r0;
        if (2 * p + 1 < n) out[2 * p + 1] = r1;
    }
}

// Source index along an axis for output position v + offset: the
// reference's single periodic wrap (valid for cut <= n), or the shift into
// the padded noise grid of a non-periodic axis.
__device__ __forceinline__ int src_index(int v, int n, int per, int cut) {
    if (per) return v > n - 1 ? v - n : (v < 0 ? v + n : v);
    return v + cut;
}

// x pass: one workgroup = kRowOut consecutive outputs of one (y', z') row.
__global__ __launch_bounds__(kRowOut) void k_grf_x(const double *__restrict__ in, double *__restrict__ out,
                                                   const double *__restrict__ w, int nx, int nxt, int per,
                                                   int cut, long rows, int segs) {
    extern __shared__ double win[];
    const long row = blockIdx.x / segs;
    const int x0 = (int)(blockIdx.x % segs) * kRowOut;
    if (row >= rows) return;
    const double *src = in + row * nxt;
    const int span = kRowOut + 2 * cut;
    for (int k = threadIdx.x; k < span; k += kRowOut) {
        const int v = x0 - cut + k;  // output-grid position of this window slot
        if (v < nx + cut) win[k] = src[src_index(v, nx, per, cut)];
    }
    __syncthreads();
    const int x = x0 + threadIdx.x;
    if (x >= nx) return;
    double s = 0;
    for (int i = -cut; i <= cut; i++) s += w[i < 0 ? -i : i] * win[threadIdx.x + cut + i];
    out[row * nx + x] = s;
}

// Geometry of a y or z pass.  Element (x, a, o) of the input sits at
// x + sa_in a + so_in o, of the output at x + sa_out a + so_out o.
struct AxisPass {
    int nx, na_out, per, cut, no;
    long sa_in, so_in, sa_out, so_out;
    double avg, scale;  // out = avg + scale v (the z pass); 0, 1 otherwise
};

// y/z pass with the window in LDS: a tile of kColX columns x kColA outputs
// for one o; 256 threads = kColX lanes x 4 groups.
__global__ __launch_bounds__(256) void k_grf_axis_lds(const double *__restrict__ in, double *__restrict__ out,
                                                      const double *__restrict__ w, AxisPass P, int tiles_x,
                                                      int tiles_a) {
    extern __shared__ double win[];  // [kColA + 2 cut][kColX]
    const int bx = blockIdx.x % tiles_x;
    const int ba = (blockIdx.x / tiles_x) % tiles_a;
    const long o = blockIdx.x / ((long)tiles_x * tiles_a);
    if (o >= P.no) return;
    const int lane = threadIdx.x % kColX, grp = threadIdx.x / kColX;
    const int x = bx * kColX + lane, a0 = ba * kColA;
    const int span = kColA + 2 * P.cut;
    for (int k = grp; k < span; k += 4) {
        const int v = a0 - P.cut + k;
        if (x < P.nx && v < P.na_out + P.cut)
            win[k * kColX + lane] = in[x + P.sa_in * src_index(v, P.na_out, P.per, P.cut) + P.so_in * o];
    }
    __syncthreads();
    if (x >= P.nx) return;
    for (int m = grp; m < kColA; m += 4) {
        const int a = a0 + m;
        if (a >= P.na_out) break;
        double s = 0;
        for (int j = -P.cut; j <= P.cut; j++) s += w[j < 0 ? -j : j] * win[(m + P.cut + j) * kColX + lane];
        out[x + P.sa_out * a + P.so_out * o] = P.avg + P.scale * s;
    }
}

// y/z pass for large cutoffs: one thread per output, window read through L2.
__global__ __launch_bounds__(256) void k_grf_axis_direct(const double *__restrict__ in, double *__restrict__ out,
                                                         const double *__restrict__ w, AxisPass P) {
    const long total = (long)P.nx * P.na_out * P.no;
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < total; q += (long)gridDim.x * 256) {
        const int x = (int)(q % P.nx);
        const int a = (int)((q / P.nx) % P.na_out);
        const long o = q / ((long)P.nx * P.na_out);
        double s = 0;
        for (int j = -P.cut; j <= P.cut; j++)
            s += w[j < 0 ? -j : j] * in[x + P.sa_in * src_index(a + j, P.na_out, P.per, P.cut) + P.so_in * o];
        out[x + P.sa_out * a + P.so_out * o] = P.avg + P.scale * s;
    }
}

struct Geom {
    int N[3], prd[3], nt[3], cut;
    double llinv;
    long nnoise() const { return (long)nt[0] * nt[1] * nt[2]; }
    long ncell() const { return (long)N[0] * N[1] * N[2]; }
};

thread_local std::string g_grf_err;

int gfail(int code, const char *msg) {
    g_grf_err = msg;
    return code;
}

int make_geom(const int N[3], const int prd[3], const stz_grf_params *g, Geom &G) {
    if (!N || !prd || !g) return gfail(STZMG_ERR_ARG, "null argument");
    if (!(g->conv_l > 0)) return gfail(STZMG_ERR_ARG, "conv_l must be positive");
    // gauss_fac and cut are int members of the reference, initialised in
    // this order (src/shear_sim_3d.hh:140-142, src/shear_sim_3d.cpp:57)
    const int gf = (int)g->gauss_fac;
    const double cutd = gf * g->conv_l;
    if (!(cutd >= 0) || cutd > 4096) return gfail(STZMG_ERR_ARG, "cutoff out of range");
    G.cut = (int)cutd;
    G.llinv = 1. / g->conv_l;
    for (int a = 0; a < 3; a++) {
        if (N[a] < 1) return gfail(STZMG_ERR_ARG, "grid sizes must be positive");
        G.N[a] = N[a];
        G.prd[a] = prd[a] ? 1 : 0;
        // the reference's single periodic wrap reaches cut <= N
        if (G.prd[a] && G.cut > N[a]) return gfail(STZMG_ERR_ARG, "cutoff exceeds a periodic grid size");
        G.nt[a] = G.prd[a] ? N[a] : N[a] + 2 * G.cut;
    }
    return 0;
}

// gauss_normal_factor_long (src/shear_sim_3d.cpp:3071-3078), same order.
double normal_factor(int cut, double llinv) {
    double g = 0;
    for (int ii = -cut; ii <= cut; ii++)
        for (int jj = -cut; jj <= cut; jj++)
            for (int kk = -cut; kk <= cut; kk++)
                g += exp(-llinv * (ii * ii + jj * jj + kk * kk)) * exp(-llinv * (ii * ii + jj * jj + kk * kk));
    return sqrt(g);
}

struct DevBuf {
    double *p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() {
        if (p) hipFree(p);
    }
};

#define GHIP(call)                                                           \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess) {                                              \
            g_grf_err = std::string(#call) + ": " + hipGetErrorString(e_);   \
            return STZMG_ERR_HIP;                                            \
        }                                                                    \
    } while (0)

// Buffers and launches of one configuration: stage 0 draws the noise,
// stages 1-3 are the x, y and z passes.
struct Plan {
    Geom G{};
    double chi_avg = 0, chi_1 = 0;
    uint64_t seed = 0;
    DevBuf w, noise, t1, t2;

    int setup(const Geom &geom, const stz_grf_params *g, bool own_noise, hipStream_t st) {
        G = geom;
        const int cut = G.cut;
        chi_avg = g->chi_avg;
        chi_1 = g->sigma / normal_factor(cut, G.llinv);
        seed = (uint64_t)g->seed;
        if (sizeof(double) * (kRowOut + 2 * cut) > 64 * 1024)
            return gfail(STZMG_ERR_UNSUPPORTED, "cutoff too large for the x pass");
        std::vector<double> wh(cut + 1);
        for (int i = 0; i <= cut; i++) wh[i] = exp(-G.llinv * (double)(i * i));
        const int nx = G.N[0], ny = G.N[1];
        GHIP(hipMalloc((void **)&w.p, sizeof(double) * (cut + 1)));
        GHIP(hipMalloc((void **)&t1.p, sizeof(double) * (size_t)nx * G.nt[1] * G.nt[2]));
        GHIP(hipMalloc((void **)&t2.p, sizeof(double) * (size_t)nx * ny * G.nt[2]));
        if (own_noise) GHIP(hipMalloc((void **)&noise.p, sizeof(double) * (size_t)G.nnoise()));
        GHIP(hipMemcpyAsync(w.p, wh.data(), sizeof(double) * (cut + 1), hipMemcpyHostToDevice, st));
        GHIP(hipStreamSynchronize(st));  // wh is released on return
        return 0;
    }

    int enqueue(int stage, const double *noise_dev, double *chi, hipStream_t st) const {
        const int cut = G.cut;
        const int nx = G.N[0], ny = G.N[1], nz = G.N[2];
        const int nyt = G.nt[1], nzt = G.nt[2];
        if (stage == 0) {
            const long nn = G.nnoise(), np = (nn + 1) / 2;
            const unsigned nb = (unsigned)std::min<long>((np + 255) / 256, 1L << 16);
            hipLaunchKernelGGL(k_grf_noise, dim3(nb), dim3(256), 0, st, seed, nn, noise.p);
        } else if (stage == 1) {
            const int segs = (nx + kRowOut - 1) / kRowOut;
            const long rows = (long)nyt * nzt;
            if (rows * segs > 0x7fffffffL) return gfail(STZMG_ERR_UNSUPPORTED, "grid too large");
            hipLaunchKernelGGL(k_grf_x, dim3((unsigned)(rows * segs)), dim3(kRowOut),
                               sizeof(double) * (kRowOut + 2 * cut), st, noise_dev, t1.p, (const double *)w.p, nx,
                               G.nt[0], G.prd[0], cut, rows, segs);
        } else {
            const AxisPass P = stage == 2
                ? AxisPass{nx, ny, G.prd[1], cut, nzt, nx, (long)nx * nyt, nx, (long)nx * ny, 0.0, 1.0}
                : AxisPass{nx, nz, G.prd[2], cut, ny, (long)nx * ny, nx, (long)nx * ny, nx, chi_avg, chi_1};
            const double *in = stage == 2 ? t1.p : t2.p;
            double *out = stage == 2 ? t2.p : chi;
            if (cut <= kLdsCut) {
                const int tx = (P.nx + kColX - 1) / kColX, ta = (P.na_out + kColA - 1) / kColA;
                const long nb = (long)tx * ta * P.no;
                if (nb > 0x7fffffffL) return gfail(STZMG_ERR_UNSUPPORTED, "grid too large");
                hipLaunchKernelGGL(k_grf_axis_lds, dim3((unsigned)nb), dim3(256),
                                   sizeof(double) * kColX * (kColA + 2 * cut), st, in, out, (const double *)w.p, P,
                                   tx, ta);
            } else {
                const long tot = (long)P.nx * P.na_out * P.no;
                const unsigned nb = (unsigned)std::min<long>((tot + 255) / 256, 1L << 20);
                hipLaunchKernelGGL(k_grf_axis_direct, dim3(nb), dim3(256), 0, st, in, out, (const double *)w.p, P);
            }
        }
        GHIP(hipGetLastError());
        return 0;
    }
};

// chi (device, Nx Ny Nz) from noise (device, the padded grid; drawn from
// the seed when null) on stream st.
int grf_device(const Geom &G, const stz_grf_params *g, const double *noise_dev, double *chi, hipStream_t st) {
    Plan P;
    if (int rc = P.setup(G, g, noise_dev == nullptr, st)) return rc;
    if (!noise_dev) {
        if (int rc = P.enqueue(0, nullptr, nullptr, st)) return rc;
        noise_dev = P.noise.p;
    }
    for (int stage = 1; stage <= 3; stage++)
        if (int rc = P.enqueue(stage, noise_dev, chi, st)) return rc;
    GHIP(hipStreamSynchronize(st));
    return 0;
}

// Name of the reference's cache file (src/shear_sim_3d.cpp:2812-2816).
std::string cache_name(const Geom &G, const stz_grf_params *g) {
    const double gnf = normal_factor(G.cut, G.llinv);
    const double chi_1 = g->sigma / gnf;
    const int human_mu = (int)(g->TZ * g->chi_avg);
    const int human_sig = (int)(g->TZ * chi_1 * gnf + 1);
    const char *pref = (g->sim_case == 19 && !G.prd[0] && !G.prd[1]) ? "f" : (G.prd[2] ? "t" : "");
    char buf[256];
    snprintf(buf, sizeof(buf), "%scg_%d_%d_%d_mu%d_sig%d.dat", pref, G.N[0], G.N[1], G.N[2], human_mu, human_sig);
    return buf;
}

}  // namespace

extern "C" {

void stz_grf_params_default(stz_grf_params *g) {
    if (!g) return;
    // the shear_sim driver's values (src/shear_sim.cpp:8, :83-86)
    g->gauss_fac = 5;
    g->conv_l = 5;
    g->TZ = 21000;
    g->chi_avg = 550. / 21000;
    g->sigma = 15. / 21000;
    g->sim_case = 0;
    g->seed = 0;
}

const char *stz_grf_last_error(void) { return g_grf_err.c_str(); }

int stz_grf_geometry(const int N[3], const int prd[3], const stz_grf_params *g, int *cut, long long *noise_size,
                     double *chi_1) {
    Geom G;
    if (int rc = make_geom(N, prd, g, G)) return rc;
    if (cut) *cut = G.cut;
    if (noise_size) *noise_size = G.nnoise();
    if (chi_1) *chi_1 = g->sigma / normal_factor(G.cut, G.llinv);
    return 0;
}

int stz_grf_cache_name(const int N[3], const int prd[3], const stz_grf_params *g, char *buf, int len) {
    Geom G;
    if (int rc = make_geom(N, prd, g, G)) return rc;
    const std::string s = cache_name(G, g);
    if (!buf || len < (int)s.size() + 1) return gfail(STZMG_ERR_ARG, "name buffer too small");
    memcpy(buf, s.c_str(), s.size() + 1);
    return 0;
}

int stz_grf_field(const int N[3], const int prd[3], const stz_grf_params *g, const double *noise, double *chi) {
    Geom G;
    if (int rc = make_geom(N, prd, g, G)) return rc;
    if (!chi) return gfail(STZMG_ERR_ARG, "null output");
    hipStream_t st = nullptr;
    GHIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { hipStreamDestroy(s); }
    } guard{st};
    DevBuf nd, cd;
    if (noise) {
        GHIP(hipMalloc((void **)&nd.p, sizeof(double) * (size_t)G.nnoise()));
        GHIP(hipMemcpyAsync(nd.p, noise, sizeof(double) * (size_t)G.nnoise(), hipMemcpyHostToDevice, st));
    }
    GHIP(hipMalloc((void **)&cd.p, sizeof(double) * (size_t)G.ncell()));
    if (int rc = grf_device(G, g, nd.p, cd.p, st)) return rc;
    GHIP(hipMemcpyAsync(chi, cd.p, sizeof(double) * (size_t)G.ncell(), hipMemcpyDeviceToHost, st));
    GHIP(hipStreamSynchronize(st));
    return 0;
}

int stz_grf_field_device(const int N[3], const int prd[3], const stz_grf_params *g, const double *noise_dev,
                         double *chi_dev, void *stream) {
    Geom G;
    if (int rc = make_geom(N, prd, g, G)) return rc;
    if (!chi_dev) return gfail(STZMG_ERR_ARG, "null output");
    return grf_device(G, g, noise_dev, chi_dev, (hipStream_t)stream);
}

int stz_grf_time_kernels(const int N[3], const int prd[3], const stz_grf_params *g, int reps, double ms[4]) {
    Geom G;
    if (int rc = make_geom(N, prd, g, G)) return rc;
    if (!ms || reps < 1) return gfail(STZMG_ERR_ARG, "bad timing arguments");
    hipStream_t st = nullptr;
    GHIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct Guard {
        hipStream_t s;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Guard() {
            if (e0) hipEventDestroy(e0);
            if (e1) hipEventDestroy(e1);
            hipStreamDestroy(s);
        }
    } guard{st};
    GHIP(hipEventCreate(&guard.e0));
    GHIP(hipEventCreate(&guard.e1));
    Plan P;
    DevBuf chi;
    if (int rc = P.setup(G, g, true, st)) return rc;
    GHIP(hipMalloc((void **)&chi.p, sizeof(double) * (size_t)G.ncell()));
    for (int stage = 0; stage <= 3; stage++)
        if (int rc = P.enqueue(stage, P.noise.p, chi.p, st)) return rc;
    for (int stage = 0; stage <= 3; stage++) {
        GHIP(hipEventRecord(guard.e0, st));
        for (int q = 0; q < reps; q++)
            if (int rc = P.enqueue(stage, P.noise.p, chi.p, st)) return rc;
        GHIP(hipEventRecord(guard.e1, st));
        GHIP(hipEventSynchronize(guard.e1));
        float el = 0;
        GHIP(hipEventElapsedTime(&el, guard.e0, guard.e1));
        ms[stage] = el / reps;
    }
    return 0;
}

int stz_set_random_initial_conditions(stz_t s, const stz_grf_params *g, const double *noise, const char *cache_dir) {
    if (!s || !g) return STZMG_ERR_ARG;
    const int N[3] = {s->g.Nx, s->g.Ny, s->g.Nz};
    const int prd[3] = {1, 1, s->g.pz ? 1 : 0};
    Geom G;
    if (int rc = make_geom(N, prd, g, G)) return stz_detail::fail(s, rc, "random initial conditions: %s", g_grf_err.c_str());
    const long nc = G.ncell();
    std::string path;
    if (cache_dir) path = std::string(cache_dir) + "/" + cache_name(G, g);
    // an existing cache is loaded instead of convolving (as the reference does)
    if (!path.empty()) {
        if (FILE *f = fopen(path.c_str(), "rb")) {
            std::vector<double> h(nc);
            const size_t got = fread(h.data(), sizeof(double), (size_t)nc, f);
            fclose(f);
            if (got != (size_t)nc)
                return stz_detail::fail(s, STZMG_ERR_ARG, "cache %s holds %zu of %ld values", path.c_str(), got, nc);
            hipError_t e = hipMemcpyAsync(s->chi, h.data(), sizeof(double) * nc, hipMemcpyHostToDevice, s->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
            if (e != hipSuccess) return stz_detail::fail(s, STZMG_ERR_HIP, "chi upload: %s", hipGetErrorString(e));
            return 0;
        }
    }
    DevBuf nd;
    if (noise) {
        hipError_t e = hipMalloc((void **)&nd.p, sizeof(double) * (size_t)G.nnoise());
        if (e == hipSuccess)
            e = hipMemcpyAsync(nd.p, noise, sizeof(double) * (size_t)G.nnoise(), hipMemcpyHostToDevice, s->stream);
        if (e != hipSuccess) return stz_detail::fail(s, STZMG_ERR_HIP, "noise upload: %s", hipGetErrorString(e));
    }
    if (int rc = grf_device(G, g, nd.p, s->chi, s->stream))
        return stz_detail::fail(s, rc, "random initial conditions: %s", g_grf_err.c_str());
    if (!path.empty()) {
        std::vector<double> h(nc);
        hipError_t e = hipMemcpyAsync(h.data(), s->chi, sizeof(double) * nc, hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) return stz_detail::fail(s, STZMG_ERR_HIP, "chi download: %s", hipGetErrorString(e));
        FILE *f = fopen(path.c_str(), "wb");
        if (!f) return stz_detail::fail(s, STZMG_ERR_ARG, "cannot write %s: %s", path.c_str(), strerror(errno));
        const size_t put = fwrite(h.data(), sizeof(double), (size_t)nc, f);
        fclose(f);
        if (put != (size_t)nc) return stz_detail::fail(s, STZMG_ERR_ARG, "short write to %s", path.c_str());
    }
    return 0;
}

}  // extern "C"
