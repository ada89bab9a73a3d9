// Output of the quasi-static simulation in the reference's file formats
// (include/stz.h, stz_output_cube ... stz_write_files).
//
// * Cubes (output_cube, src/shear_sim_3d.cpp:1963-2161, with the values of
//   Field::fval, src/field.hh:144-178): one float32 per grid point, x
//   slowest, then y, z fastest, appended to <dir>/<name>_grid.<frame>; the
//   velocity and reference-map components carry Nz + 1 z points (the top
//   plate plane, or the periodic image of plane 0).  The state is SoA with x
//   fastest, so the
//   kernel transposes (z, x) tiles of one y row through LDS to write whole
//   z runs.
// * Wall tractions (compute_net_force, src/shear_sim_3d.cpp:1061-1126;
//   src/trans_sim_3d.cpp:1575-1660): the stress extrapolated to the top
//   wall, 1.5 s(Nz-1) - 0.5 s(Nz-2), and on the bottom 1.5 s(x) - 0.5 s(x+1)
//   (the reference pairs the bottom cell with its +x neighbour, grido + 1,
//   :1068); the transformed step reports the physical stress T s T^t.
//   traction.dat gets "%.16f" x 7 (time first) or, transformed, "%g" x 6;
//   the untransformed step also appends the six doubles to trac_grid.bin.
// * proj_mag.dat (:714-724): "%g %g" (time, sum of modsq of the projection's
//   stress increments), from the partial sums of the stress update.
#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/stz.h"
#include "stz_internal.hpp"

using namespace stz;

namespace {

constexpr int kCubeNames = 16;
const char *const kCubeName[kCubeNames] = {"u",   "v",   "w",   "s11", "s12", "s13",  "s22",  "s23",
                                           "s33", "p",   "dev", "chi", "tem", "xi_x", "xi_y", "xi_z"};

inline bool vertex_mode(int mode) { return mode <= 2 || mode >= 13; }

constexpr int kTile = 32;  // (z, x) transpose tile; 256 threads = 32 x 8

// The value of cube mode `mode` at cell c (modes 3-12), as Field::fval
// computes it (pressure and deviatoric norm without fused multiply-adds, so
// that they round as the reference's do).
__device__ __forceinline__ double cell_value(int mode, const double *__restrict__ sig, const double *__restrict__ chi,
                                             long Nc, long c) {
#pragma clang fp contract(off)
    if (mode <= 8) return sig[(mode - 3) * Nc + c];
    if (mode == 11) return chi[c];
    if (mode == 12) return chi[c] * 21000;
    const double s11 = sig[c], s22 = sig[3 * Nc + c], s33 = sig[5 * Nc + c];
    const double p = -(s11 + s22 + s33) * (1 / 3.);
    if (mode == 9) return p;
    const double s12 = sig[Nc + c], s13 = sig[2 * Nc + c], s23 = sig[4 * Nc + c];
    const double t11 = s11 + p, t22 = s22 + p, t33 = s33 + p;
    return sqrt(0.5 * (t11 * t11 + t22 * t22 + t33 * t33) + s12 * s12 + s13 * s13 + s23 * s23);
}

// out[(x Ny + y) nzo + z] = value(x, y, z) for one y per block row of tiles.
// Grid: (ceil(Nx / 32), Ny, ceil(nzo / 32)).  Vertex modes (velocity 0-2,
// reference map 13-15) have Nz + 1 planes; with z periodic plane Nz is the
// periodic image of plane 0, whose z reference map is shifted by lz.
__global__ __launch_bounds__(256) void k_cube(Grid g, int mode, int nzo, const double *__restrict__ vel,
                                              const double *__restrict__ sig, const double *__restrict__ chi,
                                              const double *__restrict__ xi, double lz, float *__restrict__ out) {
    __shared__ float tile[kTile][kTile + 1];
    const int x0 = blockIdx.x * kTile, y = blockIdx.y, z0 = blockIdx.z * kTile;
    const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;  // ty: 0..7
    for (int zz = ty; zz < kTile; zz += 8) {
        const int x = x0 + tx, z = z0 + zz;
        float v = 0;
        if (x < g.Nx && z < nzo) {
            if (mode <= 2) {
                const int zw = z < g.Nzv ? z : z - g.Nzv;  // periodic: plane Nz is plane 0
                v = (float)vel[mode * g.Nv + x + (long)g.Nx * (y + (long)g.Ny * zw)];
            } else if (mode >= 13) {
                const int zw = z < g.Nzv ? z : z - g.Nzv;
                const double r = xi[(mode - 13) * g.Nv + x + (long)g.Nx * (y + (long)g.Ny * zw)];
                v = (float)(mode == 15 && zw != z ? r + lz : r);
            } else {
                v = (float)cell_value(mode, sig, chi, g.Nc, x + (long)g.Nx * (y + (long)g.Ny * z));
            }
        }
        tile[zz][tx] = v;
    }
    __syncthreads();
    const int tz = threadIdx.x % kTile, tx2 = threadIdx.x / kTile;
    for (int xx = tx2; xx < kTile; xx += 8) {
        const int x = x0 + xx, z = z0 + tz;
        if (x < g.Nx && z < nzo) out[((long)x * g.Ny + y) * nzo + z] = tile[tz][xx];
    }
}

// Wall tractions: block partials part[6 b + q], q = top 13 23 33, bottom
// 13 23 33, over the Nx Ny wall columns (unscaled sums).
__global__ __launch_bounds__(256) void k_traction(Grid g, const double *__restrict__ sig, int trans, Xform X,
                                                  double *__restrict__ part) {
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    double v[6] = {0, 0, 0, 0, 0, 0};
    const long nxy = (long)g.Nx * g.Ny;
    if (q < nxy) {
        const int i = (int)(q % g.Nx), j = (int)(q / g.Nx);
        const long ct = q + nxy * (g.Nz - 1), cm = q + nxy * (g.Nz - 2);
        const long cb = q, cbp = (i + 1 < g.Nx ? i + 1 : 0) + (long)g.Nx * j;
        for (int w = 0; w < 2; w++) {
            const long a = w == 0 ? ct : cb, b = w == 0 ? cm : cbp;
            double e[6];
#pragma unroll
            for (int k = 0; k < 6; k++) e[k] = 1.5 * sig[k * g.Nc + a] - .5 * sig[k * g.Nc + b];
            if (!trans) {
                v[3 * w] = e[2];
                v[3 * w + 1] = e[4];
                v[3 * w + 2] = e[5];
            } else {
                const double M[9] = {e[0], e[1], e[2], e[1], e[3], e[4], e[2], e[4], e[5]};
                double Tt[9], tmp[9], P[9];
                m3t(X.T, Tt);
                m3mul(X.T, M, tmp);
                m3mul(tmp, Tt, P);
                v[3 * w] = P[2];
                v[3 * w + 1] = P[5];
                v[3 * w + 2] = P[8];
            }
        }
    }
    __shared__ double red[6][4];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        double t = v[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
        if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = t;
    }
    __syncthreads();
    if (threadIdx.x < 6) part[6 * (long)blockIdx.x + threadIdx.x] = red[threadIdx.x][0] + red[threadIdx.x][1] +
                                                                     red[threadIdx.x][2] + red[threadIdx.x][3];
}

int cube_size(const stz_t s, int mode, long *n) {
    if (mode < 0 || mode >= kCubeNames) return stz_detail::fail(s, STZMG_ERR_ARG, "cube mode %d: modes 0-15", mode);
    if (mode >= 13 && !s->xi)
        return stz_detail::fail(s, STZMG_ERR_STATE,
                                "cube mode %d: the reference map is kept by the direct step only (stz_init_ref)", mode);
    const Grid &g = s->g;
    *n = (long)g.Nx * g.Ny * (vertex_mode(mode) ? g.Nz + 1 : g.Nz);
    return 0;
}

#define OHIP(s, call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return stz_detail::fail((s), STZMG_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

int cube_host(stz_t s, int mode, std::vector<float> &h) {
    long n = 0;
    if (int rc = cube_size(s, mode, &n)) return rc;
    const Grid &g = s->g;
    const int nzo = (int)(n / ((long)g.Nx * g.Ny));
    float *d = reinterpret_cast<float *>(s->stage);  // 6 max(Nc, Nv) doubles: room for any cube
    hipLaunchKernelGGL(k_cube, dim3((g.Nx + kTile - 1) / kTile, g.Ny, (nzo + kTile - 1) / kTile), dim3(256), 0,
                       s->stream, g, mode, nzo, (const double *)s->vel, (const double *)s->sig,
                       (const double *)s->chi, (const double *)s->xi, s->p.bz - s->p.az, d);
    OHIP(s, hipGetLastError());
    h.resize(n);
    OHIP(s, hipMemcpyAsync(h.data(), d, sizeof(float) * n, hipMemcpyDeviceToHost, s->stream));
    OHIP(s, hipStreamSynchronize(s->stream));
    return 0;
}

int net_force(stz_t s, double out[6]) {
    const Grid &g = s->g;
    if (g.Nz < 2) return stz_detail::fail(s, STZMG_ERR_ARG, "net force needs two cell planes");
    const long nxy = (long)g.Nx * g.Ny;
    const unsigned nbk = (unsigned)((nxy + 255) / 256);
    const bool trans = s->p.transform != STZ_T_NONE;
    Xform X{};
    if (trans) X = stz_detail::xform(s, s->t);
    double *part = s->stage;  // 6 per block, well inside the staging buffer
    hipLaunchKernelGGL(k_traction, dim3(nbk), dim3(256), 0, s->stream, g, (const double *)s->sig, trans ? 1 : 0, X,
                       part);
    OHIP(s, hipGetLastError());
    std::vector<double> h(6 * (size_t)nbk);
    OHIP(s, hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s->stream));
    OHIP(s, hipStreamSynchronize(s->stream));
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (unsigned b = 0; b < nbk; b++)
        for (int q = 0; q < 6; q++) v[q] += h[6 * (size_t)b + q];
    const double dxdy = g.dx * g.dy, mdxdy = -g.dx * g.dy;
    for (int q = 0; q < 3; q++) out[q] = v[q] * dxdy;
    for (int q = 3; q < 6; q++) out[q] = v[q] * mdxdy;
    return 0;
}

int ensure_dir(stz_t s, const char *dir) {
    if (!dir || !*dir) return stz_detail::fail(s, STZMG_ERR_ARG, "no output directory");
    if (mkdir(dir, 0700) != 0 && errno != EEXIST)
        return stz_detail::fail(s, STZMG_ERR_ARG, "cannot create %s: %s", dir, strerror(errno));
    return 0;
}

int append(stz_t s, const std::string &path, const void *data, size_t bytes, const char *mode = "a") {
    FILE *f = fopen(path.c_str(), mode);
    if (!f) return stz_detail::fail(s, STZMG_ERR_ARG, "cannot open %s: %s", path.c_str(), strerror(errno));
    const size_t put = bytes ? fwrite(data, 1, bytes, f) : 0;
    const int rc = fclose(f);
    if (put != bytes || rc != 0) return stz_detail::fail(s, STZMG_ERR_ARG, "short write to %s", path.c_str());
    return 0;
}

int write_traction(stz_t s, const char *dir) {
    if (int rc = ensure_dir(s, dir)) return rc;
    double v[6];
    if (int rc = net_force(s, v)) return rc;
    char line[512];
    const std::string d(dir);
    if (s->p.transform != STZ_T_NONE) {
        snprintf(line, sizeof(line), "%g\t%g\t%g\t%g\t%g\t%g\n", v[0], v[1], v[2], v[3], v[4], v[5]);
        return append(s, d + "/traction.dat", line, strlen(line));
    }
    snprintf(line, sizeof(line), "%.16f\t%.16f\t%.16f\t%.16f\t%.16f\t%.16f\t%.16f\n", s->t, v[0], v[1], v[2], v[3],
             v[4], v[5]);
    if (int rc = append(s, d + "/traction.dat", line, strlen(line))) return rc;
    return append(s, d + "/trac_grid.bin", v, sizeof(v));
}

int write_cube(stz_t s, const char *dir, const char *name, int mode, int frame) {
    if (int rc = ensure_dir(s, dir)) return rc;
    std::vector<float> h;
    if (int rc = cube_host(s, mode, h)) return rc;
    char path[4096];
    snprintf(path, sizeof(path), "%s/%s_grid.%d", dir, name, frame);
    return append(s, path, h.data(), sizeof(float) * h.size());
}

}  // namespace

namespace stz_detail {

int step_output(stz_t s) {
    if (!s->out_flags) return 0;
    const char *dir = s->out_dir.c_str();
    if ((s->out_flags & STZ_OUT_PROJ_MAG) && s->p.transform == STZ_T_NONE) {
        if (int rc = ensure_dir(s, dir)) return rc;
        char line[128];
        snprintf(line, sizeof(line), "%g %g\n", s->t, s->proj_mag);
        // the reference opens proj_mag.dat with "w" on its first line
        if (int rc = append(s, s->out_dir + "/proj_mag.dat", line, strlen(line), s->pm_writes ? "a" : "w"))
            return rc;
        s->pm_writes++;
    }
    if (s->out_flags & STZ_OUT_TRACTION) return write_traction(s, dir);
    return 0;
}

}  // namespace stz_detail

extern "C" {

int stz_output_cube(stz_t s, int mode, float *out, long long n) {
    if (!s) return STZMG_ERR_ARG;
    long need = 0;
    if (int rc = cube_size(s, mode, &need)) return rc;
    if (!out) return (int)STZMG_ERR_ARG;
    if (n < need) return stz_detail::fail(s, STZMG_ERR_ARG, "cube needs %ld values, buffer holds %lld", need, n);
    std::vector<float> h;
    if (int rc = cube_host(s, mode, h)) return rc;
    memcpy(out, h.data(), sizeof(float) * h.size());
    return 0;
}

long long stz_cube_size(stz_t s, int mode) {
    long n = 0;
    if (!s || cube_size(s, mode, &n)) return -1;
    return n;
}

int stz_write_cube(stz_t s, const char *dir, int mode, int frame) {
    if (!s) return STZMG_ERR_ARG;
    if (mode < 0 || mode >= kCubeNames) {
        long n;
        return cube_size(s, mode, &n);
    }
    return write_cube(s, dir, kCubeName[mode], mode, frame);
}

int stz_net_force(stz_t s, double out[6]) {
    if (!s || !out) return STZMG_ERR_ARG;
    return net_force(s, out);
}

int stz_write_traction(stz_t s, const char *dir) {
    if (!s) return STZMG_ERR_ARG;
    return write_traction(s, dir);
}

double stz_proj_mag(stz_t s) { return s ? s->proj_mag : 0.0; }

int stz_set_output(stz_t s, const char *dir, int flags) {
    if (!s) return STZMG_ERR_ARG;
    if (flags & ~(STZ_OUT_TRACTION | STZ_OUT_PROJ_MAG)) return stz_detail::fail(s, STZMG_ERR_ARG, "unknown flags");
    if (flags) {
        if (int rc = ensure_dir(s, dir)) return rc;
        s->out_dir = dir;
    }
    s->out_flags = flags;
    return 0;
}

int stz_output_time_kernel(stz_t s, int mode, int reps, double *ms) {
    if (!s || reps < 1 || !ms) return STZMG_ERR_ARG;
    long n = 0;
    if (int rc = cube_size(s, mode, &n)) return rc;
    const Grid &g = s->g;
    const int nzo = (int)(n / ((long)g.Nx * g.Ny));
    float *d = reinterpret_cast<float *>(s->stage);
    const dim3 grid((g.Nx + kTile - 1) / kTile, g.Ny, (nzo + kTile - 1) / kTile);
    hipEvent_t e0, e1;
    OHIP(s, hipEventCreate(&e0));
    OHIP(s, hipEventCreate(&e1));
    auto launch = [&] {
        hipLaunchKernelGGL(k_cube, grid, dim3(256), 0, s->stream, g, mode, nzo, (const double *)s->vel,
                           (const double *)s->sig, (const double *)s->chi, (const double *)s->xi,
                           s->p.bz - s->p.az, d);
    };
    launch();  // warm
    hipEventRecord(e0, s->stream);
    for (int r = 0; r < reps; r++) launch();
    hipEventRecord(e1, s->stream);
    hipError_t e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipGetLastError();
    float t = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (e != hipSuccess) return stz_detail::fail(s, STZMG_ERR_HIP, "cube kernel timing: %s", hipGetErrorString(e));
    *ms = (double)t / reps;
    return 0;
}

int stz_write_files(stz_t s, const char *dir, int frame, int n_tp, int fflags) {
    if (!s) return STZMG_ERR_ARG;
    if (n_tp < 10 || frame < 0)
        return stz_detail::fail(s, STZMG_ERR_ARG, "write_files needs n_tp >= 10 (frames every n_tp/10) and frame >= 0");
    if (fflags >> kCubeNames) return stz_detail::fail(s, STZMG_ERR_ARG, "fflags: unknown bits above 15");
    if ((fflags >> 13) && !s->xi)
        return stz_detail::fail(s, STZMG_ERR_STATE,
                                "fflags: bits 13-15 select the reference map, kept by the direct step only");
    // write_files (src/shear_sim_3d.cpp:1938-1960): bit `mode` of fflags, the
    // velocity and stress every n_tp/10 frames, chi when (n_tp/10) % frame
    // == 0 for frame > 0 (the reference's order of the operands)
    const int every = n_tp / 10;
    for (int mode = 0; mode < kCubeNames; mode++) {
        bool on = (fflags >> mode) & 1;
        if (mode <= 8) on = on || frame % every == 0;
        if (mode == 11) on = on || (frame > 0 && every % frame == 0);
        if (on)
            if (int rc = write_cube(s, dir, kCubeName[mode], mode, frame)) return rc;
    }
    return 0;
}

}  // extern "C"
