// Vector kernels of the GMRES solves: the flexible GMRES of
// STZMG_SOLVE_FP32_GMRES (FP64 Krylov basis, FP32 preconditioned vectors)
// and the FP32 inner GMRES of STZMG_SOLVE_FP32_GMRES_IR (FP32 basis).
// Vectors are 3 N values of type TV (SoA vec3 fields of level 0); a basis
// of nv vectors is stored contiguously, vector j at V + j * n3.  Sums are
// formed in FP64 as per-block partials and summed by k_reduce_cols in a fixed
// order, so results are reproducible.
#pragma once

#include "mg_kernels.hpp"

namespace stzmg {

constexpr int kKrylovMax = 16;       // largest restart length
constexpr int kDotBlocks = 1024;     // blocks of the grid-stride dot kernels

// Grid-stride steps per pass of the multi-vector kernels (FP32 vectors: 4,
// FP64: 2; bounded by the registers of the 17-vector instances).
template <typename TV>
struct KrylovUnroll {
    static constexpr int U = sizeof(TV) == 4 ? 4 : 2;
};

// part[b * NV + j] = sum over this block's share of w . V_j, j < NV.
template <int NV, typename TV = double>
__global__ __launch_bounds__(kThreads) void k_mdot(const TV *__restrict__ w, const TV *__restrict__ V, long n3,
                                                   double *__restrict__ part) {
    constexpr int nv = NV;
    double acc[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) acc[j] = 0;
    // U stride steps per pass with every load issued first (more loads in
    // flight per wave); the terms are added in the order of the plain
    // grid-stride loop, so the sums are the same
    constexpr int U = KrylovUnroll<TV>::U;
    const long S = (long)gridDim.x * kThreads;
    long i = (long)blockIdx.x * kThreads + threadIdx.x;
    for (; i + (U - 1) * S < n3; i += U * S) {
        TV wv[U], vv[U][NV];
#pragma unroll
        for (int u = 0; u < U; u++) {
            wv[u] = w[i + u * S];
#pragma unroll
            for (int j = 0; j < NV; j++) vv[u][j] = V[(long)j * n3 + i + u * S];
        }
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int j = 0; j < NV; j++) acc[j] += (double)wv[u] * (double)vv[u][j];
    }
    for (; i < n3; i += S) {
        const double wi = (double)w[i];
#pragma unroll
        for (int j = 0; j < NV; j++) acc[j] += wi * (double)V[(long)j * n3 + i];
    }
    __shared__ double red[kThreads / 64];
#pragma unroll
    for (int j = 0; j < nv; j++) {
        double v = acc[j];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0;
            for (int q = 0; q < kThreads / 64; q++) t += red[q];
            part[(long)blockIdx.x * nv + j] = t;
        }
        __syncthreads();
    }
}

// w -= sum_j coef[j] V_j (j < NV) in FP64, stored as TV, then part[b] =
// block share of |w|^2 (of the stored values).
template <int NV, typename TV = double>
__global__ __launch_bounds__(kThreads) void k_maxpy_norm(TV *__restrict__ w, const TV *__restrict__ V,
                                                         const double *__restrict__ coef, long n3,
                                                         double *__restrict__ part) {
    double c[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) c[j] = coef[j];
    double acc = 0;
    // U stride steps per pass, loads first, terms in the plain loop's order
    constexpr int U = KrylovUnroll<TV>::U;
    const long S = (long)gridDim.x * kThreads;
    long i = (long)blockIdx.x * kThreads + threadIdx.x;
    for (; i + (U - 1) * S < n3; i += U * S) {
        TV wv[U], vv[U][NV];
#pragma unroll
        for (int u = 0; u < U; u++) {
            wv[u] = w[i + u * S];
#pragma unroll
            for (int j = 0; j < NV; j++) vv[u][j] = V[(long)j * n3 + i + u * S];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            double wi = (double)wv[u];
#pragma unroll
            for (int j = 0; j < NV; j++) wi -= c[j] * (double)vv[u][j];
            const TV ws = (TV)wi;
            w[i + u * S] = ws;
            acc += (double)ws * (double)ws;
        }
    }
    for (; i < n3; i += S) {
        double wi = (double)w[i];
#pragma unroll
        for (int j = 0; j < NV; j++) wi -= c[j] * (double)V[(long)j * n3 + i];
        const TV ws = (TV)wi;
        w[i] = ws;
        acc += (double)ws * (double)ws;
    }
    __shared__ double red[kThreads / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int q = 0; q < kThreads / 64; q++) t += red[q];
        part[blockIdx.x] = t;
    }
}

// Launch of the NV-vector kernels for a run-time count nv (1..kKrylovMax+1).
template <int NV, typename TV = double>
void launch_mdot_n(hipStream_t st, const TV *w, const TV *V, int nv, long n3, double *part) {
    if constexpr (NV <= kKrylovMax + 1) {
        if (nv == NV) {
            hipLaunchKernelGGL((k_mdot<NV, TV>), dim3(kDotBlocks), dim3(kThreads), 0, st, w, V, n3, part);
            return;
        }
        launch_mdot_n<NV + 1, TV>(st, w, V, nv, n3, part);
    }
}
template <int NV, typename TV = double>
void launch_maxpy_n(hipStream_t st, TV *w, const TV *V, const double *coef, int nv, long n3, double *part) {
    if constexpr (NV <= kKrylovMax + 1) {
        if (nv == NV) {
            hipLaunchKernelGGL((k_maxpy_norm<NV, TV>), dim3(kDotBlocks), dim3(kThreads), 0, st, w, V, coef, n3,
                               part);
            return;
        }
        launch_maxpy_n<NV + 1, TV>(st, w, V, coef, nv, n3, part);
    }
}

// out[j] = sum_b part[b * nv + j] (one block per column, fixed order),
// times scale[j] when scale is given.
__global__ __launch_bounds__(kThreads) void k_reduce_cols(const double *__restrict__ part, int nblocks, int nv,
                                                          double *__restrict__ out,
                                                          const double *__restrict__ scale) {
    const int j = blockIdx.x;
    double t = 0;
    for (int b = threadIdx.x; b < nblocks; b += kThreads) t += part[(long)b * nv + j];
    __shared__ double red[kThreads];
    red[threadIdx.x] = t;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = scale ? red[0] * scale[j] : red[0];
}

// dst = a src (in FP64, stored as TD; kEW elements per thread, grid
// nblk_ew(n3)).
template <typename TD = double, typename TS = double>
__global__ __launch_bounds__(kThreads) void k_scale(TD *__restrict__ dst, const TS *__restrict__ src, double a,
                                                    long n3) {
    const long b0 = (long)blockIdx.x * (kThreads * kEW) + threadIdx.x;
    TS v[kEW];
#pragma unroll
    for (int u = 0; u < kEW; u++) {
        const long i = b0 + (long)u * kThreads;
        v[u] = i < n3 ? src[i] : TS(0);
    }
#pragma unroll
    for (int u = 0; u < kEW; u++) {
        const long i = b0 + (long)u * kThreads;
        if (i < n3) dst[i] = (TD)(a * (double)v[u]);
    }
}

// x += sum_j y[j] Z_j (j < nv; kEW elements per thread, grid nblk_ew(n3)).
// The loads of one basis vector for all kEW elements are issued together;
// each element's terms are added in the order j = 0, 1, ...
template <typename TZ = float>
__global__ __launch_bounds__(kThreads) void k_xupdate(double *__restrict__ x, const TZ *__restrict__ Z,
                                                      const double *__restrict__ y, int nv, long n3) {
    const long b0 = (long)blockIdx.x * (kThreads * kEW) + threadIdx.x;
    double s[kEW];
#pragma unroll
    for (int u = 0; u < kEW; u++) {
        const long i = b0 + (long)u * kThreads;
        s[u] = i < n3 ? x[i] : 0.0;
    }
    for (int j = 0; j < nv; j++) {
        const double yj = y[j];
        const TZ *Zj = Z + (long)j * n3;
        TZ z[kEW];
#pragma unroll
        for (int u = 0; u < kEW; u++) {
            const long i = b0 + (long)u * kThreads;
            z[u] = i < n3 ? Zj[i] : TZ(0);
        }
#pragma unroll
        for (int u = 0; u < kEW; u++) s[u] += yj * (double)z[u];
    }
#pragma unroll
    for (int u = 0; u < kEW; u++) {
        const long i = b0 + (long)u * kThreads;
        if (i < n3) x[i] = s[u];
    }
}

}  // namespace stzmg
