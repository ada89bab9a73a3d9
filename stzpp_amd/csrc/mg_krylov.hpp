// Vector kernels of the flexible GMRES solve (FP64 Krylov basis, FP32
// preconditioned vectors).  Vectors are 3 N doubles (SoA vec3 fields of
// level 0); a basis of nv vectors is stored contiguously, vector j at
// V + j * n3.  Dot products are formed as per-block partials and summed by
// k_reduce_cols in a fixed order, so results are reproducible.
#pragma once

#include "mg_kernels.hpp"

namespace stzmg {

constexpr int kKrylovMax = 32;       // largest restart length
constexpr int kDotBlocks = 1024;     // blocks of the grid-stride dot kernels

// part[b * nv + j] = sum over this block's share of w . V_j, j < nv.
__global__ __launch_bounds__(kThreads) void k_mdot(const double *__restrict__ w, const double *__restrict__ V,
                                                   int nv, long n3, double *__restrict__ part) {
    double acc[kKrylovMax + 1];
#pragma unroll
    for (int j = 0; j <= kKrylovMax; j++) acc[j] = 0;
    for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < n3; i += (long)gridDim.x * kThreads) {
        const double wi = w[i];
#pragma unroll
        for (int j = 0; j <= kKrylovMax; j++)
            if (j < nv) acc[j] += wi * V[(long)j * n3 + i];
    }
    __shared__ double red[kThreads / 64];
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

// w -= sum_j coef[j] V_j (j < nv), then part[b] = block share of |w|^2.
__global__ __launch_bounds__(kThreads) void k_maxpy_norm(double *__restrict__ w, const double *__restrict__ V,
                                                         const double *__restrict__ coef, int nv, long n3,
                                                         double *__restrict__ part) {
    double c[kKrylovMax + 1];
#pragma unroll
    for (int j = 0; j <= kKrylovMax; j++) c[j] = j < nv ? coef[j] : 0.0;
    double acc = 0;
    for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < n3; i += (long)gridDim.x * kThreads) {
        double wi = w[i];
#pragma unroll
        for (int j = 0; j <= kKrylovMax; j++)
            if (j < nv) wi -= c[j] * V[(long)j * n3 + i];
        w[i] = wi;
        acc += wi * wi;
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

// out[j] = sum_b part[b * nv + j] (one block per column, fixed order).
__global__ __launch_bounds__(kThreads) void k_reduce_cols(const double *__restrict__ part, int nblocks, int nv,
                                                          double *__restrict__ out) {
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
    if (threadIdx.x == 0) out[j] = red[0];
}

// dst = a src.
__global__ __launch_bounds__(kThreads) void k_scale(double *__restrict__ dst, const double *__restrict__ src,
                                                    double a, long n3) {
    const long i = (long)blockIdx.x * kThreads + threadIdx.x;
    if (i < n3) dst[i] = a * src[i];
}

// x += sum_j y[j] Z_j with FP32 Z (j < nv).
__global__ __launch_bounds__(kThreads) void k_xupdate(double *__restrict__ x, const float *__restrict__ Z,
                                                      const double *__restrict__ y, int nv, long n3) {
    const long i = (long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n3) return;
    double s = x[i];
    for (int j = 0; j < nv; j++) s += y[j] * (double)Z[(long)j * n3 + i];
    x[i] = s;
}

}  // namespace stzmg
