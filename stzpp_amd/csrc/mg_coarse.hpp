// Coarse end of the V-cycle in one workgroup.
//
// Levels of a few thousand to ~33k points give a colour pass too little work
// to fill the GPU, and the reference schedule spends many sweeps there (24 on
// the coarsest level, v+l on level l).  Launched kernel by kernel, each of
// those passes costs a full dispatch.  k_coarse_cycle runs the whole
// sub-cycle below a given level -- smoothing, residual, restriction, the
// bottom solve, interpolation -- inside one 1024-thread workgroup, with
// __syncthreads() between dependent passes.  The operations and their order
// are those of the multi-launch path, so the results are the same.
#pragma once

#include <type_traits>

#include "mg_kernels.hpp"

namespace stzmg {

constexpr int kCoarseThreads = 1024;
constexpr int kMaxCoarse = 8;
// Levels with at most this many points go to the single-workgroup kernel.
// Measured at 256^3 and 128^3 (V-cycle time, MI355X): one workgroup wins on
// the coarsest level (<= 1024 points by construction), breaks even at 16^3
// and loses at 32^3, where one CU's memory latency dominates.
constexpr long kCoarseMaxPoints = 1024;

template <typename T>
struct CoarseLevel {
    Dims d;
    T *x, *r, *t;
    const T *ops, *inv;
    double *x64, *r64;          // FP64 fields (bottom solve of an FP32 cycle)
    const double *ops64, *inv64;
    int down, up;               // sweeps before restriction / after interpolation
};

template <typename T>
struct CoarseArgs {
    int nl;                     // levels handled (relative 0 .. nl-1)
    int bottom;                 // sweeps on the last one
    int table_vals;             // doubles of dynamic LDS before the bottom fields;
                                // -1: the bottom fields stay in global memory
    CoarseLevel<T> L[kMaxCoarse];
};

// Per-plane operator tables (and centre inverses) of a level are staged in
// LDS: points of different planes share a wavefront here, so the table
// cannot come through the scalar cache as in the one-plane-per-block kernels.
// Per-point tables stay in global memory.
template <typename T, int MODE>
__device__ void stage_tables(const Dims &d, const T *ops, const T *inv, T *lds, const T **ops_s,
                             const T **inv_s) {
    if (MODE != 0) {
        *ops_s = ops;
        *inv_s = inv;
        return;
    }
    const int na = d.o * kBlk, ni = d.o * 9;
    for (int q = threadIdx.x; q < na; q += blockDim.x) lds[q] = ops[q];
    for (int q = threadIdx.x; q < ni; q += blockDim.x) lds[na + q] = inv[q];
    __syncthreads();
    *ops_s = lds;
    *inv_s = lds + na;
}

template <typename T, int MODE>
__device__ void wg_gs_sweep(const Dims &d, T *x, const T *r, const T *ops, const T *inv) {
    const int pxo = d.px && (d.m & 1), pyo = d.py && (d.n & 1), pzo = d.pz && (d.o & 1);
    for (int cz = 0; cz < 2 + pzo; cz++)
        for (int cy = 0; cy < 2 + pyo; cy++)
            for (int cx = 0; cx < 2 + pxo; cx++) {
                const int cntx = col_count(d.m, cx, d.px), cnty = col_count(d.n, cy, d.py),
                          cntz = col_count(d.o, cz, d.pz);
                const int tot = cntx * cnty * cntz;
                for (int t = threadIdx.x; t < tot; t += blockDim.x) {
                    const int ii = t % cntx, rest = t / cntx;
                    const int jj = rest % cnty, kk = rest / cnty;
                    gs_point<T, MODE>(d, x, r, ops, inv, col_coord(ii, cx, d.m), col_coord(jj, cy, d.n),
                                      col_coord(kk, cz, d.o));
                }
                __syncthreads();
            }
}

// Sum over a row of 16 lanes, left in every lane: DPP row permutes (quad
// swaps, then the half-row and row mirrors) instead of LDS-routed shuffles.
// All 64 lanes of the wave must be active.
__device__ __forceinline__ double row16_sum(double v) {
    auto step = [](double x, auto ctrl) {
        constexpr int C = decltype(ctrl)::value;
        const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), C, 0xF, 0xF, false);
        const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), C, 0xF, 0xF, false);
        return x + __hiloint2double(hi, lo);
    };
    v = step(v, std::integral_constant<int, 0xB1>{});   // quad_perm [1, 0, 3, 2]
    v = step(v, std::integral_constant<int, 0x4E>{});   // quad_perm [2, 3, 0, 1]
    v = step(v, std::integral_constant<int, 0x141>{});  // row_half_mirror
    v = step(v, std::integral_constant<int, 0x140>{});  // row_mirror
    return v;
}
__device__ __forceinline__ float row16_sum(float v) {
    auto step = [](float x, auto ctrl) {
        constexpr int C = decltype(ctrl)::value;
        return x + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), C, 0xF, 0xF, false));
    };
    v = step(v, std::integral_constant<int, 0xB1>{});
    v = step(v, std::integral_constant<int, 0x4E>{});
    v = step(v, std::integral_constant<int, 0x141>{});
    v = step(v, std::integral_constant<int, 0x140>{});
    return v;
}

// Bottom solve: nsweeps colour sweeps of one level with x and r held in
// LDS (xs, rs: 3 N values each) and 16 lanes per point.  Each lane sums up to
// two of the 26 off-centre blocks, the 16 partial sums are added across the
// lane row (row16_sum), and the first lane of the group applies the inverse
// centre block.  Waves without a point in a pass skip it.  Colour order and
// per-point update are those of wg_gs_sweep; only the summation order inside
// a point differs.  On exit x holds the result.
constexpr int kBottomLanes = 16;
static_assert(kBottomLanes == 16, "row16_sum reduces rows of 16 lanes");
template <typename T, int MODE>
__device__ void wg_bottom_sweeps(const Dims &d, const T *ops, const T *inv, int nsweeps, T *xs, const T *rs) {
    const long N = d.N;
    const int sl = threadIdx.x % kBottomLanes, grp = threadIdx.x / kBottomLanes;
    const int ngrp = blockDim.x / kBottomLanes;
    const int pxo = d.px && (d.m & 1), pyo = d.py && (d.n & 1), pzo = d.pz && (d.o & 1);
    for (int sw = 0; sw < nsweeps; sw++)
        for (int cz = 0; cz < 2 + pzo; cz++)
            for (int cy = 0; cy < 2 + pyo; cy++)
                for (int cx = 0; cx < 2 + pxo; cx++) {
                    const int cntx = col_count(d.m, cx, d.px), cnty = col_count(d.n, cy, d.py),
                              cntz = col_count(d.o, cz, d.pz);
                    const int tot = cntx * cnty * cntz;
                    for (int base = 0; base < tot; base += ngrp) {
                        const int t = base + grp;
                        // first group of this wave (uniform): a wave past the
                        // pass's points has nothing to sum
                        if (base + (int)(threadIdx.x / 64) * (64 / kBottomLanes) >= tot) continue;
                        T a0 = 0, a1 = 0, a2 = 0;
                        long p = 0, c = 0;
                        if (t < tot) {
                            const int ii = t % cntx, rest = t / cntx;
                            const int jj = rest % cnty, kk = rest / cnty;
                            const int i = col_coord(ii, cx, d.m), j = col_coord(jj, cy, d.n), k = col_coord(kk, cz, d.o);
                            p = i + (long)d.m * (j + (long)d.n * k);
                            c = cls_of<MODE>(p, k);
                            const T *A = ops + c * kBlk;
#pragma unroll
                            for (int h2 = 0; h2 < 2; h2++) {
                                const int b = sl + kBottomLanes * h2;
                                if (b >= 27 || b == 13) continue;
                                const int ni = nb(i, b % 3 - 1, d.m, d.px), nj = nb(j, (b / 3) % 3 - 1, d.n, d.py),
                                          nk = nb(k, b / 9 - 1, d.o, d.pz);
                                if (ni < 0 || nj < 0 || nk < 0) continue;
                                const long q = ni + (long)d.m * (nj + (long)d.n * nk);
                                mv(A + 9 * b, xs[q], xs[q + N], xs[q + 2 * N], a0, a1, a2);
                            }
                        }
                        a0 = row16_sum(a0);
                        a1 = row16_sum(a1);
                        a2 = row16_sum(a2);
                        if (t < tot && sl == 0) {
                            const T e0 = rs[p] - a0, e1 = rs[p + N] - a1, e2 = rs[p + 2 * N] - a2;
                            const T *B = inv + c * 9;
                            xs[p] = B[0] * e0 + B[1] * e1 + B[2] * e2;
                            xs[p + N] = B[3] * e0 + B[4] * e1 + B[5] * e2;
                            xs[p + 2 * N] = B[6] * e0 + B[7] * e1 + B[8] * e2;
                        }
                    }
                    __syncthreads();
                }
}

template <typename T, int MODE>
__device__ void wg_bottom_solve(const Dims &d, T *x, const T *r, const T *ops, const T *inv, int nsweeps, T *xs,
                                T *rs) {
    const int n3 = 3 * (int)d.N;
    for (int q = threadIdx.x; q < n3; q += blockDim.x) {
        xs[q] = x[q];
        rs[q] = r[q];
    }
    __syncthreads();
    wg_bottom_sweeps<T, MODE>(d, ops, inv, nsweeps, xs, rs);
    for (int q = threadIdx.x; q < n3; q += blockDim.x) x[q] = xs[q];
    __syncthreads();
}

// The coarsest level is always entered with x = 0 (the restriction clears
// it), and nsweeps colour sweeps from x = 0 are a fixed linear map of r:
// x = Q r with Q of order 3 N.  k_bottom_map computes column j of Q -- the
// sweeps applied to r = e_j -- in workgroup j with the same per-point update
// as wg_bottom_solve, and stores it at q[i 3N + j]; k_bottom_apply then
// replaces the single-workgroup bottom solve of every V-cycle by a
// matrix-vector product spread over the GPU.  Dynamic LDS: the staged
// per-plane tables (MODE 0; table_vals doubles) and x, r of the level.
template <int MODE>
__global__ __launch_bounds__(kCoarseThreads) void k_bottom_map(Dims d, const double *ops, const double *inv,
                                                               int nsweeps, int table_vals, double *q) {
    extern __shared__ double coarse_lds[];
    const double *ops_s, *inv_s;
    stage_tables<double, MODE>(d, ops, inv, coarse_lds, &ops_s, &inv_s);
    const int n3 = 3 * (int)d.N;
    double *xs = coarse_lds + table_vals, *rs = xs + n3;
    const int j = blockIdx.x;
    for (int t = threadIdx.x; t < n3; t += blockDim.x) {
        xs[t] = 0.0;
        rs[t] = t == j ? 1.0 : 0.0;
    }
    __syncthreads();
    wg_bottom_sweeps<double, MODE>(d, ops_s, inv_s, nsweeps, xs, rs);
    for (int i = threadIdx.x; i < n3; i += blockDim.x) q[(long)i * n3 + j] = xs[i];
}

// x = Q r on the coarsest level (3 N values of type T, component-major):
// one wave per row, FP64 sums.  Grid: ceil(3 N / 4) workgroups of 256.
constexpr int kApplyThreads = 256;
template <typename T>
__global__ __launch_bounds__(kApplyThreads) void k_bottom_apply(const double *__restrict__ q,
                                                                const T *__restrict__ r, T *__restrict__ x,
                                                                int n3) {
    const int row = blockIdx.x * (kApplyThreads / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= n3) return;  // whole waves leave together
    const double *qr = q + (long)row * n3;
    double s = 0;
    for (int j = lane; j < n3; j += 64) s += qr[j] * (double)r[j];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) x[row] = (T)s;
}

template <typename T, int MODE>
__device__ void wg_resid(const Dims &d, const T *x, const T *r, const T *ops, T *out) {
    const int mn = d.m * d.n;
    for (int p = threadIdx.x; p < (int)d.N; p += blockDim.x) {
        const int k = p / mn, rem = p - k * mn, j = rem / d.m, i = rem - j * d.m;
        const T *A = ops + cls_of<MODE>(p, k) * kBlk;
        T y0 = 0, y1 = 0, y2 = 0;
        stencil_sum<T, false>(d, x, A, i, j, k, y0, y1, y2);
        out[p] = r[p] - y0;
        out[p + d.N] = r[p + d.N] - y1;
        out[p + 2 * d.N] = r[p + 2 * d.N] - y2;
    }
    __syncthreads();
}

template <typename T>
__device__ void wg_restrict(const Dims &F, const Dims &C, const T *res, T *rc, T *xc) {
    const int mn = C.m * C.n;
    for (int p = threadIdx.x; p < (int)C.N; p += blockDim.x) {
        const int K = p / mn, rem = p - K * mn, J = rem / C.m, I = rem - J * C.m;
        restrict_point<T>(F, C, res, rc, xc, I, J, K);
    }
    __syncthreads();
}

template <typename T>
__device__ void wg_interp(const Dims &F, const Dims &C, const T *xc, T *xf) {
    const int mn = F.m * F.n;
    for (int p = threadIdx.x; p < (int)F.N; p += blockDim.x) {
        const int k = p / mn, rem = p - k * mn, j = rem / F.m, i = rem - j * F.m;
        interp_point<T>(F, C, xc, xf, i, j, k);
    }
    __syncthreads();
}

template <typename A, typename B>
__device__ void wg_convert(const A *a, B *b, long n) {
    for (long q = threadIdx.x; q < n; q += blockDim.x) b[q] = (B)a[q];
    __syncthreads();
}

// Dynamic LDS: the largest staged table, (243 + 9) * o_max values of the
// widest type used (double, none for per-point tables), then x and r of the
// bottom level (a.table_vals >= 0), given at launch.
template <typename T, int MODE>
__global__ __launch_bounds__(kCoarseThreads) void k_coarse_cycle(CoarseArgs<T> a) {
    extern __shared__ double coarse_lds[];
    T *lds = reinterpret_cast<T *>(coarse_lds);
    const T *ops_s, *inv_s;
    for (int i = 0; i < a.nl - 1; i++) {
        CoarseLevel<T> &L = a.L[i];
        stage_tables<T, MODE>(L.d, L.ops, L.inv, lds, &ops_s, &inv_s);
        for (int q = 0; q < L.down; q++) wg_gs_sweep<T, MODE>(L.d, L.x, L.r, ops_s, inv_s);
        wg_resid<T, MODE>(L.d, L.x, L.r, ops_s, L.t);
        wg_restrict<T>(L.d, a.L[i + 1].d, L.t, a.L[i + 1].r, a.L[i + 1].x);
    }
    CoarseLevel<T> &B = a.L[a.nl - 1];
    if (sizeof(T) == 4 && B.x64 != nullptr) {
        // The coarsest solve of an FP32 cycle runs in FP64.
        const long n3 = 3 * B.d.N;
        wg_convert<T, double>(B.r, B.r64, n3);
        for (long q = threadIdx.x; q < n3; q += blockDim.x) B.x64[q] = 0.0;
        __syncthreads();
        const double *ops_d, *inv_d;
        stage_tables<double, MODE>(B.d, B.ops64, B.inv64, coarse_lds, &ops_d, &inv_d);
        if (a.table_vals >= 0) {
            double *xs = coarse_lds + a.table_vals;
            wg_bottom_solve<double, MODE>(B.d, B.x64, B.r64, ops_d, inv_d, a.bottom, xs, xs + n3);
        } else {
            for (int q = 0; q < a.bottom; q++) wg_gs_sweep<double, MODE>(B.d, B.x64, B.r64, ops_d, inv_d);
        }
        wg_convert<double, T>(B.x64, B.x, n3);
    } else {
        stage_tables<T, MODE>(B.d, B.ops, B.inv, lds, &ops_s, &inv_s);
        if (a.table_vals >= 0) {
            T *xs = reinterpret_cast<T *>(coarse_lds + a.table_vals);
            wg_bottom_solve<T, MODE>(B.d, B.x, B.r, ops_s, inv_s, a.bottom, xs, xs + 3 * B.d.N);
        } else {
            for (int q = 0; q < a.bottom; q++) wg_gs_sweep<T, MODE>(B.d, B.x, B.r, ops_s, inv_s);
        }
    }
    for (int i = a.nl - 1; i > 0; i--) {
        CoarseLevel<T> &F = a.L[i - 1];
        wg_interp<T>(F.d, a.L[i].d, a.L[i].x, F.x);
        stage_tables<T, MODE>(F.d, F.ops, F.inv, lds, &ops_s, &inv_s);
        for (int q = 0; q < F.up; q++) wg_gs_sweep<T, MODE>(F.d, F.x, F.r, ops_s, inv_s);
    }
}

// LDS bytes of the staged per-plane tables for levels with at most omax
// planes, and of the bottom fields (x and r of at most kCoarseMaxPoints
// points, in double).
inline size_t coarse_lds_bytes(int omax) { return sizeof(double) * (size_t)(kBlk + 9) * omax; }
inline size_t bottom_lds_bytes(long n) { return sizeof(double) * 6 * (size_t)n; }

}  // namespace stzmg
