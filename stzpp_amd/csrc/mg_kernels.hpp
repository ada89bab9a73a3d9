// Device kernels of the multigrid projection solver (gfx950, HIP).
//
// Data layout in HBM, per level:
//   vec3 fields   SoA: component c of point p at f[c*N + p], p = i + m*(j + n*k)
//                 (x fastest, so a wavefront covers one x-row segment of one
//                 z plane and a per-plane stencil is wavefront-uniform)
//   operator      ops[cls*243 + 9*b + e]: 27 blocks of 3x3 per class;
//                 cls = k (PLANE) or p (POINT)
//   inverse       inv[cls*9 + e]: inverse of the centre block, precomputed
//                 once per operator (the reference recomputes it per point and
//                 sweep, region.cc:1708 via mg3d_inverse)
//
// Kernels are templated on the arithmetic type T (double or float) and on the
// class mode, so the FP32 smoother and the FP64 outer loop share one source.
#pragma once

#include <hip/hip_runtime.h>

namespace stzmg {

constexpr int kBlk = 243;
constexpr int kThreads = 256;

struct Dims {
    int m, n, o;
    int px, py, pz;
    long N;
};

// Neighbour coordinate with periodic wrap; -1 marks a point outside a
// non-periodic dimension (its stencil block is zero by construction).
__device__ __forceinline__ int nb(int i, int di, int n, int prd) {
    int q = i + di;
    if (q < 0) return prd ? q + n : -1;
    if (q >= n) return prd ? q - n : -1;
    return q;
}

template <int MODE>
__device__ __forceinline__ long cls_of(long p, int k) {
    return MODE == 0 ? (long)k : p;
}

// 3x3 block times vec3, accumulated.
template <typename T>
__device__ __forceinline__ void mv(const T *__restrict__ a, T x0, T x1, T x2, T &y0, T &y1, T &y2) {
    y0 += a[0] * x0 + a[1] * x1 + a[2] * x2;
    y1 += a[3] * x0 + a[4] * x1 + a[5] * x2;
    y2 += a[6] * x0 + a[7] * x1 + a[8] * x2;
}

// Sum over the 27 (or 26 without the centre) blocks of A_b x_{p+off(b)}.
template <typename T, bool SKIP_CENTRE>
__device__ __forceinline__ void stencil_sum(const Dims &d, const T *__restrict__ x,
                                            const T *__restrict__ A, int i, int j, int k,
                                            T &y0, T &y1, T &y2) {
    const long N = d.N;
    int ix[3] = {nb(i, -1, d.m, d.px), i, nb(i, 1, d.m, d.px)};
    int jx[3] = {nb(j, -1, d.n, d.py), j, nb(j, 1, d.n, d.py)};
    int kx[3] = {nb(k, -1, d.o, d.pz), k, nb(k, 1, d.o, d.pz)};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        if (kx[c] < 0) continue;
#pragma unroll
        for (int b = 0; b < 3; b++) {
            if (jx[b] < 0) continue;
            const long row = (long)d.m * (jx[b] + (long)d.n * kx[c]);
#pragma unroll
            for (int a = 0; a < 3; a++) {
                if (SKIP_CENTRE && a == 1 && b == 1 && c == 1) continue;
                if (ix[a] < 0) continue;
                const long q = row + ix[a];
                mv(A + 9 * (a + 3 * b + 9 * c), x[q], x[q + N], x[q + 2 * N], y0, y1, y2);
            }
        }
    }
}

// Block-wide sum of two doubles; thread 0 writes the block partials.
__device__ __forceinline__ void block_sum2(double a, double b, double *__restrict__ part, long slot) {
    __shared__ double sa[kThreads / 64], sb[kThreads / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off, 64);
        b += __shfl_down(b, off, 64);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        sa[w] = a;
        sb[w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ta = 0, tb = 0;
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) {
            ta += sa[q];
            tb += sb[q];
        }
        part[2 * slot] = ta;
        part[2 * slot + 1] = tb;
    }
}

// out = r - A x (APPLY: out = A x).  Grid: (ceil(m*n/256), o).  With NORM,
// per-block partials of sum |r|^4 and sum |r|^2 go to part.
template <typename T, int MODE, bool NORM, bool APPLY>
__global__ __launch_bounds__(kThreads) void k_resid(Dims d, const T *__restrict__ x,
                                                    const T *__restrict__ r,
                                                    const T *__restrict__ ops,
                                                    T *__restrict__ out,
                                                    double *__restrict__ part) {
    const int k = blockIdx.y;
    const long mn = (long)d.m * d.n;
    const long pl = (long)blockIdx.x * kThreads + threadIdx.x;
    double q4 = 0, q2 = 0;
    if (pl < mn) {
        const int j = (int)(pl / d.m), i = (int)(pl - (long)j * d.m);
        const long p = pl + mn * k;
        const T *A = ops + cls_of<MODE>(p, k) * kBlk;
        T y0 = 0, y1 = 0, y2 = 0;
        stencil_sum<T, false>(d, x, A, i, j, k, y0, y1, y2);
        T o0, o1, o2;
        if (APPLY) {
            o0 = y0; o1 = y1; o2 = y2;
        } else {
            o0 = r[p] - y0;
            o1 = r[p + d.N] - y1;
            o2 = r[p + 2 * d.N] - y2;
        }
        out[p] = o0;
        out[p + d.N] = o1;
        out[p + 2 * d.N] = o2;
        if (NORM) {
            q2 = (double)o0 * o0 + (double)o1 * o1 + (double)o2 * o2;
            q4 = q2 * q2;
        }
    }
    if (NORM) block_sum2(q4, q2, part, (long)blockIdx.y * gridDim.x + blockIdx.x);
}

// Colour of coordinate i along a dimension of n vertices: i mod 2, with a
// third colour for the last vertex of an odd periodic dimension.
__host__ __device__ __forceinline__ int col_count(int n, int c, int prd) {
    const bool special = prd && (n & 1);
    if (c == 2) return special ? 1 : 0;
    int cnt = (n - c + 1) / 2;
    if (special && c == 0) cnt -= 1;
    return cnt;
}

__device__ __forceinline__ int col_coord(int t, int c, int n) { return c == 2 ? n - 1 : c + 2 * t; }

// Block Gauss-Seidel update of one point:
//   x_p = inv(A_pp) (r_p - sum_{b != centre} A_b x_{p+b})
template <typename T, int MODE>
__device__ __forceinline__ void gs_point(const Dims &d, T *__restrict__ x, const T *__restrict__ r,
                                         const T *__restrict__ ops, const T *__restrict__ inv, int i, int j,
                                         int k) {
    const long p = i + (long)d.m * (j + (long)d.n * k);
    const long c = cls_of<MODE>(p, k);
    const T *A = ops + c * kBlk;
    T y0 = 0, y1 = 0, y2 = 0;
    stencil_sum<T, true>(d, x, A, i, j, k, y0, y1, y2);
    const T e0 = r[p] - y0, e1 = r[p + d.N] - y1, e2 = r[p + 2 * d.N] - y2;
    const T *B = inv + c * 9;
    x[p] = B[0] * e0 + B[1] * e1 + B[2] * e2;
    x[p + d.N] = B[3] * e0 + B[4] * e1 + B[5] * e2;
    x[p + 2 * d.N] = B[6] * e0 + B[7] * e1 + B[8] * e2;
}

// One colour pass of block Gauss-Seidel over the points of colour
// (cx,cy,cz).  Points of one colour are never stencil neighbours, so the pass
// is order-independent.  Grid: (ceil(cntx*cnty/256), cntz).
template <typename T, int MODE>
__global__ __launch_bounds__(kThreads) void k_gs(Dims d, T *__restrict__ x,
                                                 const T *__restrict__ r,
                                                 const T *__restrict__ ops,
                                                 const T *__restrict__ inv, int cx, int cy,
                                                 int cz, int cntx, int cnty) {
    const int k = col_coord(blockIdx.y, cz, d.o);
    const long t = (long)blockIdx.x * kThreads + threadIdx.x;
    if (t >= (long)cntx * cnty) return;
    const int jj = (int)(t / cntx), ii = (int)(t - (long)jj * cntx);
    gs_point<T, MODE>(d, x, r, ops, inv, col_coord(ii, cx, d.m), col_coord(jj, cy, d.n), k);
}

// 1-D restriction weight of fine vertex 2I+e for coarse vertex I (vertex
// coarsening, periodic-even or non-periodic-odd sizes): 1 on the coarse
// vertex, 1/2 on the two midpoints; -1 marks an index outside the domain.
__device__ __forceinline__ int fine_of(int I, int e, int nf, int prd) {
    int f = 2 * I + e;
    if (f < 0) return prd ? f + nf : -1;
    if (f >= nf) return prd ? f - nf : -1;
    return f;
}

// r_c = (1/8) P^T res_f at coarse vertex (I,J,K) and x_c = 0
// (region::restriction, region.cc:1767-1812 and multigrid.cc:444).
template <typename T>
__device__ __forceinline__ void restrict_point(const Dims &F, const Dims &C, const T *__restrict__ res,
                                               T *__restrict__ rc, T *__restrict__ xc, int I, int J, int K) {
    T s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
    for (int ek = -1; ek <= 1; ek++) {
        const int fk = fine_of(K, ek, F.o, F.pz);
        if (fk < 0) continue;
        const T wk = ek == 0 ? T(1) : T(0.5);
#pragma unroll
        for (int ej = -1; ej <= 1; ej++) {
            const int fj = fine_of(J, ej, F.n, F.py);
            if (fj < 0) continue;
            const T wj = ej == 0 ? wk : T(0.5) * wk;
#pragma unroll
            for (int ei = -1; ei <= 1; ei++) {
                const int fi = fine_of(I, ei, F.m, F.px);
                if (fi < 0) continue;
                const T w = T(0.125) * (ei == 0 ? wj : T(0.5) * wj);
                const long q = fi + (long)F.m * (fj + (long)F.n * fk);
                s0 += w * res[q];
                s1 += w * res[q + F.N];
                s2 += w * res[q + 2 * F.N];
            }
        }
    }
    const long p = I + (long)C.m * (J + (long)C.n * K);
    rc[p] = s0;
    rc[p + C.N] = s1;
    rc[p + 2 * C.N] = s2;
    xc[p] = 0;
    xc[p + C.N] = 0;
    xc[p + 2 * C.N] = 0;
}

// The same restriction, one thread per coarse column (I, J) marching over
// kchunk coarse planes.  Fine plane 2K+1 is the +z plane of coarse plane K
// and the -z plane of K+1 with the same weights, so its nine-point partial
// sum -- formed from zero in restrict_point's order -- is carried in
// registers instead of being fetched again.  Each coarse sum is the carried
// partial plus the 0 and +z planes' terms in restrict_point's order, so the
// results are identical.
// Grid: (ceil(mc*nc/256), ceil(oc/kchunk)).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_restrict_march(Dims F, Dims C, const T *__restrict__ res,
                                                             T *__restrict__ rc, T *__restrict__ xc, int kchunk) {
    const long pl = (long)blockIdx.x * kThreads + threadIdx.x;
    if (pl >= (long)C.m * C.n) return;
    const int J = (int)(pl / C.m), I = (int)(pl - (long)J * C.m);
    const int K0 = blockIdx.y * kchunk;
    const int K1 = K0 + kchunk < C.o ? K0 + kchunk : C.o;
    int fj[3], fi[3];
#pragma unroll
    for (int e = 0; e < 3; e++) {
        fj[e] = fine_of(J, e - 1, F.n, F.py);
        fi[e] = fine_of(I, e - 1, F.m, F.px);
    }
    // terms of fine plane fk with z weight wk added to (s0, s1, s2), and
    // with PART also to the partial (q0, q1, q2) that starts from zero
    auto plane = [&](int fk, T wk, T &s0, T &s1, T &s2, bool part, T &q0, T &q1, T &q2) {
#pragma unroll
        for (int b = 0; b < 3; b++) {
            if (fj[b] < 0) continue;
            const T wj = b == 1 ? wk : T(0.5) * wk;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                if (fi[a] < 0) continue;
                const T w = T(0.125) * (a == 1 ? wj : T(0.5) * wj);
                const long q = fi[a] + (long)F.m * (fj[b] + (long)F.n * fk);
                const T v0 = res[q], v1 = res[q + F.N], v2 = res[q + 2 * F.N];
                s0 += w * v0;
                s1 += w * v1;
                s2 += w * v2;
                if (part) {
                    q0 += w * v0;
                    q1 += w * v1;
                    q2 += w * v2;
                }
            }
        }
    };
    T c0 = 0, c1 = 0, c2 = 0, d0 = 0, d1 = 0, d2 = 0;  // carried partial of the -z plane
    {
        const int fk = fine_of(K0, -1, F.o, F.pz);
        if (fk >= 0) plane(fk, T(0.5), c0, c1, c2, false, d0, d1, d2);
    }
    for (int K = K0; K < K1; K++) {
        T s0 = c0, s1 = c1, s2 = c2;
        plane(2 * K, T(1), s0, s1, s2, false, d0, d1, d2);
        const int fk = fine_of(K, 1, F.o, F.pz);
        T q0 = 0, q1 = 0, q2 = 0;
        if (fk >= 0) plane(fk, T(0.5), s0, s1, s2, true, q0, q1, q2);
        c0 = q0;
        c1 = q1;
        c2 = q2;
        const long p = I + (long)C.m * (J + (long)C.n * K);
        rc[p] = s0;
        rc[p + C.N] = s1;
        rc[p + 2 * C.N] = s2;
        xc[p] = 0;
        xc[p + C.N] = 0;
        xc[p + 2 * C.N] = 0;
    }
}

// Coarse parents of fine vertex f along one dimension: f even sits on coarse
// f/2; f odd is the midpoint of f/2 and f/2+1 (wrapped when periodic).
template <typename T>
__device__ __forceinline__ int parents1(int f, int nc, int prd, int *P, T *w) {
    P[0] = f >> 1;
    if (!(f & 1)) {
        w[0] = T(1);
        return 1;
    }
    int h = (f >> 1) + 1;
    if (prd && h >= nc) h -= nc;
    P[1] = h;
    w[0] = w[1] = T(0.5);
    return 2;
}

// x_f += P x_c at fine vertex (i,j,k), trilinear vertex interpolation
// (region::interpolation, region.cc:1817-1843, :2062-2121).
template <typename T>
__device__ __forceinline__ void interp_point(const Dims &F, const Dims &C, const T *__restrict__ xc,
                                             T *__restrict__ xf, int i, int j, int k) {
    int Ki[2], Ji[2], Ii[2];
    T wk[2], wj[2], wi[2];
    const int nk = parents1<T>(k, C.o, C.pz, Ki, wk);
    const int nj = parents1<T>(j, C.n, C.py, Ji, wj);
    const int ni = parents1<T>(i, C.m, C.px, Ii, wi);
    T s0 = 0, s1 = 0, s2 = 0;
    for (int c = 0; c < nk; c++)
        for (int b = 0; b < nj; b++)
            for (int a = 0; a < ni; a++) {
                const T w = wi[a] * wj[b] * wk[c];
                const long q = Ii[a] + (long)C.m * (Ji[b] + (long)C.n * Ki[c]);
                s0 += w * xc[q];
                s1 += w * xc[q + C.N];
                s2 += w * xc[q + 2 * C.N];
            }
    const long p = i + (long)F.m * (j + (long)F.n * k);
    xf[p] += s0;
    xf[p + F.N] += s1;
    xf[p + 2 * F.N] += s2;
}

// The same x_f += P x_c, one thread per coarse vertex (I, J, K) and the up
// to eight fine vertices 2(I, J, K) + (0|1, 0|1, 0|1) whose parents lie in
// {I, I+1} x {J, J+1} x {K, K+1}: the eight coarse values are loaded once
// instead of once per fine vertex.  Each fine sum runs over the same parents
// in the same order with the same weights as interp_point, so the results
// are identical.  Requires f = 2c (periodic) or f = 2c - 1 (walled) per
// dimension, the only sizes the hierarchy builds.
// Grid: (ceil(mc*nc/256), oc).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_interp_cells(Dims F, Dims C, const T *__restrict__ xc,
                                                           T *__restrict__ xf) {
    const long pl = (long)blockIdx.x * kThreads + threadIdx.x;
    if (pl >= (long)C.m * C.n) return;
    const int J = (int)(pl / C.m), I = (int)(pl - (long)J * C.m), K = blockIdx.y;
    // second parent along each dimension (used only when the odd fine
    // vertex exists: then I + 1 < C.m unless periodic, where it wraps)
    const int I1 = I + 1 < C.m ? I + 1 : 0, J1 = J + 1 < C.n ? J + 1 : 0, K1 = K + 1 < C.o ? K + 1 : 0;
    const bool ox = 2 * I + 1 < F.m, oy = 2 * J + 1 < F.n, oz = 2 * K + 1 < F.o;
    const int Ii[2] = {I, I1}, Ji[2] = {J, J1}, Ki[2] = {K, K1};
    T cv[2][2][2][3];
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int a = 0; a < 2; a++) {
                const bool need = (c == 0 || oz) && (b == 0 || oy) && (a == 0 || ox);
                const long q = Ii[a] + (long)C.m * (Ji[b] + (long)C.n * Ki[c]);
#pragma unroll
                for (int e = 0; e < 3; e++) cv[c][b][a][e] = need ? xc[q + e * C.N] : T(0);
            }
    // x-periodic fine rows have even length, so the pair 2I, 2I+1 is one
    // aligned two-value access per component
    struct alignas(2 * sizeof(T)) Pair {
        T v[2];
    };
    const bool pair = ox && !(F.m & 1);
#pragma unroll
    for (int fk = 0; fk < 2; fk++)
#pragma unroll
        for (int fj = 0; fj < 2; fj++) {
            if ((fk && !oz) || (fj && !oy)) continue;
            T sv[2][3];
#pragma unroll
            for (int fi = 0; fi < 2; fi++) {
                // parents of this fine vertex: 1 (even) or 2 (odd) per dimension
                const int nk = fk + 1, nj = fj + 1, ni = fi + 1;
                const T wk = fk ? T(0.5) : T(1), wj = fj ? T(0.5) : T(1), wi = fi ? T(0.5) : T(1);
                T s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
                for (int c = 0; c < 2; c++)
#pragma unroll
                    for (int b = 0; b < 2; b++)
#pragma unroll
                        for (int a = 0; a < 2; a++) {
                            if (c >= nk || b >= nj || a >= ni) continue;
                            const T w = wi * wj * wk;
                            s0 += w * cv[c][b][a][0];
                            s1 += w * cv[c][b][a][1];
                            s2 += w * cv[c][b][a][2];
                        }
                sv[fi][0] = s0;
                sv[fi][1] = s1;
                sv[fi][2] = s2;
            }
            const long p = 2 * I + (long)F.m * ((2 * J + fj) + (long)F.n * (2 * K + fk));
            if (pair) {
#pragma unroll
                for (int e = 0; e < 3; e++) {
                    Pair *q = reinterpret_cast<Pair *>(xf + p + e * F.N);
                    Pair u = *q;
                    u.v[0] += sv[0][e];
                    u.v[1] += sv[1][e];
                    *q = u;
                }
            } else {
#pragma unroll
                for (int e = 0; e < 3; e++) {
                    xf[p + e * F.N] += sv[0][e];
                    if (ox) xf[p + 1 + e * F.N] += sv[1][e];
                }
            }
        }
}

// Galerkin coarse operator A_c = (1/8) P^T A_f P (region::compute_rats,
// region.cc:1912-1966 with A_red/A_add, :2486-2578).  One thread per
// (coarse representative, coarse block D).  For coarse vertex I and fine
// candidate f = 2I+e (weight 1 for e=0, 1/2 otherwise) with fine block d,
// the column g = f+d feeds coarse column I+D with weight 1 if g = 2(I+D),
// 1/2 if |g - 2(I+D)| = 1.  PLANE levels use representatives (0,0,K).
__device__ __forceinline__ double pw(int s) { return s == 0 ? 1.0 : ((s == 1 || s == -1) ? 0.5 : 0.0); }

// One entry of one coarse block per thread (thread t: class t / 243, block
// (t % 243) / 9, entry t % 9).  Each entry is summed over the fine parents
// and fine blocks in one fixed order, so the result does not depend on how
// the entries are spread over threads.
template <int MODE_F, int MODE_C>
__global__ void k_rat(Dims F, Dims C, const double *__restrict__ opsF, double *__restrict__ opsC,
                      long nrep) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrep * kBlk) return;
    const long rep = t / kBlk;
    const int DE = (int)(t - rep * kBlk);
    const int D = DE / 9, e = DE - 9 * D;
    const int Di = D % 3 - 1, Dj = (D / 3) % 3 - 1, Dk = D / 9 - 1;
    int I, J, K;
    if (MODE_C == 0) {
        I = 0; J = 0; K = (int)rep;
    } else {
        K = (int)(rep / ((long)C.m * C.n));
        const long rem = rep - (long)K * C.m * C.n;
        J = (int)(rem / C.m);
        I = (int)(rem - (long)J * C.m);
    }
    // Coarse column out of a non-periodic domain: the block stays zero.
    const bool okc = (C.px || (I + Di >= 0 && I + Di < C.m)) &&
                     (C.py || (J + Dj >= 0 && J + Dj < C.n)) &&
                     (C.pz || (K + Dk >= 0 && K + Dk < C.o));
    double acc = 0;
    if (okc) {
        for (int ek = -1; ek <= 1; ek++) {
            const int fk = fine_of(K, ek, F.o, F.pz);
            if (fk < 0) continue;
            for (int ej = -1; ej <= 1; ej++) {
                const int fj = fine_of(J, ej, F.n, F.py);
                if (fj < 0) continue;
                for (int ei = -1; ei <= 1; ei++) {
                    const int fi = fine_of(I, ei, F.m, F.px);
                    if (fi < 0) continue;
                    const double wr = 0.125 * pw(ei) * pw(ej) * pw(ek);
                    const long pf = fi + (long)F.m * (fj + (long)F.n * fk);
                    const double *A = opsF + cls_of<MODE_F>(pf, fk) * kBlk + e;
                    for (int dk = -1; dk <= 1; dk++) {
                        if (!F.pz && (fk + dk < 0 || fk + dk >= F.o)) continue;
                        const double w3 = pw(ek + dk - 2 * Dk);
                        if (w3 == 0) continue;
                        for (int dj = -1; dj <= 1; dj++) {
                            if (!F.py && (fj + dj < 0 || fj + dj >= F.n)) continue;
                            const double w2 = w3 * pw(ej + dj - 2 * Dj);
                            if (w2 == 0) continue;
                            for (int di = -1; di <= 1; di++) {
                                if (!F.px && (fi + di < 0 || fi + di >= F.m)) continue;
                                const double w = wr * w2 * pw(ei + di - 2 * Di);
                                if (w == 0) continue;
                                acc += w * A[9 * ((di + 1) + 3 * (dj + 1) + 9 * (dk + 1))];
                            }
                        }
                    }
                }
            }
        }
    }
    opsC[rep * kBlk + DE] = acc;
}

// Rounding-level asymmetry of a Galerkin level, removed: when every mirrored
// entry pair (block o, block -o) of every plane differs by at most rtol x
// the level's largest entry, both entries become their mean, so the level
// is centrosymmetric bit for bit and its sweeps add each pair of opposite
// neighbours before one multiply.  The product of a centrosymmetric fine
// operator with the symmetric transfers is centrosymmetric; k_rat's fixed
// summation order leaves mirrored entries apart by rounding only (relative
// to the largest entry: 2e-16 to 1.1e-15 on the transformed operator's
// Galerkin levels, 64^3 periodic), while the walls of a walled level make
// them differ by percents, and such a level is left as it is.  One
// workgroup; ncls tables of kBlk values.
constexpr int kSymThreads = 1024;
constexpr double kCentroTol = 1e-13;  // rtol of k_centro_sym as the library uses it
__global__ __launch_bounds__(kSymThreads) void k_centro_sym(double *__restrict__ ops, long ncls, double rtol) {
    __shared__ double red[kSymThreads / 64];
    const long ne = ncls * kBlk, np = ncls * 117;  // 13 blocks x 9 entries per plane
    double m = 0;
    for (long q = threadIdx.x; q < ne; q += kSymThreads) m = fmax(m, fabs(ops[q]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    double scale = 0;
#pragma unroll
    for (int w = 0; w < kSymThreads / 64; w++) scale = fmax(scale, red[w]);
    const double lim = rtol * scale;
    int bad = 0;
    for (long q = threadIdx.x; q < np; q += kSymThreads) {
        const long c = q / 117, r = q - c * 117, b = r / 9, e = r - b * 9;
        const double a = ops[c * kBlk + 9 * b + e], z = ops[c * kBlk + 9 * (26 - b) + e];
        if (!(fabs(a - z) <= lim)) bad = 1;  // a NaN also leaves the level as it is
    }
    if (__syncthreads_or(bad)) return;
    for (long q = threadIdx.x; q < np; q += kSymThreads) {
        const long c = q / 117, r = q - c * 117, b = r / 9, e = r - b * 9;
        double &a = ops[c * kBlk + 9 * b + e], &z = ops[c * kBlk + 9 * (26 - b) + e];
        const double v = 0.5 * (a + z);
        a = v;
        z = v;
    }
}

// Zero pattern and centrosymmetry of a per-plane level, so that the host
// reads 28 words instead of the tables: out[b] (b < 27) = OR over the
// planes of the nonzero entries of block b (bit e for entry e, a NaN
// counting as nonzero), out[27] = 1 when block(o) equals block(-o) bit for
// bit in every plane.  One workgroup of kThreads.
constexpr int kCfgWords = 28;
__global__ __launch_bounds__(kThreads) void k_level_config(const double *__restrict__ ops, long ncls,
                                                           unsigned *__restrict__ out) {
    __shared__ unsigned used[27];
    if (threadIdx.x < 27) used[threadIdx.x] = 0;
    __syncthreads();
    const long ne = ncls * kBlk, np = ncls * 117;  // 13 blocks x 9 entries per plane
    for (long q = threadIdx.x; q < ne; q += kThreads)
        if (ops[q] != 0.0) {
            const int r = (int)(q % kBlk);
            atomicOr(&used[r / 9], 1u << (r % 9));
        }
    int bad = 0;
    for (long q = threadIdx.x; q < np; q += kThreads) {
        const long c = q / 117, r = q - c * 117, b = r / 9, e = r - b * 9;
        if (ops[c * kBlk + 9 * b + e] != ops[c * kBlk + 9 * (26 - b) + e]) bad = 1;
    }
    bad = __syncthreads_or(bad);  // (a barrier: the LDS masks are complete)
    if (threadIdx.x < 27) out[threadIdx.x] = used[threadIdx.x];
    if (threadIdx.x == 0) out[27] = bad ? 0u : 1u;
}

// Centre-block inverse by cofactors over the determinant (src/mat3.hh:177-183)
// and FP32 copies of the table.  flag[0] is set when a determinant is zero
// or not finite.
__global__ void k_inverse(const double *__restrict__ ops, long ncls, double *__restrict__ inv64,
                          float *__restrict__ inv32, int *__restrict__ flag) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncls) return;
    const double *a = ops + c * kBlk + 9 * 13;
    const double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[2] * a[7] - a[1] * a[8],
                 c02 = a[1] * a[5] - a[2] * a[4];
    const double c10 = a[5] * a[6] - a[3] * a[8], c11 = a[0] * a[8] - a[2] * a[6],
                 c12 = a[2] * a[3] - a[0] * a[5];
    const double c20 = a[3] * a[7] - a[4] * a[6], c21 = a[1] * a[6] - a[0] * a[7],
                 c22 = a[0] * a[4] - a[1] * a[3];
    const double det = a[0] * c00 + a[3] * c01 + a[6] * c02;
    if (!(det != 0.0) || !isfinite(det)) flag[0] = 1;
    const double id = 1.0 / det;
    const double v[9] = {c00 * id, c01 * id, c02 * id, c10 * id, c11 * id,
                         c12 * id, c20 * id, c21 * id, c22 * id};
#pragma unroll
    for (int e = 0; e < 9; e++) {
        inv64[c * 9 + e] = v[e];
        inv32[c * 9 + e] = (float)v[e];
    }
}

__global__ void k_cast_d2f(const double *__restrict__ a, float *__restrict__ b, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = (float)a[t];
}

// Deterministic final reduction of per-block partial pairs.
__global__ void k_reduce_partials(const double *__restrict__ part, long nb, double *__restrict__ out) {
    __shared__ double sa[1024], sb[1024];
    double a = 0, b = 0;
    for (long q = threadIdx.x; q < nb; q += blockDim.x) {
        a += part[2 * q];
        b += part[2 * q + 1];
    }
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int s = blockDim.x >> 1; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sa[threadIdx.x] += sa[threadIdx.x + s];
            sb[threadIdx.x] += sb[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sa[0];
        out[1] = sb[0];
    }
}

// AoS (3 per point) <-> SoA (component-major).
__global__ void k_aos_to_soa(const double *__restrict__ a, double *__restrict__ s, long N) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    s[p] = a[3 * p];
    s[p + N] = a[3 * p + 1];
    s[p + 2 * N] = a[3 * p + 2];
}
__global__ void k_soa_to_aos(const double *__restrict__ s, double *__restrict__ a, long N) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    a[3 * p] = s[p];
    a[3 * p + 1] = s[p + N];
    a[3 * p + 2] = s[p + 2 * N];
}

// Element-wise kernels take kEW elements per thread, kThreads apart (each
// load instruction of a wave covers one contiguous segment, and a thread
// has kEW loads in flight): launch nblk_ew(n) workgroups of kThreads.
constexpr int kEW = 4;
inline unsigned nblk_ew(long n) { return (unsigned)((n + (long)kThreads * kEW - 1) / ((long)kThreads * kEW)); }

// FP32 defect d = res * inv_s (the scale from the host) and, when e is
// given, e = 0 (a V-cycle whose first sweep starts from zero without reading
// x needs no zeros in memory).
template <typename TI = double>
__global__ __launch_bounds__(kThreads) void k_defect_f32(const TI *__restrict__ res, double inv_s,
                                                         float *__restrict__ d, float *__restrict__ e, long n) {
    const long b0 = (long)blockIdx.x * (kThreads * kEW) + threadIdx.x;
    TI v[kEW];
#pragma unroll
    for (int u = 0; u < kEW; u++) {
        const long i = b0 + (long)u * kThreads;
        v[u] = i < n ? res[i] : TI(0);
    }
#pragma unroll
    for (int u = 0; u < kEW; u++) {
        const long i = b0 + (long)u * kThreads;
        if (i < n) {
            d[i] = (float)((double)v[u] * inv_s);
            if (e) e[i] = 0.0f;
        }
    }
}

// x64 += s * e32.
__global__ void k_update_f64(double *__restrict__ x, const float *__restrict__ e, double s, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) x[t] += s * (double)e[t];
}

// b = a converted (kEW elements per thread, grid nblk_ew(n)).
template <typename A, typename B>
__global__ __launch_bounds__(kThreads) void k_convert(const A *__restrict__ a, B *__restrict__ b, long n) {
    const long b0 = (long)blockIdx.x * (kThreads * kEW) + threadIdx.x;
    A v[kEW];
#pragma unroll
    for (int u = 0; u < kEW; u++) {
        const long i = b0 + (long)u * kThreads;
        v[u] = i < n ? a[i] : A(0);
    }
#pragma unroll
    for (int u = 0; u < kEW; u++) {
        const long i = b0 + (long)u * kThreads;
        if (i < n) b[i] = (B)v[u];
    }
}

template <typename T>
__global__ void k_fill(T *__restrict__ a, T v, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = v;
}

}  // namespace stzmg
