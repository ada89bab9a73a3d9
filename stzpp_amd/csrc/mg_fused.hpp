// Fused 8-colour block Gauss-Seidel for per-plane operators (x, y periodic).
//
// The colours are (i mod 2, j mod 2, k mod 2), visited with k mod 2
// outermost.  One launch updates every plane of one z parity P: planes of
// parity P only couple to each other through planes of the other parity,
// which this launch only reads.  Inside a plane the four in-plane colours
// depend on each other; a workgroup owns a TX x TY tile of the plane and
// recomputes a shrinking halo (3, 2, 1, 0 points) so that its tile comes out
// exactly as after four separate colour passes, with the plane and its two
// neighbours staged in LDS.  A sweep is two launches (P = 0, 1) instead of
// eight, and every value of x is read about twice from HBM instead of eight
// times.
//
// Tiles recompute the halo from old values, so neighbouring tiles must not
// see each other's new values: the launch reads x and writes a second
// buffer.  P = 0 reads both parities from x_cur; P = 1 reads its own parity
// from x_cur and the other (already updated) parity from x_new; both write
// x_new.  After the two launches x_new holds the sweep.
//
// The operator's zero pattern is a template parameter, so the stencil sum
// only contains the entries an operator family can have (the shear operator
// has 45 nonzero entries of 243, the transformed operator of a T with only
// t13 off the diagonal 79, Galerkin coarse operators all 27 blocks).
#pragma once

#include <utility>

#include "mg_kernels.hpp"

namespace stzmg {

enum { PAT_DENSE = 0, PAT_NOCORNER = 1, PAT_TRANS_XZ = 2, PAT_SHEAR = 3 };

// 9-bit mask of the entries (row-major 3x3) that block b may hold.
__host__ __device__ constexpr unsigned pat_mask(int pat, int b) {
    const int di = b % 3 - 1, dj = (b / 3) % 3 - 1, dk = b / 9 - 1;
    const int nz = (di != 0) + (dj != 0) + (dk != 0);
    if (pat == PAT_DENSE) return 0x1FF;
    if (nz == 3) return 0;  // corner blocks
    if (pat == PAT_NOCORNER) return 0x1FF;
    if (pat == PAT_TRANS_XZ) {
        if (nz <= 1) return 0x155;                 // centre, faces: entries 0 2 4 6 8
        if (dj == 0) return 0x155;                 // x-z edges
        if (di == 0) return 0x0A2;                 // y-z edges: entries 1 5 7
        return 0x08A;                              // x-y edges: entries 1 3 7
    }
    // PAT_SHEAR
    if (nz <= 1) return 0x111;                     // diagonal
    if (dk == 0) return 0x00A;                     // x-y edges: entries 1 3
    if (dj == 0) return 0x044;                     // x-z edges: entries 2 6
    return 0x0A0;                                  // y-z edges: entries 5 7
}

// Columns of a block that the mask touches.
__host__ __device__ constexpr unsigned pat_cols(unsigned m) {
    return ((m & 0x049) ? 1u : 0u) | ((m & 0x092) ? 2u : 0u) | ((m & 0x124) ? 4u : 0u);
}

// One plane's 27 blocks and inverse centre block (252 values; the inverse
// is left zero when inv is null) spread over the 64 lanes of a wave: entry
// q sits in lane q % 64 of word q / 64.  load() runs with all lanes active.  The
// stencil loop broadcasts a coefficient with v_readlane, a VALU operation.
// Loaded through the scalar cache instead, every group of coefficients would
// wait on the counter that the LDS reads share, draining the reads in flight.
template <typename T>
struct LaneCoef;
template <>
struct LaneCoef<float> {
    int w[4];
    __device__ __forceinline__ void load(const float *__restrict__ A, const float *__restrict__ inv) {
        const int lane = threadIdx.x & 63;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int q = lane + 64 * j;
            const float v = q < kBlk ? A[q] : (inv && q < kBlk + 9 ? inv[q - kBlk] : 0.0f);
            w[j] = __float_as_int(v);
            // pins the load to this point, where every lane is active: a
            // load sunk into a divergent branch would leave the lanes that
            // v_readlane reads unwritten
            asm volatile("" : "+v"(w[j]));
        }
    }
    template <int Q>
    __device__ __forceinline__ float get() const {
        return __int_as_float(__builtin_amdgcn_readlane(w[Q >> 6], Q & 63));
    }
};
template <>
struct LaneCoef<double> {
    int lo[4], hi[4];
    __device__ __forceinline__ void load(const double *__restrict__ A, const double *__restrict__ inv) {
        const int lane = threadIdx.x & 63;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int q = lane + 64 * j;
            const double v = q < kBlk ? A[q] : (inv && q < kBlk + 9 ? inv[q - kBlk] : 0.0);
            lo[j] = __double2loint(v);
            hi[j] = __double2hiint(v);
            asm volatile("" : "+v"(lo[j]), "+v"(hi[j]));
        }
    }
    template <int Q>
    __device__ __forceinline__ double get() const {
        return __hiloint2double(__builtin_amdgcn_readlane(hi[Q >> 6], Q & 63),
                                __builtin_amdgcn_readlane(lo[Q >> 6], Q & 63));
    }
};

__host__ __device__ constexpr int popc9(unsigned m) {
    int n = 0;
    for (int e = 0; e < 9; e++) n += (m >> e) & 1;
    return n;
}
__host__ __device__ constexpr int nth_bit(unsigned m, int n) {
    for (int e = 0; e < 9; e++)
        if ((m >> e) & 1) {
            if (n == 0) return e;
            n--;
        }
    return 0;
}

// The broadcast coefficients of a block in groups of three: the empty asm
// takes them as inputs and hands them back, so the three v_readlane of a
// group issue before the first multiply-add that uses one of them.  A
// v_readlane result read by the next VALU instruction needs two wait states
// (an s_nop); three readlanes ahead of their uses cover them without one.
// No instruction is emitted and the arithmetic is unchanged.
template <unsigned M, int G, typename T>
__device__ __forceinline__ void pin_group(T (&c)[9]) {
    constexpr int n = popc9(M), i0 = 3 * G;
    if constexpr (i0 + 2 < n)
        asm volatile("" : "+s"(c[nth_bit(M, i0)]), "+s"(c[nth_bit(M, i0 + 1)]), "+s"(c[nth_bit(M, i0 + 2)]));
    else if constexpr (i0 + 1 < n)
        asm volatile("" : "+s"(c[nth_bit(M, i0)]), "+s"(c[nth_bit(M, i0 + 1)]));
    else if constexpr (i0 < n)
        asm volatile("" : "+s"(c[nth_bit(M, i0)]));
}

#ifndef STZ_FUSED_PIN
#define STZ_FUSED_PIN 1
#endif

template <typename T, int PAT, int B, typename CF, typename LD>
__device__ __forceinline__ void acc_block(const CF &cf, const LD &ld, T &a0, T &a1, T &a2) {
    constexpr unsigned m = pat_mask(PAT, B);
    if constexpr (B != 13 && m != 0) {
        constexpr int di = B % 3 - 1, dj = (B / 3) % 3 - 1, dk = B / 9 - 1;
        constexpr unsigned cols = pat_cols(m);
        constexpr int q = 9 * B;
        T v0 = 0, v1 = 0, v2 = 0;
        if constexpr (cols & 1) v0 = ld(0, di, dj, dk);
        if constexpr (cols & 2) v1 = ld(1, di, dj, dk);
        if constexpr (cols & 4) v2 = ld(2, di, dj, dk);
        T c[9];
        if constexpr (m & 0x001) c[0] = cf.template get<q + 0>();
        if constexpr (m & 0x002) c[1] = cf.template get<q + 1>();
        if constexpr (m & 0x004) c[2] = cf.template get<q + 2>();
        if constexpr (m & 0x008) c[3] = cf.template get<q + 3>();
        if constexpr (m & 0x010) c[4] = cf.template get<q + 4>();
        if constexpr (m & 0x020) c[5] = cf.template get<q + 5>();
        if constexpr (m & 0x040) c[6] = cf.template get<q + 6>();
        if constexpr (m & 0x080) c[7] = cf.template get<q + 7>();
        if constexpr (m & 0x100) c[8] = cf.template get<q + 8>();
        if constexpr (STZ_FUSED_PIN) {
            pin_group<m, 0>(c);
            pin_group<m, 1>(c);
            pin_group<m, 2>(c);
        }
        if constexpr (m & 0x001) a0 += c[0] * v0;
        if constexpr (m & 0x002) a0 += c[1] * v1;
        if constexpr (m & 0x004) a0 += c[2] * v2;
        if constexpr (m & 0x008) a1 += c[3] * v0;
        if constexpr (m & 0x010) a1 += c[4] * v1;
        if constexpr (m & 0x020) a1 += c[5] * v2;
        if constexpr (m & 0x040) a2 += c[6] * v0;
        if constexpr (m & 0x080) a2 += c[7] * v1;
        if constexpr (m & 0x100) a2 += c[8] * v2;
    }
}

template <typename T, int PAT, typename CF, typename LD, int... Bs>
__device__ __forceinline__ void acc_blocks(const CF &cf, const LD &ld, T &a0, T &a1, T &a2,
                                           std::integer_sequence<int, Bs...>) {
    (acc_block<T, PAT, Bs>(cf, ld, a0, a1, a2), ...);
}

// Build-time tuning of the fused smoother: whether the next two planes and
// the right-hand side are prefetched into registers ahead of the arithmetic,
// and the minimum number of waves per SIMD the register allocation must
// allow.  Measured on MI355X at 256^3: prefetching raises the FP32 kernel
// to ~290 registers (one wave per SIMD) and makes a sweep 2.8x slower
// (1.30 ms against 0.47 ms); FP64 gains 3%.  Both default to off.
#ifndef STZ_FUSED_PREFETCH_X
#define STZ_FUSED_PREFETCH_X 0
#endif
#ifndef STZ_FUSED_PREFETCH_R
#define STZ_FUSED_PREFETCH_R 0
#endif
#ifndef STZ_FUSED_WAVES
#define STZ_FUSED_WAVES 1
#endif

__device__ __forceinline__ int wrap1(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Tile of this workgroup on a (gridDim.x, gridDim.y, gridDim.z) tiling.
// Workgroups are dealt to the 8 XCDs round-robin (workgroup b runs on XCD
// b % 8) and every XCD has its own L2.  Tile number (b % 8) * (nb / 8) + b / 8
// gives each XCD a contiguous run of tiles (x fastest, then y, then z), so
// the halo lines neighbouring tiles share are fetched into one L2.
constexpr int kXcds = 8;
__device__ __forceinline__ void xcd_tile(int &tx, int &ty, int &tz) {
    const int nb = gridDim.x * gridDim.y * gridDim.z;
    const int b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const int t = nb % kXcds ? b : (b % kXcds) * (nb / kXcds) + b / kXcds;
    tx = t % gridDim.x;
    ty = (t / gridDim.x) % gridDim.y;
    tz = t / (gridDim.x * gridDim.y);
}

// Geometry of in-plane sub-pass s (colour (s & 1, s >> 1), halo 3 - s) on a
// TX x TY tile: first relative x/y of the colour and the number of points.
template <int TX, int TY, int NT>
struct SubPass {
    static constexpr int h(int s) { return 3 - s; }
    static constexpr int xf(int s) { return -h(s) + (((s & 1) + h(s)) & 1); }
    static constexpr int yf(int s) { return -h(s) + (((s >> 1) + h(s)) & 1); }
    static constexpr int nx(int s) { return (TX + h(s) - 1 - xf(s)) / 2 + 1; }
    static constexpr int ny(int s) { return (TY + h(s) - 1 - yf(s)) / 2 + 1; }
    static constexpr int npts(int s) { return nx(s) * ny(s); }
    static constexpr int iters(int s) { return (npts(s) + NT - 1) / NT; }
    static constexpr int max_iters = iters(0);
};

// Grid: (m/TX, n/TY, ceil(planes of parity P / kchunk)).  Requires a PLANE
// operator, x and y periodic with m % TX == 0 and n % TY == 0.
//
// Per plane: four sub-passes on the three staged planes, write-back of the
// tile, then the next two planes are staged into the slots that fell free.
template <typename T, int PAT, int TX, int TY, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(STZ_FUSED_WAVES))) void k_gs_planes(
    Dims d, const T *__restrict__ xs, const T *__restrict__ xo, T *__restrict__ xd, const T *__restrict__ r,
    const T *__restrict__ ops, const T *__restrict__ inv, int P, int kchunk) {
    constexpr int H = 4;                 // 3 recomputed halo points + the stencil reach
    constexpr int WX = TX + 2 * H, WY = TY + 2 * H;
    constexpr int HX = WX / 2;           // x is stored split by parity
    constexpr int CSZ = WX * WY;         // one component of one plane
    constexpr int PSZ = 3 * CSZ;         // one plane slot
    constexpr int NPER = (PSZ + NT - 1) / NT;
    using SP = SubPass<TX, TY, NT>;
    constexpr int MI = SP::max_iters;
    __shared__ T lds[3 * PSZ];

    int tx, ty, tz;
    xcd_tile(tx, ty, tz);
    const int x0 = tx * TX, y0 = ty * TY;
    const int nP = (d.o - P + 1) / 2;
    const int kk0 = tz * kchunk;
    const int kk1 = kk0 + kchunk < nP ? kk0 + kchunk : nP;
    if (kk0 >= kk1) return;
    const int tid = threadIdx.x;

    auto li = [&](int c, int s, int yy, int xx) { return s * PSZ + c * CSZ + yy * WX + (xx & 1) * HX + (xx >> 1); };

    auto fetch = [&](const T *__restrict__ src, int kg, T (&v)[NPER]) {
        bool valid = true;
        int kw = kg;
        if (kg < 0 || kg >= d.o) {
            if (d.pz) kw = kg < 0 ? kg + d.o : kg - d.o;
            else valid = false;
        }
        const long base = (long)d.m * d.n * kw;
#pragma unroll
        for (int u = 0; u < NPER; u++) {
            const int q = tid + u * NT;
            T val = 0;
            if (valid && q < PSZ) {
                const int c = q / CSZ, rem = q - c * CSZ, yy = rem / WX, xx = rem - yy * WX;
                const int gi = wrap1(x0 + xx - H, d.m), gj = wrap1(y0 + yy - H, d.n);
                val = src[c * d.N + base + gi + (long)d.m * gj];
            }
            v[u] = val;
        }
    };
    auto store = [&](int s, const T (&v)[NPER]) {
#pragma unroll
        for (int u = 0; u < NPER; u++) {
            const int q = tid + u * NT;
            if (q < PSZ) {
                const int c = q / CSZ, rem = q - c * CSZ, yy = rem / WX, xx = rem - yy * WX;
                lds[li(c, s, yy, xx)] = v[u];
            }
        }
    };

    int sm = 0, sc = 1, sp = 2;  // slots of planes k-1, k, k+1
    T va[NPER], vb[NPER];
    {
        const int k = P + 2 * kk0;
        fetch(xo, k - 1, va);
        fetch(xs, k, vb);
        store(sm, va);
        store(sc, vb);
        fetch(xo, k + 1, va);
        store(sp, va);
    }
    __syncthreads();
    for (int kk = kk0; kk < kk1; kk++) {
        const int k = P + 2 * kk;
        const long base = (long)d.m * d.n * k;
        const bool more = kk + 1 < kk1;
#if STZ_FUSED_PREFETCH_X
        if (more) {
            fetch(xs, k + 2, va);
            fetch(xo, k + 3, vb);
        }
#endif
        auto rload = [&](int s, int t, T (&v)[3]) {
            const int ty = t / SP::nx(s), tx = t - ty * SP::nx(s);
            const int gi = wrap1(x0 + SP::xf(s) + 2 * tx, d.m), gj = wrap1(y0 + SP::yf(s) + 2 * ty, d.n);
            const long p = base + gi + (long)d.m * gj;
            v[0] = r[p];
            v[1] = r[p + d.N];
            v[2] = r[p + 2 * d.N];
        };
        // right-hand side of every point this thread computes in this plane
        T rr[4][MI][3];
#if STZ_FUSED_PREFETCH_R
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
            for (int u = 0; u < MI; u++) {
                const int t = tid + u * NT;
                if (u < SP::iters(s) && t < SP::npts(s)) rload(s, t, rr[s][u]);
            }
#endif
        LaneCoef<T> cf;
        cf.load(ops + (long)k * kBlk, inv + (long)k * 9);
#pragma unroll
        for (int s = 0; s < 4; s++) {
#pragma unroll
            for (int u = 0; u < MI; u++) {
                const int t = tid + u * NT;
                if (u < SP::iters(s) && t < SP::npts(s)) {
                    const int ty = t / SP::nx(s), tx = t - ty * SP::nx(s);
                    const int xx = SP::xf(s) + 2 * tx + H, yy = SP::yf(s) + 2 * ty + H;
                    auto ld = [&](int c, int di, int dj, int dk) -> T {
                        const int sl = dk < 0 ? sm : (dk > 0 ? sp : sc);
                        return lds[li(c, sl, yy + dj, xx + di)];
                    };
#if !STZ_FUSED_PREFETCH_R
                    rload(s, t, rr[s][u]);
#endif
                    T a0 = 0, a1 = 0, a2 = 0;
                    acc_blocks<T, PAT>(cf, ld, a0, a1, a2, std::make_integer_sequence<int, 27>{});
                    const T e0 = rr[s][u][0] - a0, e1 = rr[s][u][1] - a1, e2 = rr[s][u][2] - a2;
                    constexpr int I = kBlk;
                    lds[li(0, sc, yy, xx)] =
                        cf.template get<I + 0>() * e0 + cf.template get<I + 1>() * e1 + cf.template get<I + 2>() * e2;
                    lds[li(1, sc, yy, xx)] =
                        cf.template get<I + 3>() * e0 + cf.template get<I + 4>() * e1 + cf.template get<I + 5>() * e2;
                    lds[li(2, sc, yy, xx)] =
                        cf.template get<I + 6>() * e0 + cf.template get<I + 7>() * e1 + cf.template get<I + 8>() * e2;
                }
            }
            __syncthreads();
        }
        for (int q = tid; q < 3 * TX * TY; q += NT) {
            const int c = q / (TX * TY), rem = q - c * (TX * TY), yy = rem / TX, xx = rem - yy * TX;
            xd[c * d.N + base + (x0 + xx) + (long)d.m * (y0 + yy)] = lds[li(c, sc, yy + H, xx + H)];
        }
        if (more) {
            __syncthreads();
            const int nsm = sp, nsc = sm, nsp = sc;
#if !STZ_FUSED_PREFETCH_X
            fetch(xs, k + 2, va);
            store(nsc, va);
            fetch(xo, k + 3, va);
            store(nsp, va);
#else
            store(nsc, va);
            store(nsp, vb);
#endif
            sm = nsm;
            sc = nsc;
            sp = nsp;
            __syncthreads();
        }
    }
}

}  // namespace stzmg
