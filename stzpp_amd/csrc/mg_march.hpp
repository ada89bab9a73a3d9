// Z-marching fused 8-colour block Gauss-Seidel for per-plane operators on
// levels periodic in x and y, with x a multiple of the tile width (8, 16, 32
// or 64).
//
// Same colour order and the same halo recomputation as k_gs_planes
// (mg_fused.hpp): one launch updates every plane of one z parity P, a
// workgroup owns a TX x TY tile and recomputes a shrinking halo (3, 2, 1, 0)
// over the four in-plane colours, reading x_cur / x_new and writing x_new.
// What differs is how the work reaches the hardware:
//
// * The planes k-1 and k+1 (the other parity) do not change during the
//   launch, so their part of the stencil sum is taken once per point before
//   the in-plane sub-passes ("phase Z"); the sub-passes then only add the
//   eight in-plane blocks.
// * Coefficients live in scalar registers: each phase loads its nonzero
//   coefficients in groups of at most MarchGroup<T>::regs registers through the
//   scalar cache, one wait per group, and the multiply-adds take them as
//   scalar operands (no per-point broadcast instruction).
// * Staging reads one window row per wave instruction: row addresses are
//   scalar, the per-lane column offsets (with the periodic wrap) are computed
//   once per workgroup, so a staged value costs one load and one LDS store.
// * The global loads of the next plane's right-hand side and of the next two
//   staged planes are issued before the arithmetic of the current plane and
//   stored to LDS after it.  A thread owns a horizontal pair of points (one
//   of sub-pass 0 and one of 1, likewise 2 and 3): the pair's right-hand
//   side is one two-value load per component, and its new values go from
//   registers to x_new as one two-value store after the second sub-pass.
// * A sweep that starts from x = 0 stages the planes it knows are zero
//   without reading them (zflags).
#pragma once

#include <type_traits>
#include <utility>

#include "mg_fused.hpp"

namespace stzmg {

template <int I>
using IC = std::integral_constant<int, I>;

template <typename F, int... Is>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, Is...>) {
    (f(IC<Is>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Pattern keys of the marching kernel: a zero pattern PAT_* (0..3), plus
// kMarchSym when every plane's stencil is centrosymmetric (block(-o) equal
// to block(o) bit for bit, as the shear and transformed stencils and their
// per-plane wall rows are).  Such a stencil sums each opposite pair of
// neighbours first, x(p+o) + x(p-o), and multiplies once: half the
// multiply-adds and half the coefficients.
constexpr int kMarchSym = 4;
__host__ __device__ constexpr bool march_sym(int key) { return key >= kMarchSym; }

// Blocks of the two phases: Z = the 18 blocks with dk != 0 (centrosymmetric:
// the 9 with dk = -1), then the eight in-plane blocks (centrosymmetric: the 4
// before the centre) and the inverse centre block (index 27, 9 entries).
__host__ __device__ constexpr int march_block(int key, int phase, int i) {
    if (march_sym(key)) return phase == 0 ? i : (i < 4 ? 9 + i : 27);
    return phase == 0 ? (i < 9 ? i : 9 + i) : (i < 4 ? 9 + i : (i < 8 ? 10 + i : 27));
}
__host__ __device__ constexpr int march_len(int key, int phase) {
    return march_sym(key) ? (phase == 0 ? 9 : 5) : (phase == 0 ? 18 : 9);
}
__host__ __device__ constexpr unsigned march_mask(int key, int b) {
    return b == 27 ? 0x1FFu : pat_mask(key % kMarchSym, b);
}

// Greedy grouping of a phase's blocks: a new group starts where adding the
// block would exceed G coefficients.
__host__ __device__ constexpr int march_group(int pat, int phase, int G, int i) {
    int g = 0, run = 0;
    for (int q = 0; q <= i; q++) {
        const int nz = popc9(march_mask(pat, march_block(pat, phase, q)));
        if (run > 0 && run + nz > G) {
            g++;
            run = 0;
        }
        run += nz;
    }
    return g;
}
__host__ __device__ constexpr int march_ngroups(int pat, int phase, int G) {
    return march_group(pat, phase, G, march_len(pat, phase) - 1) + 1;
}
// Coefficients in group g, and position of block i's first coefficient in it.
__host__ __device__ constexpr int march_gsize(int pat, int phase, int G, int g) {
    int n = 0;
    for (int q = 0; q < march_len(pat, phase); q++)
        if (march_group(pat, phase, G, q) == g) n += popc9(march_mask(pat, march_block(pat, phase, q)));
    return n;
}
__host__ __device__ constexpr int march_pos(int pat, int phase, int G, int i) {
    const int g = march_group(pat, phase, G, i);
    int n = 0;
    for (int q = 0; q < i; q++)
        if (march_group(pat, phase, G, q) == g) n += popc9(march_mask(pat, march_block(pat, phase, q)));
    return n;
}
// rank of entry e among the set bits of m
__host__ __device__ constexpr int rank9(unsigned m, int e) {
    int n = 0;
    for (int q = 0; q < e; q++) n += (m >> q) & 1;
    return n;
}

// Packed coefficient table of a plane: the nonzero entries of pattern PAT in
// phase order (phase Z blocks, then the in-plane blocks, then the inverse
// centre block), so every group is a contiguous run (wide scalar loads).
__host__ __device__ constexpr int march_count(int pat, int phase) {
    int n = 0;
    for (int q = 0; q < march_len(pat, phase); q++) n += popc9(march_mask(pat, march_block(pat, phase, q)));
    return n;
}
__host__ __device__ constexpr int march_stride(int pat) { return march_count(pat, 0) + march_count(pat, 1); }
// Source of packed entry j: 9 b + e for operator entries (b < 27), 243 + e
// for the inverse block.
__host__ __device__ constexpr int march_src(int pat, int j) {
    for (int ph = 0; ph < 2; ph++)
        for (int q = 0; q < march_len(pat, ph); q++) {
            const int b = march_block(pat, ph, q);
            const unsigned m = march_mask(pat, b);
            const int n = popc9(m);
            if (j < n) return b == 27 ? 243 + nth_bit(m, j) : 9 * b + nth_bit(m, j);
            j -= n;
        }
    return 0;
}
// Offset of group g of a phase in the packed table.
__host__ __device__ constexpr int march_goff(int pat, int phase, int G, int g) {
    int n = phase == 0 ? 0 : march_count(pat, 0);
    for (int q = 0; q < g; q++) n += march_gsize(pat, phase, G, q);
    return n;
}

// pk[c * stride + j] = entry march_src(j) of class c (operator or inverse).
template <typename T>
__global__ void k_pack_march(const T *__restrict__ ops, const T *__restrict__ inv, const int *__restrict__ src,
                             int stride, long ncls, T *__restrict__ pk) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncls * stride) return;
    const long c = t / stride;
    const int j = (int)(t - c * stride), q = src[j];
    pk[t] = q < kBlk ? ops[c * kBlk + q] : inv[c * 9 + (q - kBlk)];
}

// Element at a 32-bit byte offset from a base pointer: with the base
// uniform, the access takes the scalar-base plus vector-offset form (no
// 64-bit address arithmetic per access).
template <typename T>
__device__ __forceinline__ const T &at_bytes(const T *base, unsigned off) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
}
template <typename T>
__device__ __forceinline__ T &at_bytes(T *base, unsigned off) {
    return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + off);
}

// Scalar registers per coefficient group (a double takes two).
template <typename T>
struct MarchGroup {
    static constexpr int regs = 48;
    static constexpr int G = regs / (int)(sizeof(T) / 4);
};

// Loads group g of a phase into scalar registers: the addresses are uniform
// and the operator read-only, so the loads go through the scalar cache, and
// every value passes through an empty asm that takes it in a scalar
// register, so the group is complete and held there from this point on
// (one wait for the group).
template <typename T, int PAT, int PHASE, int GRP, int N>
__device__ __forceinline__ void march_load(const T *__restrict__ pk, T (&c)[N]) {
    constexpr int G = MarchGroup<T>::G;
    constexpr int off = march_goff(PAT, PHASE, G, GRP);
    static_for<N>([&](auto q) { c[q] = pk[off + decltype(q)::value]; });
#pragma unroll
    for (int q = 0; q < N; q++) asm volatile("" : "+s"(c[q]));
}

// Adds block (phase, i) of group GRP times the neighbour values of one point.
// ld(c, di, dj, dk) reads component c of the neighbour.
template <typename T, int PAT, int PHASE, int I, int N, typename LD>
__device__ __forceinline__ void march_block_acc(const T (&c)[N], const LD &ld, T &a0, T &a1, T &a2) {
    constexpr int G = MarchGroup<T>::G;
    constexpr int b = march_block(PAT, PHASE, I);
    constexpr unsigned m = march_mask(PAT, b);
    if constexpr (b != 27 && m != 0) {
        constexpr int di = b % 3 - 1, dj = (b / 3) % 3 - 1, dk = b / 9 - 1;
        constexpr unsigned cols = pat_cols(m);
        constexpr int q0 = march_pos(PAT, PHASE, G, I);
        T v0 = 0, v1 = 0, v2 = 0;
        if constexpr (march_sym(PAT)) {
            // the opposite neighbour has the same block
            if constexpr (cols & 1) v0 = ld(0, di, dj, dk) + ld(0, -di, -dj, -dk);
            if constexpr (cols & 2) v1 = ld(1, di, dj, dk) + ld(1, -di, -dj, -dk);
            if constexpr (cols & 4) v2 = ld(2, di, dj, dk) + ld(2, -di, -dj, -dk);
        } else {
            if constexpr (cols & 1) v0 = ld(0, di, dj, dk);
            if constexpr (cols & 2) v1 = ld(1, di, dj, dk);
            if constexpr (cols & 4) v2 = ld(2, di, dj, dk);
        }
        if constexpr (m & 0x001) a0 += c[q0 + rank9(m, 0)] * v0;
        if constexpr (m & 0x002) a0 += c[q0 + rank9(m, 1)] * v1;
        if constexpr (m & 0x004) a0 += c[q0 + rank9(m, 2)] * v2;
        if constexpr (m & 0x008) a1 += c[q0 + rank9(m, 3)] * v0;
        if constexpr (m & 0x010) a1 += c[q0 + rank9(m, 4)] * v1;
        if constexpr (m & 0x020) a1 += c[q0 + rank9(m, 5)] * v2;
        if constexpr (m & 0x040) a2 += c[q0 + rank9(m, 6)] * v0;
        if constexpr (m & 0x080) a2 += c[q0 + rank9(m, 7)] * v1;
        if constexpr (m & 0x100) a2 += c[q0 + rank9(m, 8)] * v2;
    }
}

// Threads that own a point of sub-pass s under the lane mapping of
// k_gs_march (a prefix of the workgroup).
template <int TX, int TY, int NT>
constexpr int march_lanes(int s) {
    using SP = SubPass<TX, TY, NT>;
    if (TX != 64) return SP::npts(s);
    const int ny = SP::ny(s), extra = (SP::nx(s) - 32) * ny;
    return 32 * (ny + (extra + 31) / 32);
}

// Waves per SIMD that the LDS of one workgroup (three staged window planes)
// allows on a CU of 160 KiB: the register allocation is held to that
// occupancy.
// The 16-wide tiles serve small levels (16^3 to 32^3 points), whose launches
// are latency-bound: they take the registers of two waves per SIMD rather
// than spill to reach the occupancy their LDS would allow.
template <typename T, int TX, int TY, int NT>
constexpr int march_waves() {
    constexpr int lds = 3 * 3 * (TX + 7) * (TY + 8) * (int)sizeof(T);
    constexpr int wgs = (160 * 1024) / lds;
    constexpr int w = wgs * NT / 64 / 4;
    constexpr int cap = TX == 16 ? 2 : 8;
    return w < 1 ? 1 : (w > cap ? cap : w);
}

// Grid: (m/TX, n/TY, ceil(planes of parity P / kchunk)).  Requires a PLANE
// operator, x and y periodic, m % TX == 0, n % TY == 0, TX in {8, 16, 32, 64}
// and every sub-pass of the tile at most NT points.  pk: the packed table of
// the level (k_pack_march), one row of march_stride(PAT) values per plane.
// zflags: bit 0 = the planes of xs are all zero, bit 1 = those of xo are
// (a sweep that starts from x = 0): such planes are staged as zeros without
// reading them, so the buffer's contents do not matter.
template <typename T, int PAT, int TX, int TY, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(march_waves<T, TX, TY, NT>()))) void k_gs_march(Dims d, const T *__restrict__ xs, const T *__restrict__ xo,
                                                 T *__restrict__ xd, const T *__restrict__ r,
                                                 const T *__restrict__ pk, int P, int kchunk, int zflags) {
    // The window holds what the four sub-passes read: their points reach
    // at most 3 (x) beyond the tile and their stencils one more, i.e. x in
    // [-3, TX + 3] (origin HXL = 3 points left of the tile), and y in
    // [-4, TY + 4) (origin H = 4; an even number of rows keeps the row
    // staging regular).  At 64 x 8 in FP32 the three slots then take
    // 40,896 bytes: four workgroups per CU.
    constexpr int H = 4, HXL = 3;
    constexpr int WX = TX + 2 * HXL + 1, WY = TY + 2 * H;
    constexpr int HX = (WX + 1) / 2;       // x is stored split by parity
    constexpr int CSZ = WX * WY;           // one component of one plane
    constexpr int PSZ = 3 * CSZ;           // one plane slot
    constexpr int NW = NT / 64;            // waves
    static_assert(TX == 8 || TX == 16 || TX == 32 || TX == 64, "tile width");
    static_assert(TY % 2 == 0 && NT % 64 == 0, "tile shape");
    // chunk A: the tile's columns of every window row in 16-byte loads;
    // chunk B: the WX - TX halo columns, one element per thread and load
    constexpr int VW = 16 / (int)sizeof(T);  // values per 16-byte load
    constexpr int LPR = TX / VW;             // lanes per window row
    constexpr int RPI = 64 / LPR;            // window rows per wave instruction
    static_assert(TX % VW == 0 && 64 % LPR == 0, "tile width");
    constexpr int CB = WX - TX;
    constexpr int NA = (3 * WY + NW * RPI - 1) / (NW * RPI);
    constexpr int EB = 3 * WY * CB, NB = (EB + NT - 1) / NT;
    struct alignas(16) Vec {
        T v[VW];
    };
    using SP = SubPass<TX, TY, NT>;
    static_assert(march_lanes<TX, TY, NT>(0) <= NT && march_lanes<TX, TY, NT>(1) <= NT &&
                      march_lanes<TX, TY, NT>(2) <= NT && march_lanes<TX, TY, NT>(3) <= NT,
                  "one point per thread and sub-pass");
    constexpr int G = MarchGroup<T>::G;
    __shared__ T lds[3 * PSZ];

    int tx, ty, tz;
    xcd_tile(tx, ty, tz);
    const int x0 = tx * TX, y0 = ty * TY;
    const int nP = (d.o - P + 1) / 2;
    const int kk0 = tz * kchunk;
    const int kk1 = kk0 + kchunk < nP ? kk0 + kchunk : nP;
    if (kk0 >= kk1) return;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const long mn = (long)d.m * d.n;

    // ---- per-thread geometry, once per workgroup ----
    // chunk A: the tile's own TX columns of each window row, VW consecutive
    // values per lane (one 16-byte load: the row segment starts at the tile's
    // aligned x0), LPR lanes per row and RPI rows per wave instruction.
    // Instruction i of wave wv covers rows (wv + NW i) RPI + lane / LPR
    // (component-major, 3 WY rows); per lane the global and LDS offsets of
    // each instruction are computed once.
    unsigned ga[NA];
    int la[NA];
    const int col4 = lane % LPR;
#pragma unroll
    for (int i = 0; i < NA; i++) {
        const int row0 = (wv + NW * i) * RPI + lane / LPR;
        const int row = row0 < 3 * WY ? row0 : 3 * WY - 1;
        const int c = row / WY, yy = row - c * WY;
        ga[i] = (unsigned)(sizeof(T) * (c * d.N + (long)d.m * wrap1(y0 - H + yy, d.n) + x0 + col4 * VW));
        la[i] = c * CSZ + yy * WX + ((HXL + col4 * VW) >> 1);
    }
    // chunk B: the halo columns [0, HXL) and [HXL + TX, WX), up to NB
    // elements per thread
    unsigned vb_off[NB];
    int lb_off[NB];
#pragma unroll
    for (int j = 0; j < NB; j++) {
        const int e = tid + NT * j < EB ? tid + NT * j : EB - 1;
        const int row = e / CB, cb = e % CB, col = cb < HXL ? cb : TX + cb;
        const int c = row / WY, yy = row - c * WY;
        vb_off[j] = (unsigned)(sizeof(T) *
                               (c * d.N + (long)d.m * wrap1(y0 - H + yy, d.n) + wrap1(x0 - HXL + col, d.m)));
        lb_off[j] = c * CSZ + yy * WX + (col & 1) * HX + (col >> 1);
    }
    // sub-pass points, dealt in pairs: a thread that owns point (x, y) of
    // sub-pass 2g (x even) also owns (x + 1, y) of sub-pass 2g + 1 where that
    // point exists (sub-passes 1 and 3 start one x after 0 and 2, with no
    // more columns and no more rows beyond the shared ones).  The pair's
    // right-hand sides are one aligned two-value load per component and its
    // new values one two-value store, once the second is computed: full cache
    // lines, where points of one colour alone fill every other value.  A
    // pair lies inside the tile or outside it as a whole (x and the tile
    // edges are even), and a pair inside has its second point.  Per thread:
    // LDS index of each point (component 0, slot 0), whether it owns it, and
    // per pair the offset in a plane of r and x and whether it is inside.
    // Lanes map to pairs so that a half-wave reads one LDS row segment: with
    // TX = 64 (32..35 pairs per row) half-wave h < ny takes the first 32
    // pairs of row h and the columns beyond go to the half-waves after;
    // otherwise pairs are dealt in order.
    static_assert(SP::xf(1) == SP::xf(0) + 1 && SP::yf(1) == SP::yf(0) && SP::nx(1) <= SP::nx(0) &&
                      SP::ny(1) <= SP::ny(0),
                  "sub-pass 1 pairs with sub-pass 0");
    static_assert(SP::xf(3) == SP::xf(2) + 1 && SP::yf(3) == SP::yf(2) + 2 && SP::nx(3) <= SP::nx(2) &&
                      SP::ny(3) + 1 <= SP::ny(2),
                  "sub-pass 3 pairs with sub-pass 2");
    struct alignas(2 * sizeof(T)) Pair {
        T v[2];
    };
    int pb[4];
    unsigned ro[2], wo[2];
    bool own[4], inside[2];
    static_for<2>([&](auto g_) {
        constexpr int g = decltype(g_)::value;
        constexpr int s = 2 * g;                                // first sub-pass of the pair
        constexpr int dy = (SP::yf(s + 1) - SP::yf(s)) / 2;   // row offset of the second
        constexpr int nx = SP::nx(s), ny = SP::ny(s);
        int txx, tyy;
        bool ok;
        if constexpr (TX == 64) {
            constexpr int ex = nx - 32;
            const int hw = tid >> 5, l = tid & 31;
            if (hw < ny) {
                ok = true;
                txx = l;
                tyy = hw;
            } else {
                const int e = l + 32 * (hw - ny);
                ok = e < ex * ny;
                const int ee = ok ? e : 0;
                txx = 32 + (ex > 0 ? ee % (ex > 0 ? ex : 1) : 0);
                tyy = ex > 0 ? ee / (ex > 0 ? ex : 1) : 0;
                if (!ok) txx = 0;
            }
        } else {
            ok = tid < SP::npts(s);
            const int t = ok ? tid : 0;
            tyy = t / nx;
            txx = t - tyy * nx;
        }
        const int xx = SP::xf(s) + 2 * txx + HXL, yy = SP::yf(s) + 2 * tyy + H;
        pb[s] = yy * WX + (xx & 1) * HX + (xx >> 1);
        pb[s + 1] = yy * WX + ((xx + 1) & 1) * HX + ((xx + 1) >> 1);
        own[s] = ok;
        own[s + 1] = ok && txx < SP::nx(s + 1) && tyy >= dy && tyy - dy < SP::ny(s + 1);
        ro[g] = (unsigned)(sizeof(T) * (wrap1(x0 + xx - HXL, d.m) + (long)d.m * wrap1(y0 + yy - H, d.n)));
        inside[g] = ok && xx >= HXL && xx < HXL + TX && yy >= H && yy < H + TY;
        wo[g] = (unsigned)(sizeof(T) * ((x0 + xx - HXL) + (long)d.m * (y0 + yy - H)));
    });

    // ---- staging ----
    struct Stage {
        Vec a[NA];
        T b[NB];
    };
    const bool zs = (zflags & 1) != 0, zo = (zflags & 2) != 0;
    auto fetch = [&](const T *__restrict__ src, int kg, Stage &v, bool zero) {
        bool valid = !zero;
        int kw = kg;
        if (kg < 0 || kg >= d.o) {
            if (d.pz) kw = kg < 0 ? kg + d.o : kg - d.o;
            else valid = false;
        }
        if (valid) {
            const T *base = src + mn * kw;
#pragma unroll
            for (int i = 0; i < NA; i++) v.a[i] = at_bytes(reinterpret_cast<const Vec *>(base), ga[i]);
#pragma unroll
            for (int j = 0; j < NB; j++) v.b[j] = at_bytes(base, vb_off[j]);
        } else {
#pragma unroll
            for (int i = 0; i < NA; i++)
#pragma unroll
                for (int e = 0; e < VW; e++) v.a[i].v[e] = 0;
#pragma unroll
            for (int j = 0; j < NB; j++) v.b[j] = 0;
        }
    };
    // element e of a chunk-A load sits at window x = HXL + col4 VW + e,
    // parity-split: half (HXL + e) & 1, index la + ((HXL & 1) + e) / 2
    auto store = [&](int slot, const Stage &v) {
        T *sl = lds + slot * PSZ;
#pragma unroll
        for (int i = 0; i < NA; i++)
            if ((3 * WY) % (NW * RPI) == 0 || (wv + NW * i) * RPI + lane / LPR < 3 * WY) {
#pragma unroll
                for (int e = 0; e < VW; e++)
                    sl[la[i] + ((HXL + e) & 1) * HX + ((HXL & 1) + e) / 2] = v.a[i].v[e];
            }
#pragma unroll
        for (int j = 0; j < NB; j++)
            if (EB % NT == 0 || tid + NT * j < EB) sl[lb_off[j]] = v.b[j];
    };

    // right-hand side of the four points of plane kr (two pairs)
    auto load_r = [&](int kr, T (&v)[4][3]) {
        const T *rk = r + mn * kr;
        static_for<2>([&](auto g) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const Pair q = at_bytes(reinterpret_cast<const Pair *>(rk + c * d.N), ro[g]);
                v[2 * g][c] = q.v[0];
                v[2 * g + 1][c] = q.v[1];
            }
        });
    };

    // The first three planes and the first right-hand side are all in
    // flight before any of them is stored (one memory round trip, not
    // three in a row).
    int sm = 0, sc = 1, sp = 2;  // slots of planes k-1, k, k+1
    T rr[4][3];
    {
        Stage v0, v1, v2;
        const int k = P + 2 * kk0;
        fetch(xo, k - 1, v0, zo);
        fetch(xs, k, v1, zs);
        fetch(xo, k + 1, v2, zo);
        load_r(k, rr);
        store(sm, v0);
        store(sc, v1);
        store(sp, v2);
    }
    __syncthreads();
    // complete here: the wait the compiler places before the first use of rr
    // inside the loop would otherwise count the prefetches issued there too
#pragma unroll
    for (int s = 0; s < 4; s++) asm volatile("" : "+v"(rr[s][0]), "+v"(rr[s][1]), "+v"(rr[s][2]));

    // One plane.  The next plane's right-hand side and the next two staged
    // planes (k + 2, k + 3) are issued first and stay in flight during this
    // plane's arithmetic; the planes go to LDS at the end.  The loads
    // complete in order, so each wait covers only what is due.  (Issuing the
    // planes two steps ahead into a second register set measured the same.)
    auto step = [&](int kk, Stage &n1, Stage &n2) {
        const int k = P + 2 * kk;
        const bool more = kk + 1 < kk1;
        T rn[4][3];
        if (more) {
            load_r(k + 2, rn);
            fetch(xs, k + 2, n1, zs);
            fetch(xo, k + 3, n2, zo);
        }
        const T *A = pk + (long)k * march_stride(PAT);

        // phase Z: planes k-1 and k+1.  The two points of a pair are
        // addressed from the first one's window index (the second sits a
        // compile-time distance away), so the cells their neighbourhoods
        // share are read from LDS once (not in the 8-wide class, whose one
        // wave would then hold more values than its registers).
        constexpr bool SHARE = TX != 8;
        T acc[4][3];
        static_for<4>([&](auto s) { acc[s][0] = acc[s][1] = acc[s][2] = 0; });
        {
            int bm[4], bp[4];
            static_for<4>([&](auto s) {
                bm[s] = pb[s] + sm * PSZ;
                bp[s] = pb[s] + sp * PSZ;
            });
            static_for<march_ngroups(PAT, 0, G)>([&](auto g_) {
                constexpr int GR = decltype(g_)::value;
                constexpr int NG = march_gsize(PAT, 0, G, GR);
                if constexpr (NG > 0) {
                    T c[NG];
                    march_load<T, PAT, 0, GR>(A, c);
                    static_for<2>([&](auto pg) {
                        constexpr int PG = decltype(pg)::value;
                        // waves without a point of this pair of sub-passes
                        // skip it (a wave-uniform branch)
                        if (wv * 64 >= march_lanes<TX, TY, NT>(2 * PG)) return;
                        static_for<2>([&](auto h_) {
                            constexpr int S = 2 * PG + decltype(h_)::value;
                            constexpr int xp = (SP::xf(S) + HXL) & 1;
                            // pb[S] - pb[2 PG]: x + 1 of a parity-split row
                            constexpr int dl = !SHARE || S == 2 * PG ? 0 : (xp == 0 ? 1 - HX : HX);
                            constexpr int SB = SHARE ? 2 * PG : S;
                            auto ld = [&](int cc, int di, int dj, int dk) -> T {
                                const int b0 = dk < 0 ? bm[SB] : bp[SB];
                                const int off = dj * WX + (((xp + di) & 1) - xp) * HX + ((xp + di) >> 1);
                                return lds[b0 + dl + cc * CSZ + off];
                            };
                            static_for<march_len(PAT, 0)>([&](auto i) {
                                if constexpr (march_group(PAT, 0, G, decltype(i)::value) == GR)
                                    march_block_acc<T, PAT, 0, decltype(i)::value>(c, ld, acc[S][0], acc[S][1],
                                                                                    acc[S][2]);
                            });
                        });
                    });
                }
            });
        }

        // the phase-Z sums are complete here: the arithmetic is not sunk
        // into the sub-passes, where its coefficients would stay live
#pragma unroll
        for (int s = 0; s < 4; s++) asm volatile("" : "+v"(acc[s][0]), "+v"(acc[s][1]), "+v"(acc[s][2]));

        // in-plane sub-passes on the centre slot.  A new value goes to LDS
        // (the later sub-passes read it); for a pair of the tile itself the
        // first value waits in registers and the second sub-pass stores both
        // to x_new.
        T *xk = xd + mn * k;
        T hold[3] = {0, 0, 0};
        auto put = [&](auto s_, T v0, T v1, T v2) {
            constexpr int S = decltype(s_)::value;
            const int bc = pb[S] + sc * PSZ;
            lds[bc] = v0;
            lds[bc + CSZ] = v1;
            lds[bc + 2 * CSZ] = v2;
            if constexpr (S % 2 == 0) {
                hold[0] = v0;
                hold[1] = v1;
                hold[2] = v2;
            } else if (inside[S / 2]) {
                at_bytes(reinterpret_cast<Pair *>(xk), wo[S / 2]) = Pair{{hold[0], v0}};
                at_bytes(reinterpret_cast<Pair *>(xk + d.N), wo[S / 2]) = Pair{{hold[1], v1}};
                at_bytes(reinterpret_cast<Pair *>(xk + 2 * d.N), wo[S / 2]) = Pair{{hold[2], v2}};
            }
        };
        constexpr int NG1 = march_ngroups(PAT, 1, G);
        auto inplane = [&](auto s, auto &&coef) {
            constexpr int S = decltype(s)::value;
            const int bc = pb[S] + sc * PSZ;
            constexpr int xp = (SP::xf(S) + HXL) & 1;
            auto ld = [&](int cc, int di, int dj, int) -> T {
                const int off = dj * WX + (((xp + di) & 1) - xp) * HX + ((xp + di) >> 1);
                return lds[bc + cc * CSZ + off];
            };
            coef(ld);
            (void)bc;
        };
        if constexpr (NG1 == 1) {
            // one group: loaded once for the four sub-passes
            constexpr int NC = march_gsize(PAT, 1, G, 0);
            T c[NC];
            march_load<T, PAT, 1, 0>(A, c);
            static_for<4>([&](auto s) {
                constexpr int S = decltype(s)::value;
                if (own[S]) {
                    inplane(s, [&](auto &ld) {
                        static_for<march_len(PAT, 1) - 1>([&](auto i) {
                            march_block_acc<T, PAT, 1, decltype(i)::value>(c, ld, acc[S][0], acc[S][1], acc[S][2]);
                        });
                    });
                    const T e0 = rr[S][0] - acc[S][0], e1 = rr[S][1] - acc[S][1], e2 = rr[S][2] - acc[S][2];
                    constexpr int qi = march_pos(PAT, 1, G, march_len(PAT, 1) - 1);
                    put(s, c[qi + 0] * e0 + c[qi + 1] * e1 + c[qi + 2] * e2,
                        c[qi + 3] * e0 + c[qi + 4] * e1 + c[qi + 5] * e2,
                        c[qi + 6] * e0 + c[qi + 7] * e1 + c[qi + 8] * e2);
                }
                __syncthreads();
            });
        } else {
            static_for<4>([&](auto s) {
                constexpr int S = decltype(s)::value;
                if (own[S]) {
                    T e0 = 0, e1 = 0, e2 = 0;
                    static_for<NG1>([&](auto g_) {
                        constexpr int GR = decltype(g_)::value;
                        constexpr int NC = march_gsize(PAT, 1, G, GR);
                        T c[NC];
                        march_load<T, PAT, 1, GR>(A, c);
                        inplane(s, [&](auto &ld) {
                            static_for<march_len(PAT, 1) - 1>([&](auto i) {
                                if constexpr (march_group(PAT, 1, G, decltype(i)::value) == GR)
                                    march_block_acc<T, PAT, 1, decltype(i)::value>(c, ld, acc[S][0], acc[S][1],
                                                                                    acc[S][2]);
                            });
                        });
                        if constexpr (GR == NG1 - 1) {
                            // the group that holds the inverse block is the last
                            constexpr int qi = march_pos(PAT, 1, G, march_len(PAT, 1) - 1);
                            const T f0 = rr[S][0] - acc[S][0], f1 = rr[S][1] - acc[S][1],
                                    f2 = rr[S][2] - acc[S][2];
                            e0 = c[qi + 0] * f0 + c[qi + 1] * f1 + c[qi + 2] * f2;
                            e1 = c[qi + 3] * f0 + c[qi + 4] * f1 + c[qi + 5] * f2;
                            e2 = c[qi + 6] * f0 + c[qi + 7] * f1 + c[qi + 8] * f2;
                        }
                    });
                    put(s, e0, e1, e2);
                }
                __syncthreads();
            });
        }

        if (more) {
            // planes k-1 and k are free after the last sub-pass's barrier
            const int nsm = sp, nsc = sm, nsp = sc;
            store(nsc, n1);
            store(nsp, n2);
#pragma unroll
            for (int s = 0; s < 4; s++)
#pragma unroll
                for (int c = 0; c < 3; c++) rr[s][c] = rn[s][c];
            sm = nsm;
            sc = nsc;
            sp = nsp;
            __syncthreads();
        }
    };

    Stage n1, n2;
    for (int kk = kk0; kk < kk1; kk++) step(kk, n1, n2);
}

}  // namespace stzmg
