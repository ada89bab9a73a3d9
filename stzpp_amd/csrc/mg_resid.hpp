// Plane-marching residual out = r - A x (APPLY: out = A x) for per-plane
// operators on levels periodic in x and y, with the machinery of the
// marching smoother (mg_march.hpp):
//
// * A TX x TY tile walks a chunk of z planes with the planes k-1, k, k+1 of
//   x in LDS (a window of (TX + 2) x (TY + 2) points per component) and the
//   plane k+2 in flight in registers while plane k is computed.
// * Staging reads the tile's columns of each window row with 16-byte loads
//   and the two halo columns per element; every offset, with the periodic
//   wrap, is computed once per workgroup.
// * Coefficients come from a packed per-plane table (the smoother's, with
//   the centre block in place of its inverse) through the scalar cache in
//   groups of scalar registers, and centrosymmetric stencils add each pair
//   of opposite neighbours before one multiply.
// * A thread owns R vertically adjacent points of one column, so the window
//   rows they share are read from LDS once.
//
// Reference: the residual of mg_3d_template/region.cc:1009-1055.
#pragma once

#include "mg_march.hpp"

namespace stzmg {

// Grid: (m/TX, n/TY, ceil(o / kchunk)).  Requires a PLANE operator, x and y
// periodic, m % TX == 0, n % TY == 0.  pk: one row of march_stride(PAT)
// values per plane, the last nine the centre block.  With NORM every
// workgroup writes its partial sums of |r|^4 and |r|^2 to part (also a
// workgroup without planes, with zeros).  With APPLY and xcopy given, x is
// also copied to xcopy.
template <typename T, int PAT, bool NORM, bool APPLY, int TX, int TY, int NT>
__global__ __launch_bounds__(NT) void k_resid_march(Dims d, const T *__restrict__ x, const T *__restrict__ r,
                                                    const T *__restrict__ pk, T *__restrict__ out,
                                                    double *__restrict__ part, int kchunk, T *__restrict__ xcopy) {
    static_assert(NT == kThreads, "block_sum2 assumes kThreads threads");
    static_assert(NT % TX == 0 && (TX * TY) % NT == 0, "thread layout");
    constexpr int WX = TX + 2, WY = TY + 2, CSZ = WX * WY, PSZ = 3 * CSZ;
    constexpr int R = TX * TY / NT;  // rows per thread
    constexpr int NW = NT / 64;
    // chunk A: the tile's TX columns of each window row, VW values per lane
    // (one 16-byte load), LPR lanes per row, RPI rows per wave instruction;
    // chunk B: the two halo columns (x = -1 and x = TX), one value per load
    constexpr int VW = 16 / (int)sizeof(T);
    constexpr int LPR = TX / VW;
    constexpr int RPI = 64 / LPR;
    static_assert(TX % VW == 0 && 64 % LPR == 0, "tile width");
    constexpr int NA = (3 * WY + NW * RPI - 1) / (NW * RPI);
    constexpr int EB = 3 * WY * 2, NB = (EB + NT - 1) / NT;
    constexpr int G = MarchGroup<T>::G;
    struct alignas(16) Vec {
        T v[VW];
    };
    __shared__ T lds[3 * PSZ];

    int tx, ty, tz;
    xcd_tile(tx, ty, tz);
    const int x0 = tx * TX, y0 = ty * TY;
    const int k0 = tz * kchunk;
    const int k1 = k0 + kchunk < d.o ? k0 + kchunk : d.o;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const long mn = (long)d.m * d.n;

    // ---- staging geometry, once per workgroup ----
    unsigned ga[NA];
    int la[NA];
    const int col4 = lane % LPR;
#pragma unroll
    for (int i = 0; i < NA; i++) {
        const int row0 = (wv + NW * i) * RPI + lane / LPR;
        const int row = row0 < 3 * WY ? row0 : 3 * WY - 1;
        const int c = row / WY, yy = row - c * WY;
        ga[i] = (unsigned)(sizeof(T) * (c * d.N + (long)d.m * wrap1(y0 - 1 + yy, d.n) + x0 + col4 * VW));
        la[i] = c * CSZ + yy * WX + 1 + col4 * VW;
    }
    unsigned vb_off[NB];
    int lb_off[NB];
#pragma unroll
    for (int j = 0; j < NB; j++) {
        const int e = tid + NT * j < EB ? tid + NT * j : EB - 1;
        const int row = e >> 1, col = (e & 1) ? WX - 1 : 0;
        const int c = row / WY, yy = row - c * WY;
        vb_off[j] = (unsigned)(sizeof(T) *
                               (c * d.N + (long)d.m * wrap1(y0 - 1 + yy, d.n) + wrap1(x0 - 1 + col, d.m)));
        lb_off[j] = c * CSZ + yy * WX + col;
    }
    struct Stage {
        Vec a[NA];
        T b[NB];
    };
    auto fetch = [&](int kg, Stage &v) {
        bool valid = true;
        int kw = kg;
        if (kg < 0 || kg >= d.o) {
            if (d.pz) kw = kg < 0 ? kg + d.o : kg - d.o;
            else valid = false;
        }
        if (valid) {
            const T *base = x + mn * kw;
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
    auto store = [&](int slot, const Stage &v) {
        T *sl = lds + slot * PSZ;
#pragma unroll
        for (int i = 0; i < NA; i++)
            if ((3 * WY) % (NW * RPI) == 0 || (wv + NW * i) * RPI + lane / LPR < 3 * WY) {
#pragma unroll
                for (int e = 0; e < VW; e++) sl[la[i] + e] = v.a[i].v[e];
            }
#pragma unroll
        for (int j = 0; j < NB; j++)
            if (EB % NT == 0 || tid + NT * j < EB) sl[lb_off[j]] = v.b[j];
    };

    // the thread's points: column xx, rows yb .. yb + R - 1 of the tile
    const int xx = tid % TX, yb = (tid / TX) * R;
    // byte offsets in a plane: row u of the thread at po + u rstep
    const unsigned po = (unsigned)(sizeof(T) * ((x0 + xx) + (long)d.m * (y0 + yb)));
    const unsigned rstep = (unsigned)(sizeof(T) * d.m);

    double q4 = 0, q2 = 0;
    int sm = 0, sc = 1, sp = 2;
    if (k0 < k1) {
        // the three planes in flight before any is stored
        Stage v0, v1, v2;
        fetch(k0 - 1, v0);
        fetch(k0, v1);
        fetch(k0 + 1, v2);
        store(sm, v0);
        store(sc, v1);
        store(sp, v2);
    }
    __syncthreads();
    for (int k = k0; k < k1; k++) {
        const bool more = k + 1 < k1;
        Stage nx;
        if (more) fetch(k + 2, nx);
        const long base = mn * k;
        T rv[R][3];
        if constexpr (!APPLY) {
#pragma unroll
            for (int u = 0; u < R; u++)
#pragma unroll
                for (int c = 0; c < 3; c++) rv[u][c] = at_bytes(r + base + c * d.N, po + u * rstep);
        }
        const T *A = pk + (long)k * march_stride(PAT);
        T acc[R][3];
#pragma unroll
        for (int u = 0; u < R; u++) acc[u][0] = acc[u][1] = acc[u][2] = 0;

        // phase Z (planes k - 1, k + 1), then the in-plane blocks
        static_for<2>([&](auto ph_) {
            constexpr int PH = decltype(ph_)::value;
            static_for<march_ngroups(PAT, PH, G)>([&](auto g_) {
                constexpr int GR = decltype(g_)::value;
                constexpr int NC = march_gsize(PAT, PH, G, GR);
                if constexpr (NC > 0) {
                    T c[NC];
                    march_load<T, PAT, PH, GR>(A, c);
#pragma unroll
                    for (int u = 0; u < R; u++) {
                        const int b0 = (yb + u + 1) * WX + xx + 1;
                        auto ld = [&](int cc, int di, int dj, int dk) -> T {
                            const int sl = dk < 0 ? sm : (dk > 0 ? sp : sc);
                            return lds[sl * PSZ + cc * CSZ + b0 + dj * WX + di];
                        };
                        static_for<march_len(PAT, PH)>([&](auto i) {
                            if constexpr (march_group(PAT, PH, G, decltype(i)::value) == GR)
                                march_block_acc<T, PAT, PH, decltype(i)::value>(c, ld, acc[u][0], acc[u][1],
                                                                                 acc[u][2]);
                        });
                        if constexpr (PH == 1 && GR == march_ngroups(PAT, 1, G) - 1) {
                            // the centre block closes the last group
                            constexpr int qc = march_pos(PAT, 1, G, march_len(PAT, 1) - 1);
                            const T v0 = lds[sc * PSZ + b0], v1 = lds[sc * PSZ + CSZ + b0],
                                    v2 = lds[sc * PSZ + 2 * CSZ + b0];
                            acc[u][0] += c[qc + 0] * v0 + c[qc + 1] * v1 + c[qc + 2] * v2;
                            acc[u][1] += c[qc + 3] * v0 + c[qc + 4] * v1 + c[qc + 5] * v2;
                            acc[u][2] += c[qc + 6] * v0 + c[qc + 7] * v1 + c[qc + 8] * v2;
                            if constexpr (APPLY) {
                                if (xcopy) {
                                    const unsigned o = po + u * rstep;
                                    at_bytes(xcopy + base, o) = v0;
                                    at_bytes(xcopy + base + d.N, o) = v1;
                                    at_bytes(xcopy + base + 2 * d.N, o) = v2;
                                }
                            }
                        }
                    }
                }
            });
        });

#pragma unroll
        for (int u = 0; u < R; u++) {
            T o0 = acc[u][0], o1 = acc[u][1], o2 = acc[u][2];
            if constexpr (!APPLY) {
                o0 = rv[u][0] - o0;
                o1 = rv[u][1] - o1;
                o2 = rv[u][2] - o2;
            }
            const unsigned o = po + u * rstep;
            at_bytes(out + base, o) = o0;
            at_bytes(out + base + d.N, o) = o1;
            at_bytes(out + base + 2 * d.N, o) = o2;
            if constexpr (NORM) {
                const double s = (double)o0 * o0 + (double)o1 * o1 + (double)o2 * o2;
                q2 += s;
                q4 += s * s;
            }
        }
        if (more) {
            __syncthreads();  // every read of plane k - 1 is done
            store(sm, nx);
            const int t = sm;
            sm = sc;
            sc = sp;
            sp = t;
            __syncthreads();
        }
    }
    if constexpr (NORM)
        block_sum2(q4, q2, part, (long)blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z));
}

// The restriction of k_restrict_march (r_c = (1/8) P^T res, x_c = 0; the
// reference's region::restriction, region.cc:1767-1812) with the fine
// residual staged in LDS, for levels periodic in x and y.  A workgroup owns
// kRestrictTX x kRestrictTY coarse columns and walks a chunk of coarse
// planes through the fine planes 2 K0 - 1 .. 2 K1 - 1 in order: each fine
// plane's window ((2 kRestrictTX + 1) x (2 kRestrictTY + 1) points from
// fine (2 I0 - 1, 2 J0 - 1)) is loaded once, its tile columns with 16-byte
// loads and its halo column per element, the next plane in flight while
// the current one is summed, two LDS slots used in turn.  Each coarse sum
// is formed as in k_restrict_march: the carried partial of the shared odd
// plane, then the terms of planes 2K and 2K + 1 in restrict_point's order,
// the latter also into the next carried partial -- the results are the
// same.  Grid: (mc / kRestrictTX, nc / kRestrictTY, ceil(oc / kchunk)).
constexpr int kRestrictTX = 32, kRestrictTY = kThreads / kRestrictTX;
template <typename T>
__global__ __launch_bounds__(kThreads) void k_restrict_tile(Dims F, Dims C, const T *__restrict__ res,
                                                            T *__restrict__ rc, T *__restrict__ xc, int kchunk) {
    constexpr int RX = kRestrictTX, RY = kRestrictTY;
    constexpr int WX = 2 * RX + 1, WY = 2 * RY + 1, CS = WX * WY, PS = 3 * CS;
    constexpr int VW = 16 / (int)sizeof(T);  // values per 16-byte load
    constexpr int LPR = 2 * RX / VW;         // lanes per window row
    constexpr int RPI = 64 / LPR;            // window rows per wave instruction
    constexpr int NW = kThreads / 64;
    static_assert((2 * RX) % VW == 0 && 64 % LPR == 0, "tile width");
    constexpr int NA = (3 * WY + NW * RPI - 1) / (NW * RPI);
    constexpr int EB = 3 * WY, NB = (EB + kThreads - 1) / kThreads;
    struct alignas(16) Vec {
        T v[VW];
    };
    __shared__ T lds[2 * PS];

    const int I0 = blockIdx.x * RX, J0 = blockIdx.y * RY;
    const int K0 = blockIdx.z * kchunk;
    const int K1 = K0 + kchunk < C.o ? K0 + kchunk : C.o;
    if (K0 >= K1) return;  // the whole workgroup
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const long mn = (long)F.m * F.n;

    // staging offsets, once per workgroup: window row = component-major,
    // columns 1 .. 2 RX (fine x = 2 I0 ..) in chunk A, column 0 (2 I0 - 1,
    // wrapped) in chunk B
    unsigned ga[NA];
    int la[NA];
    const int col = lane % LPR;
#pragma unroll
    for (int i = 0; i < NA; i++) {
        const int row0 = (wv + NW * i) * RPI + lane / LPR;
        const int row = row0 < 3 * WY ? row0 : 3 * WY - 1;
        const int c = row / WY, yy = row - c * WY;
        ga[i] = (unsigned)(sizeof(T) * (c * F.N + (long)F.m * wrap1(2 * J0 - 1 + yy, F.n) + 2 * I0 + col * VW));
        la[i] = c * CS + yy * WX + 1 + col * VW;
    }
    unsigned gb[NB];
    int lb[NB];
#pragma unroll
    for (int j = 0; j < NB; j++) {
        const int e = tid + kThreads * j < EB ? tid + kThreads * j : EB - 1;
        const int c = e / WY, yy = e - c * WY;
        gb[j] = (unsigned)(sizeof(T) * (c * F.N + (long)F.m * wrap1(2 * J0 - 1 + yy, F.n) +
                                         wrap1(2 * I0 - 1, F.m)));
        lb[j] = c * CS + yy * WX;
    }
    struct Stage {
        Vec a[NA];
        T b[NB];
    };
    auto fetch = [&](int fk, Stage &v) {
        const T *base = res + mn * fk;
#pragma unroll
        for (int i = 0; i < NA; i++) v.a[i] = at_bytes(reinterpret_cast<const Vec *>(base), ga[i]);
#pragma unroll
        for (int j = 0; j < NB; j++) v.b[j] = at_bytes(base, gb[j]);
    };
    auto store = [&](int slot, const Stage &v) {
        T *sl = lds + slot * PS;
#pragma unroll
        for (int i = 0; i < NA; i++)
            if ((3 * WY) % (NW * RPI) == 0 || (wv + NW * i) * RPI + lane / LPR < 3 * WY) {
#pragma unroll
                for (int e = 0; e < VW; e++) sl[la[i] + e] = v.a[i].v[e];
            }
#pragma unroll
        for (int j = 0; j < NB; j++)
            if (EB % kThreads == 0 || tid + kThreads * j < EB) sl[lb[j]] = v.b[j];
    };
    // fine plane f of the walk, wrapped where z is periodic; false where it
    // lies outside a walled z range (its terms are left out, as in
    // k_restrict_march)
    auto fine_plane = [&](int f, int &fw) {
        fw = f;
        if (f < 0 || f >= F.o) {
            if (!F.pz) return false;
            fw = f < 0 ? f + F.o : f - F.o;
        }
        return true;
    };

    // the terms of the fine plane in slot sl with z weight wk, in
    // restrict_point's order (fine rows -1, 0, +1, then columns)
    const int tx = tid % RX, ty = tid / RX;
    auto terms = [&](int slot, T wk, T &s0, T &s1, T &s2, bool part, T &q0, T &q1, T &q2) {
        const T *sl = lds + slot * PS;
#pragma unroll
        for (int b = 0; b < 3; b++) {
            const T wj = b == 1 ? wk : T(0.5) * wk;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const T w = T(0.125) * (a == 1 ? wj : T(0.5) * wj);
                const int q = (2 * ty + b) * WX + 2 * tx + a;
                const T v0 = sl[q], v1 = sl[CS + q], v2 = sl[2 * CS + q];
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

    const int fa = 2 * K0 - 1, fb = 2 * K1 - 1;
    T c0 = 0, c1 = 0, c2 = 0;  // carried partial of the shared odd plane
    T s0 = 0, s1 = 0, s2 = 0;
    T z0 = 0, z1 = 0, z2 = 0;  // unused second partial of the plane sums
    Stage cur, nxt;
    int fw;
    bool vcur = fine_plane(fa, fw);
    if (vcur) fetch(fw, cur);
    for (int f = fa; f <= fb; f++) {
        int fw2 = 0;
        const bool vnext = f + 1 <= fb && fine_plane(f + 1, fw2);
        if (vnext) fetch(fw2, nxt);
        const int slot = f & 1;
        // the slot's previous plane (f - 2) was summed before the barrier
        // of the last iteration
        if (vcur) store(slot, cur);
        __syncthreads();
        if (f == fa) {
            if (vcur) terms(slot, T(0.5), c0, c1, c2, false, z0, z1, z2);
        } else if (((f - fa) & 1) == 1) {
            // f = 2K (always inside the fine range)
            s0 = c0;
            s1 = c1;
            s2 = c2;
            terms(slot, T(1), s0, s1, s2, false, z0, z1, z2);
        } else {
            // f = 2K + 1: completes coarse plane K
            T q0 = 0, q1 = 0, q2 = 0;
            if (vcur) terms(slot, T(0.5), s0, s1, s2, true, q0, q1, q2);
            c0 = q0;
            c1 = q1;
            c2 = q2;
            const int K = (f - 1) / 2;
            const long p = (I0 + tx) + (long)C.m * ((J0 + ty) + (long)C.n * K);
            rc[p] = s0;
            rc[p + C.N] = s1;
            rc[p + 2 * C.N] = s2;
            xc[p] = 0;
            xc[p + C.N] = 0;
            xc[p + 2 * C.N] = 0;
        }
        cur = nxt;
        vcur = vnext;
    }
}

}  // namespace stzmg
