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

}  // namespace stzmg
