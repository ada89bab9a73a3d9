// Pass 2 of the QS step as a plane-marching stencil kernel (gfx950, FP64):
// ENO2 advection of stress and chi, the Truesdell (or transformed L and T)
// terms, the plastic update and chi diffusion, merged into new arrays.
//
// Reference loops: advection_grid_step and chi_diffusion of
// src/shear_sim_3d.cpp:409-525, :1217-1237 (plain step) and
// src/trans_sim_3d.cpp:779-812, :880-992 (transformed step); merge_updates
// src/shear_sim_3d.cpp:154-162.
//
// Work split.  A workgroup of 256 threads owns a tile of kAX x kAY cells of
// the xy plane (one wave per row, lanes along x) and marches it through a
// chunk of z planes.  Per plane k:
//   * the 7 cell fields (6 stress components, chi) of plane k, with a halo
//     of two cells in x and y, are staged in LDS: the tile's own points from
//     registers, the halo from HBM; the x and y ENO stencils and the chi
//     diffusion read them there;
//   * each thread keeps its own column's values of planes k-2 .. k+2 in
//     registers (one new plane loaded per step), which serve the z stencils;
//   * the velocity vertices of planes k and k+1 ((kAX+1) x (kAY+1) each) sit
//     in LDS, two buffers used in turn, so each vertex plane is read from
//     HBM once per chunk;
//   * the transformed step also stages chi of planes k-1 and k+1 with a halo
//     of one cell for the mixed second differences.
// Periodic images are taken when staging (x and y always, z when periodic);
// across a z wall the ghost planes are the reference's linear extrapolation
// f(-1) = 2 f(0) - f(1), f(-2) = 2 f(-1) - f(0) and its mirror at the top
// (set_proper_ghost_for_boundaries, src/shear_sim_3d.cpp:2662-2789).  Tiles
// that stick out of the grid (nx or ny not a multiple of the tile) load
// wrapped positions in the unused lanes and store nothing there.
//
// ad_Dpl follows the reference's in-place lexicographic loop: this step's
// value at the cell and at its -x/-y/-z neighbours inside the domain (dnew),
// last step's value at +x/+y/+z and at periodic images (dold), zero across a
// wall.  The transformed step's chi advection takes f(+1) as the f(+2) of
// its x derivative (src/trans_sim_3d.cpp:985) and its mixed y-z difference
// keeps the reference's two equal last terms (:802).
#pragma once

#include "stz_kernels.hpp"

namespace stz {

constexpr int kAX = 64, kAY = 4;                  // cells per tile in x, y
constexpr int kSX = kAX + 4, kSY = kAY + 4;       // staged plane with halo 2
constexpr int kVX = kAX + 1, kVY = kAY + 1;       // vertex tile
constexpr int kCX = kAX + 2, kCY = kAY + 2;       // chi planes k+-1 with halo 1
constexpr int kHalo = kSX * kSY - kAX * kAY;      // halo points of a staged plane
static_assert(kAX * kAY == kT, "one thread per tile cell");
static_assert(kAX == 64, "one wave per tile row");

// Tiles of the launch: ntx x nty tiles per plane, nzc chunks of kchunk planes.
struct AdvTiles {
    int ntx, nty, nzc, kchunk;
};

inline AdvTiles adv_tiles(const Grid &g, int resident) {
    AdvTiles t;
    t.ntx = (g.Nx + kAX - 1) / kAX;
    t.nty = (g.Ny + kAY - 1) / kAY;
    const int per = t.ntx * t.nty;
    int nzc = resident / per;
    nzc = nzc < 1 ? 1 : (nzc > g.Nz ? g.Nz : nzc);
    t.kchunk = (g.Nz + nzc - 1) / nzc;
    t.nzc = (g.Nz + t.kchunk - 1) / t.kchunk;
    return t;
}

// ((i mod n) + n) mod n for any i (staging positions of tiles larger than
// the grid reach beyond one period).
__device__ __forceinline__ int wrapm(int i, int n) {
    const int r = i % n;
    return r < 0 ? r + n : r;
}

// Staged-plane position (sx, sy) of halo point h of a kSX x kSY plane: rows
// 0, 1 and kSY-2, kSY-1 in full, then columns 0, 1, kSX-2, kSX-1 of the
// middle rows.
__device__ __forceinline__ void halo_pos(int h, int &sx, int &sy) {
    constexpr int edge = 2 * kSX;
    if (h < edge) {
        sy = h / kSX;
        sx = h - sy * kSX;
    } else if (h < 2 * edge) {
        const int t = h - edge;
        sy = kSY - 2 + t / kSX;
        sx = t - (t / kSX) * kSX;
    } else {
        const int t = h - 2 * edge;
        sy = 2 + (t >> 2);
        const int c = t & 3;
        sx = c < 2 ? c : kSX - 4 + c;
    }
}

// The transform of the step with the products every cell uses, formed once
// on the host: Tti = T^-t, TdT = T^-1 dT/dt, TdTt = (T^-1 dT/dt)^t and the
// diffusion metric J = T^-1 T^-t.
struct XformD {
    double T[9], Ti[9], dT[9], Tti[9], TdT[9], TdTt[9], J[9];
};

inline XformD xform_products(const Xform &X) {
    auto mul = [](const double *a, const double *b, double *o) {
        double r[9];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) r[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
        for (int e = 0; e < 9; e++) o[e] = r[e];
    };
    auto tr = [](const double *a, double *o) {
        const double r[9] = {a[0], a[3], a[6], a[1], a[4], a[7], a[2], a[5], a[8]};
        for (int e = 0; e < 9; e++) o[e] = r[e];
    };
    XformD D;
    for (int e = 0; e < 9; e++) {
        D.T[e] = X.T[e];
        D.Ti[e] = X.Ti[e];
        D.dT[e] = X.dT[e];
    }
    tr(X.Ti, D.Tti);
    mul(X.Ti, X.dT, D.TdT);
    tr(D.TdT, D.TdTt);
    mul(X.Ti, D.Tti, D.J);
    return D;
}

template <bool TRANS, bool PZ>
static __global__ __launch_bounds__(kT) void k_advect_march(
    Grid g, Mat M, double dt, XformD X, Box B, AdvTiles T, const double *__restrict__ vel,
    const double *__restrict__ sig, const double *__restrict__ chi, const double *__restrict__ adapt,
    const double *__restrict__ dchi, const double *__restrict__ dnew, const double *__restrict__ dold,
    double *__restrict__ sig_out, double *__restrict__ chi_out) {
    constexpr int NF = 7;
    __shared__ double sc[NF][kSY * kSX];
    __shared__ double sv[2][3][kVY * kVX];
    __shared__ double sxz[TRANS ? 2 : 1][TRANS ? kCY * kCX : 1];

    const long Lb = xcd_block();
    const int tx = (int)(Lb % T.ntx), ty = (int)((Lb / T.ntx) % T.nty), zc = (int)(Lb / ((long)T.ntx * T.nty));
    const int k0 = zc * T.kchunk;
    if (k0 >= g.Nz) return;
    const int k1 = k0 + T.kchunk < g.Nz ? k0 + T.kchunk : g.Nz;
    const int i0 = tx * kAX, j0 = ty * kAY;
    const int tid = threadIdx.x, lx = tid % kAX, ly = tid / kAX;
    const int i = i0 + lx, j = j0 + ly;
    const bool act = i < g.Nx && j < g.Ny;
    const int nx = g.Nx, ny = g.Ny, nz = g.Nz;
    const int nxy = nx * ny;
    const int col = wrapm(i, nx) + nx * wrapm(j, ny);  // load position of this lane
    const long nc = g.Nc, nv = g.Nv;  // component strides (cell and vertex indices fit in int)

    auto stage_vertices = [&](int kv) {
        const int kw = PZ ? wrapc(kv, nz) : kv;
        double(*dst)[kVY * kVX] = sv[kv & 1];
        for (int p = tid; p < kVX * kVY; p += kT) {
            const int vy = p / kVX, vx = p - vy * kVX;
            const int q = wrapm(i0 + vx, nx) + nx * wrapm(j0 + vy, ny) + nxy * kw;
#pragma unroll
            for (int d = 0; d < 3; d++) dst[d][p] = vel[d * nv + q];
        }
    };
    // field q (0..5 stress components, 6 chi) at offset p
    auto F = [&](int q, int p) -> double { return q < 6 ? sig[q * nc + p] : chi[p]; };

    // z window of this thread's column: w[q][s] holds plane k - 2 + s.
    // Across a wall the ghost planes follow the reference's extrapolation
    // f(-1) = 2 f(0) - f(1), f(-2) = 2 f(-1) - f(0) and its mirror at the
    // top, each formed from the two planes next to it in the window.
    double w[NF][5];
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const int kk = k0 - 2 + s;
        if (PZ || (kk >= 0 && kk < nz)) {
            const int p = col + nxy * (PZ ? wrapc(kk, nz) : kk);
#pragma unroll
            for (int q = 0; q < NF; q++) w[q][s] = F(q, p);
        }
    }
    if (!PZ) {
        if (k0 + 1 >= nz) {
#pragma unroll
            for (int q = 0; q < NF; q++) w[q][3] = (w[q][2] + w[q][2]) - w[q][1];
        }
#pragma unroll
        for (int s = 1; s >= 0; s--)
            if (k0 - 2 + s < 0) {
#pragma unroll
                for (int q = 0; q < NF; q++) w[q][s] = (w[q][s + 1] + w[q][s + 1]) - w[q][s + 2];
            }
    }
    stage_vertices(k0);

    const double hx = g.dx / dt, hy = g.dy / dt, hz = g.dz / dt;
    const int me = (ly + 2) * kSX + lx + 2;
    for (int k = k0; k < k1; k++) {
        if (PZ || k + 2 < nz) {
            const int p = col + nxy * (PZ ? wrapc(k + 2, nz) : k + 2);
#pragma unroll
            for (int q = 0; q < NF; q++) w[q][4] = F(q, p);
        } else {
#pragma unroll
            for (int q = 0; q < NF; q++) w[q][4] = (w[q][3] + w[q][3]) - w[q][2];
        }
        __syncthreads();  // the previous plane's LDS reads are done
#pragma unroll
        for (int q = 0; q < NF; q++) sc[q][me] = w[q][2];
        for (int h = tid; h < kHalo; h += kT) {
            int sx, sy;
            halo_pos(h, sx, sy);
            const int p = wrapm(i0 + sx - 2, nx) + nx * wrapm(j0 + sy - 2, ny) + nxy * k;
#pragma unroll
            for (int q = 0; q < NF; q++) sc[q][sy * kSX + sx] = F(q, p);
        }
        stage_vertices(k + 1);
        if constexpr (TRANS) {
            // chi of planes k-1 and k+1 with a halo of one cell; wall ghosts
            // extrapolated as in the window
            const bool gm = !PZ && k == 0, gp = !PZ && k + 1 == nz;
            const int km = PZ ? wrapc(k - 1, nz) : k - 1, kp = PZ ? wrapc(k + 1, nz) : k + 1;
            for (int p = tid; p < kCX * kCY; p += kT) {
                const int cy = p / kCX, cx = p - cy * kCX;
                const int cc = wrapm(i0 + cx - 1, nx) + nx * wrapm(j0 + cy - 1, ny);
                if (gm) {
                    const double f0 = chi[cc], f1 = chi[cc + nxy];
                    sxz[0][p] = (f0 + f0) - f1;
                } else {
                    sxz[0][p] = chi[cc + nxy * km];
                }
                if (gp) {
                    const double a = chi[cc + nxy * (nz - 1)], b = chi[cc + nxy * (nz - 2)];
                    sxz[1][p] = (a + a) - b;
                } else {
                    sxz[1][p] = chi[cc + nxy * kp];
                }
            }
        }
        __syncthreads();

        if (act) {
            const int c = i + nx * j + nxy * k;
            auto S = [&](int q, int dx, int dy) -> double { return sc[q][me + dy * kSX + dx]; };
            // vertex d of (i + a, j + b, k + e)
            auto V = [&](int d, int a, int b, int e) -> double {
                return sv[(k + e) & 1][d][(ly + b) * kVX + lx + a];
            };
            const double ix = 1. / g.dx, iy = 1. / g.dy, iz = 1. / g.dz;
            if constexpr (!TRANS) {
                // velocity gradient and cell velocity (vel_grad_o)
                double G[3][3], v[3];
                {
                    const double gx = .25 * (1.0 / g.dx), gy = .25 * (1.0 / g.dy), gz = .25 * (1.0 / g.dz);
#pragma unroll
                    for (int q = 0; q < 3; q++) {
                        const double u000 = V(q, 0, 0, 0), u100 = V(q, 1, 0, 0);
                        const double u010 = V(q, 0, 1, 0), u110 = V(q, 1, 1, 0);
                        const double u001 = V(q, 0, 0, 1), u101 = V(q, 1, 0, 1);
                        const double u011 = V(q, 0, 1, 1), u111 = V(q, 1, 1, 1);
                        G[q][0] = gx * (u100 - u000 + u110 - u010 + u101 - u001 + u111 - u011);
                        G[q][1] = gy * (u010 - u000 + u110 - u100 + u011 - u001 + u111 - u101);
                        G[q][2] = gz * (u001 - u000 + u101 - u100 + u011 - u010 + u111 - u110);
                        v[q] = 0.125 * (u000 + u100 + u001 + u101 + u010 + u110 + u011 + u111);
                    }
                }
                double ug[7];
#pragma unroll
                for (int q = 0; q < 7; q++) {
                    const double f0 = S(q, 0, 0);
                    const double ddx = eno2(hx, v[0], S(q, 2, 0), S(q, 1, 0), f0, S(q, -1, 0), S(q, -2, 0));
                    const double ddy = eno2(hy, v[1], S(q, 0, 2), S(q, 0, 1), f0, S(q, 0, -1), S(q, 0, -2));
                    const double ddz = eno2(hz, v[2], w[q][4], w[q][3], f0, w[q][1], w[q][0]);
                    ug[q] = v[0] * ddx + v[1] * ddy + v[2] * ddz;
                }
                const double a = adapt[c];
                {
                    const double s11 = S(0, 0, 0), s12 = S(1, 0, 0), s13 = S(2, 0, 0);
                    const double s22 = S(3, 0, 0), s23 = S(4, 0, 0), s33 = S(5, 0, 0);
                    const double du_dx = G[0][0], du_dy = G[0][1], du_dz = G[0][2];
                    const double dv_dx = G[1][0], dv_dy = G[1][1], dv_dz = G[1][2];
                    const double dw_dx = G[2][0], dw_dy = G[2][1], dw_dz = G[2][2];
                    const double t11 = s11 * (du_dx - dv_dy - dw_dz) + 2 * s13 * du_dz + 2 * s12 * du_dy;
                    const double t12 = s23 * du_dz + s13 * dv_dz - s12 * dw_dz + s22 * du_dy + s11 * dv_dx;
                    const double t13 = s33 * du_dz + s11 * dw_dx + s23 * du_dy + s12 * dw_dy - s13 * dv_dy;
                    const double t22 = 2 * s23 * dv_dz + s22 * (dv_dy - du_dx - dw_dz) + 2 * s12 * dv_dx;
                    const double t23 = s33 * dv_dz + s22 * dw_dy + s13 * dv_dx - s23 * du_dx + s12 * dw_dx;
                    const double t33 = s33 * (dw_dz - dv_dy - du_dx) + 2 * s23 * dw_dy + 2 * s13 * dw_dx;
                    const double third = 1. / 3.;
                    sig_out[c] = s11 + (a * third * (s33 + s22 - 2 * s11) - ug[0] + dt * t11);
                    sig_out[nc + c] = s12 + (-a * s12 - ug[1] + dt * t12);
                    sig_out[2 * nc + c] = s13 + (-a * s13 - ug[2] + dt * t13);
                    sig_out[3 * nc + c] = s22 + (a * third * (s11 - 2 * s22 + s33) - ug[3] + dt * t22);
                    sig_out[4 * nc + c] = s23 + (-a * s23 - ug[4] + dt * t23);
                    sig_out[5 * nc + c] = s33 + (a * third * (s11 + s22 - 2 * s33) - ug[5] + dt * t33);
                }
                // chi diffusion (chi_diffusion) with the in-place ad_Dpl
                // neighbours
                const double Dc = dnew[c];
                const double Dxm = i > 0 ? dnew[c - 1] : dold[c + (nx - 1)];
                const double Dxp = dold[c + (i + 1 < nx ? 1 : 1 - nx)];
                const double Dym = j > 0 ? dnew[c - nx] : dold[c + (ny - 1) * nx];
                const double Dyp = dold[c + (j + 1 < ny ? nx : -(ny - 1) * nx)];
                double Dzm, Dzp;
                if (PZ) {
                    Dzm = k > 0 ? dnew[c - nxy] : dold[c + (nz - 1) * nxy];
                    Dzp = dold[c + (k + 1 < nz ? nxy : -(nz - 1) * nxy)];
                } else {
                    Dzm = k > 0 ? dnew[c - nxy] : 0.0;
                    Dzp = k < nz - 1 ? dold[c + nxy] : 0.0;
                }
                const double x0 = S(6, 0, 0);
                const double xxm = S(6, -1, 0), xxp = S(6, 1, 0);
                const double xym = S(6, 0, -1), xyp = S(6, 0, 1);
                const double xzm = w[6][1], xzp = w[6][3];
                const double ddx = .5 * ix * (Dxp - Dxm) * (.5 * ix * (xxp - xxm)) + Dc * (ix * ix * (xxp - 2 * x0 + xxm));
                const double ddy = .5 * iy * (Dyp - Dym) * (.5 * iy * (xyp - xym)) + Dc * (iy * iy * (xyp - 2 * x0 + xym));
                const double ddz = .5 * iz * (Dzp - Dzm) * (.5 * iz * (xzp - xzm)) + Dc * (iz * iz * (xzp - 2 * x0 + xzm));
                const double diff = M.diff_l * M.diff_l * (ddx + ddy + ddz);
                double cchi = dchi[c];
                cchi -= ug[6];
                cchi += diff / M.c0;
                chi_out[c] = x0 + cchi;
            } else {
                // cell velocity
                double v[3];
#pragma unroll
                for (int q = 0; q < 3; q++)
                    v[q] = 0.125 * (V(q, 0, 0, 0) + V(q, 1, 0, 0) + V(q, 0, 0, 1) + V(q, 1, 0, 1) + V(q, 0, 1, 0) +
                                    V(q, 1, 1, 0) + V(q, 0, 1, 1) + V(q, 1, 1, 1));
                // ENO advection of sigma'
                double ug[6];
#pragma unroll
                for (int q = 0; q < 6; q++) {
                    const double f0 = S(q, 0, 0);
                    const double ddx = eno2(hx, v[0], S(q, 2, 0), S(q, 1, 0), f0, S(q, -1, 0), S(q, -2, 0));
                    const double ddy = eno2(hy, v[1], S(q, 0, 2), S(q, 0, 1), f0, S(q, 0, -1), S(q, 0, -2));
                    const double ddz = eno2(hz, v[2], w[q][4], w[q][3], f0, w[q][1], w[q][0]);
                    ug[q] = v[0] * ddx + v[1] * ddy + v[2] * ddz;
                }
                // L = T^-t grad_X(dT/dt X + T U) over the cell (compute_L)
                double L[9];
                {
                    const double P0[3] = {B.ax + ((double)i + .5) * g.dx, B.ay + ((double)j + .5) * g.dy,
                                          B.az + ((double)k + .5) * g.dz};
                    const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                           {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
                    const double pn[3] = {P0[0] == B.bx ? B.ax : P0[0] + g.dx, P0[1] == B.by ? B.ay : P0[1] + g.dy,
                                          (PZ && P0[2] == B.bz) ? B.az : P0[2] + g.dz};
                    double vv[8][3];
#pragma unroll
                    for (int n = 0; n < 8; n++) {
                        double P[3], Vl[3], av[3], cv[3];
#pragma unroll
                        for (int d = 0; d < 3; d++) {
                            P[d] = off[n][d] ? pn[d] : P0[d];
                            Vl[d] = V(d, off[n][0], off[n][1], off[n][2]);
                        }
                        m3vec(X.dT, P, av);
                        m3vec(X.T, Vl, cv);
#pragma unroll
                        for (int d = 0; d < 3; d++) vv[n][d] = av[d] + cv[d];
                    }
                    double Gm[9];
                    const double gx = .25 * (1. / g.dx), gy = .25 * (1. / g.dy), gz = .25 * (1. / g.dz);
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        Gm[d] = gx * (((((((vv[1][d] - vv[0][d]) + vv[2][d]) - vv[3][d]) + vv[5][d]) - vv[4][d]) +
                                       vv[6][d]) - vv[7][d]);
                        Gm[3 + d] = gy * (((((((vv[3][d] - vv[0][d]) + vv[2][d]) - vv[1][d]) + vv[7][d]) -
                                           vv[4][d]) + vv[6][d]) - vv[5][d]);
                        Gm[6 + d] = gz * (((((((vv[4][d] - vv[0][d]) + vv[5][d]) - vv[1][d]) + vv[7][d]) -
                                           vv[3][d]) + vv[6][d]) - vv[2][d]);
                    }
                    m3mul(X.Tti, Gm, L);
                }
                // sigma' and the physical stress T sigma' T^t (old_sig)
                double sp[9], os[9], tmp[9], A[9], Bm[9];
                {
                    const double s11 = S(0, 0, 0), s12 = S(1, 0, 0), s13 = S(2, 0, 0);
                    const double s22 = S(3, 0, 0), s23 = S(4, 0, 0), s33 = S(5, 0, 0);
                    const double qv[9] = {s11, s12, s13, s12, s22, s23, s13, s23, s33};
#pragma unroll
                    for (int e = 0; e < 9; e++) sp[e] = qv[e];
                    double Tt[9];
                    m3t(X.T, Tt);
                    m3mul(X.T, sp, tmp);
                    m3mul(tmp, Tt, os);
                }
                // L term T^-1 (-tr(L) sigma + sigma L + L^t sigma) T^-t and T
                // term -(T^-1 dT sigma' + sigma' (T^-1 dT)^t)
                const int comp[6] = {0, 1, 2, 4, 5, 8};
                double cs[6];
                {
                    double Lt[9], Lterm[9];
                    const double trL = L[0] + L[4] + L[8];
                    m3t(L, Lt);
                    m3mul(os, L, A);
                    m3mul(Lt, os, Bm);
#pragma unroll
                    for (int e = 0; e < 9; e++) tmp[e] = (os[e] * -(trL) + A[e]) + Bm[e];
                    m3mul(X.Ti, tmp, Lt);
                    m3mul(Lt, X.Tti, Lterm);
                    m3mul(X.TdT, sp, A);
                    m3mul(sp, X.TdTt, Bm);
#pragma unroll
                    for (int q = 0; q < 6; q++) {
                        const int e = comp[q];
                        cs[q] = 0.0 + (dt * (Lterm[e] + (A[e] + Bm[e]) * -1.) - ug[q]);
                    }
                }
                const double a = adapt[c];
                if (a != 0.0) {
                    const double third = 1. / 3., tr = os[0] + os[4] + os[8];
                    const double dv[9] = {os[0] - third * tr, os[1], os[2], os[1], os[4] - third * tr, os[5],
                                          os[2], os[5], os[8] - third * tr};
                    double Ta[9], P[9];
#pragma unroll
                    for (int e = 0; e < 9; e++) Ta[e] = X.Ti[e] * a;
                    m3mul(Ta, dv, tmp);
                    m3mul(tmp, X.Tti, P);
#pragma unroll
                    for (int q = 0; q < 6; q++) cs[q] -= P[comp[q]];
                }
#pragma unroll
                for (int q = 0; q < 6; q++) sig_out[q * nc + c] = S(q, 0, 0) + cs[q];
                // chi: advection (the x derivative takes f(+1) as its f(+2))
                // and diffusion with the metric J
                const double x0 = S(6, 0, 0);
                const double xxm = S(6, -1, 0), xxp = S(6, 1, 0);
                const double xym = S(6, 0, -1), xyp = S(6, 0, 1);
                const double xzm = w[6][1], xzp = w[6][3];
                const double cx = eno2(hx, v[0], xxp, xxp, x0, xxm, S(6, -2, 0));
                const double cy = eno2(hy, v[1], S(6, 0, 2), xyp, x0, xym, S(6, 0, -2));
                const double cz = eno2(hz, v[2], w[6][4], xzp, x0, xzm, w[6][0]);
                const double Dc = dnew[c];
                const double Dxm = i > 0 ? dnew[c - 1] : dold[c + (nx - 1)];
                const double Dxp = dold[c + (i + 1 < nx ? 1 : 1 - nx)];
                const double Dym = j > 0 ? dnew[c - nx] : dold[c + (ny - 1) * nx];
                const double Dyp = dold[c + (j + 1 < ny ? nx : -(ny - 1) * nx)];
                double Dzm, Dzp;
                if (PZ) {
                    Dzm = k > 0 ? dnew[c - nxy] : dold[c + (nz - 1) * nxy];
                    Dzp = dold[c + (k + 1 < nz ? nxy : -(nz - 1) * nxy)];
                } else {
                    Dzm = k > 0 ? dnew[c - nxy] : 0.0;
                    Dzp = k < nz - 1 ? dold[c + nxy] : 0.0;
                }
                auto CZ = [&](int e, int dx, int dy) -> double {  // chi at (i+dx, j+dy, k+e), e = +-1
                    return sxz[e > 0][(ly + 1 + dy) * kCX + lx + 1 + dx];
                };
                const double dDx = .5 * ix * (Dxp - Dxm), dDy = .5 * iy * (Dyp - Dym), dDz = .5 * iz * (Dzp - Dzm);
                const double gc[3] = {.5 * ix * (xxp - xxm), .5 * iy * (xyp - xym), .5 * iz * (xzp - xzm)};
                double Jg[3];
                m3vec(X.J, gc, Jg);
                const double term1 = dDx * Jg[0] + dDy * Jg[1] + dDz * Jg[2];
                const double dxx = ix * ix * (xxp - 2 * x0 + xxm);
                const double dyy = iy * iy * (xyp - 2 * x0 + xym);
                const double dzz = iz * iz * (xzp - 2 * x0 + xzm);
                const double dxy = .25 * ix * iy * (S(6, 1, 1) + S(6, -1, -1) - S(6, 1, -1) - S(6, -1, 1));
                const double dxz = .25 * ix * iz * (CZ(1, 1, 0) + CZ(-1, -1, 0) - CZ(-1, 1, 0) - CZ(1, -1, 0));
                const double ypzm = CZ(-1, 0, 1), yzpp = CZ(1, 0, 1), yzmm = CZ(-1, 0, -1);
                // the mixed y-z difference of src/trans_sim_3d.cpp:802, whose
                // last two terms are the same cell
                const double dyz = .25 * iy * iz * (yzpp + yzmm - ypzm + ypzm);
                double term2 = 0;
                term2 += X.J[0] * dxx + X.J[1] * dxy + X.J[2] * dxz;
                term2 += X.J[3] * dxy + X.J[4] * dyy + X.J[5] * dyz;
                term2 += X.J[6] * dxz + X.J[7] * dyz + X.J[8] * dzz;
                term2 *= Dc;
                const double diff = M.diff_l * M.diff_l * (term1 + term2);
                double cchi = dchi[c];
                cchi -= v[0] * cx + v[1] * cy + v[2] * cz;
                cchi += diff / M.c0;
                chi_out[c] = x0 + cchi;
            }
        }
#pragma unroll
        for (int q = 0; q < NF; q++)
#pragma unroll
            for (int s = 0; s < 4; s++) w[q][s] = w[q][s + 1];
    }
}

}  // namespace stz
