// Direct (explicit dynamics) step of shear_sim_3d (include/stz.h,
// stz_step_direct): step_forward, src/shear_sim_3d.cpp:206-213, with
// calc_updates :89-104, step_stress :774-896, step_chi :1193-1215,
// step_velocity :919-994, step_ref :997-1030 and merge_updates :154-162.
//
// Per step, on the handle's stream:
//   k_direct_plastic  STZ plasticity of every cell: adapt, the plastic chi
//                     change and this step's ad_Dpl (adapt / 2 / mu / dt sbar)
//   k_direct_cell     stress and chi: elastic, Truesdell, ENO2 advection and
//                     plastic terms; chi advection and dt chi_diffusion / c0
//                     with the reference's in-place ad_Dpl order (this step's
//                     value at -x/-y/-z inside the domain, last step's at
//                     +x/+y/+z and at periodic images, zero across a wall)
//   k_direct_vertex   velocity (dt/rho div sigma + dt kappa/rho Lap u -
//                     u.grad u) and reference map (- u.grad xi) at the
//                     vertices 0 < k < Nz; the plane k = 0 is left as it is
//                     (the reference skips zz = 0 when its x coordinate in
//                     the rank grid is 0, :101, i.e. always with one rank)
//   k_direct_plate    top plate velocity and reference map at t + dt
//                     (set_boundary_conditions, :2393-2466; with one rank in
//                     z the bottom branch is never taken)
// Ghosts are computed where they are read: periodic images (the reference
// map shifted by the domain length, :1470-1500), linear extrapolation of
// velocity across the walls (the top plate's image from the plate value,
// :2461-2463), and zero reference map beyond the walls (an extrapolated
// Field carries only the ten primary members, src/field.hh:49-112).
#include <hip/hip_runtime.h>

#include <cmath>
#include <utility>

#include "../../include/stz.h"

// Every product and sum rounds on its own, as the reference's host code
// does, so that the explicit step matches it bit for bit where no
// transcendental function enters.
#pragma clang fp contract(off)

#include "stz_internal.hpp"

using namespace stz;

namespace {

inline unsigned nbk(long n) { return (unsigned)((n + kT - 1) / kT); }

struct Dyn {
    double rho, kappa, U;
    double ax, ay, az, Lx, Ly, Lz;
};

// Velocity component at vertex (i, j, k) with the wall ghosts of the
// reference's grid: k = -1, -2 extrapolated from planes 0, 1; k = Nz + 1 from
// the top plate and plane Nz - 1.
__device__ __forceinline__ double vel_g(const double *__restrict__ u, const Grid &g, int i, int j, int k) {
    if (g.pz || (k >= 0 && k <= g.Nz)) return vertv(u, g, i, j, k);
    if (k < 0) {
        const double a = vertv(u, g, i, j, 0), b = vertv(u, g, i, j, 1);
        const double m1 = (a + a) - b;
        return k == -1 ? m1 : (m1 + m1) - a;
    }
    const double a = vertv(u, g, i, j, g.Nz), b = vertv(u, g, i, j, g.Nz - 1);
    return 2 * a - b;
}

// Reference-map component q at vertex (i, j, k): periodic images shifted by
// the domain length, zero beyond a wall.
__device__ __forceinline__ double xi_g(const double *__restrict__ xi, const Grid &g, const Dyn &D, int q, int i,
                                       int j, int k) {
    double off = 0;
    if (q == 0) off = i >= g.Nx ? D.Lx : (i < 0 ? -D.Lx : 0.0);
    if (q == 1) off = j >= g.Ny ? D.Ly : (j < 0 ? -D.Ly : 0.0);
    if (g.pz) {
        if (q == 2) off = k >= g.Nz ? D.Lz : (k < 0 ? -D.Lz : 0.0);
    } else if (k < 0 || k > g.Nz) {
        return 0.0;
    }
    const double v = vertv(xi + (long)q * g.Nv, g, i, j, k);
    return off == 0.0 ? v : v + off;
}

static __global__ __launch_bounds__(kT) void k_direct_plastic(Grid g, Mat M, double dt,
                                                              const double *__restrict__ sig,
                                                              const double *__restrict__ chi,
                                                              double *__restrict__ adapt, double *__restrict__ dchi,
                                                              double *__restrict__ dpl_new,
                                                              int *__restrict__ counters) {
    const long c = xcd_block() * kT + threadIdx.x;
    if (c >= g.Nc) return;
    double s[6];
#pragma unroll
    for (int q = 0; q < 6; q++) s[q] = sig[q * g.Nc + c];
    const double sb = sbar_of(s);
    double a = 0, dc = 0, D = 0;
    if (sb > M.sy) {
        int n;
        a = adaptive(M, sb, chi[c], dt, dc, n);
        D = a / 2. / M.mu / dt * sb;
    }
    count_wave(&counters[0], sb > M.sy);
    adapt[c] = a;
    dchi[c] = dc;
    dpl_new[c] = D;
}

static __global__ __launch_bounds__(kT) void k_direct_cell(Grid g, Mat M, double dt, const double *__restrict__ vel,
                                                           const double *__restrict__ sig,
                                                           const double *__restrict__ chi,
                                                           const double *__restrict__ adapt,
                                                           const double *__restrict__ dchi,
                                                           const double *__restrict__ dnew,
                                                           const double *__restrict__ dold,
                                                           double *__restrict__ sig_out, double *__restrict__ chi_out,
                                                           BrickGrid B) {
    int i, j, k;
    if (!brick_point(B, g.Nx, g.Ny, i, j, k)) return;
    const long c = i + (long)g.Nx * (j + (long)g.Ny * k);
    const Stencil5 S = stencil5(g, i, j, k, g.Nz);
    // away from the walls (uniform over a brick tile, which is one plane)
    const bool zin = g.pz || (k >= 2 && k <= g.Nz - 3);
    double G[3][3], v[3];
    vel_grad_o(vel, g, c, S.ox[3], S.oy[3], S.oz[3], G, v);
    const double dtlamb = dt * M.lamb, twomudt = 2 * M.mu * dt, dtmu = dt * M.mu;
    const double hx = g.dx / dt, hy = g.dy / dt, hz = g.dz / dt;
    double ug[7];
#pragma unroll
    for (int q = 0; q < 7; q++) {
        const double *f = q < 6 ? sig + q * g.Nc : chi;
        const double *fc = f + c;
        const double f0 = fc[0];
        const double ddx = eno2(hx, v[0], fc[S.ox[4]], fc[S.ox[3]], f0, fc[S.ox[1]], fc[S.ox[0]]);
        const double ddy = eno2(hy, v[1], fc[S.oy[4]], fc[S.oy[3]], f0, fc[S.oy[1]], fc[S.oy[0]]);
        double z2p, z1p, z1m, z2m;
        if (zin) {
            z2p = fc[S.oz[4]];
            z1p = fc[S.oz[3]];
            z1m = fc[S.oz[1]];
            z2m = fc[S.oz[0]];
        } else {
            z2p = cellv(f, g, i, j, k + 2);
            z1p = cellv(f, g, i, j, k + 1);
            z1m = cellv(f, g, i, j, k - 1);
            z2m = cellv(f, g, i, j, k - 2);
        }
        const double ddz = eno2(hz, v[2], z2p, z1p, f0, z1m, z2m);
        ug[q] = v[0] * ddx + v[1] * ddy + v[2] * ddz;
    }
    const double s11 = sig[c], s12 = sig[g.Nc + c], s13 = sig[2 * g.Nc + c];
    const double s22 = sig[3 * g.Nc + c], s23 = sig[4 * g.Nc + c], s33 = sig[5 * g.Nc + c];
    const double du_dx = G[0][0], du_dy = G[0][1], du_dz = G[0][2];
    const double dv_dx = G[1][0], dv_dy = G[1][1], dv_dz = G[1][2];
    const double dw_dx = G[2][0], dw_dy = G[2][1], dw_dz = G[2][2];
    const double d_trace = (du_dx + dv_dy + dw_dz);
    const double t11 = s11 * (du_dx - dv_dy - dw_dz) + 2 * s13 * du_dz + 2 * s12 * du_dy;
    const double t12 = s23 * du_dz + s13 * dv_dz - s12 * dw_dz + s22 * du_dy + s11 * dv_dx;
    const double t13 = s33 * du_dz + s11 * dw_dx + s23 * du_dy + s12 * dw_dy - s13 * dv_dy;
    const double t22 = 2 * s23 * dv_dz + s22 * (dv_dy - du_dx - dw_dz) + 2 * s12 * dv_dx;
    const double t23 = s33 * dv_dz + s22 * dw_dy + s13 * dv_dx - s23 * du_dx + s12 * dw_dx;
    const double t33 = s33 * (dw_dz - dv_dy - du_dx) + 2 * s23 * dw_dy + 2 * s13 * dw_dx;
    double cs11 = dtlamb * d_trace + twomudt * du_dx - ug[0] + dt * t11;
    double cs22 = dtlamb * d_trace + twomudt * dv_dy - ug[3] + dt * t22;
    double cs33 = dtlamb * d_trace + twomudt * dw_dz - ug[5] + dt * t33;
    double cs12 = dtmu * (du_dy + dv_dx) - ug[1] + dt * t12;
    double cs13 = dtmu * (du_dz + dw_dx) - ug[2] + dt * t13;
    double cs23 = dtmu * (dv_dz + dw_dy) - ug[4] + dt * t23;
    const double a = adapt[c];
    if (a != 0.0) {
        const double sig_trace = s11 + s22 + s33, third = 1. / 3.;
        cs11 -= a * (s11 - third * sig_trace);
        cs12 -= a * (s12);
        cs13 -= a * (s13);
        cs22 -= a * (s22 - third * sig_trace);
        cs23 -= a * (s23);
        cs33 -= a * (s33 - third * sig_trace);
    }
    sig_out[c] = s11 + cs11;
    sig_out[g.Nc + c] = s12 + cs12;
    sig_out[2 * g.Nc + c] = s13 + cs13;
    sig_out[3 * g.Nc + c] = s22 + cs22;
    sig_out[4 * g.Nc + c] = s23 + cs23;
    sig_out[5 * g.Nc + c] = s33 + cs33;
    // chi diffusion (chi_diffusion, src/shear_sim_3d.cpp:1217-1237)
    const double Dc = dnew[c];
    const double Dxm = i > 0 ? dnew[c - 1] : dold[c + S.ox[1]];
    const double Dxp = dold[c + S.ox[3]];
    const double Dym = j > 0 ? dnew[c - g.Nx] : dold[c + S.oy[1]];
    const double Dyp = dold[c + S.oy[3]];
    double Dzm, Dzp;
    if (g.pz) {
        Dzm = k > 0 ? dnew[c - (long)g.Nx * g.Ny] : dold[c + S.oz[1]];
        Dzp = dold[c + S.oz[3]];
    } else {
        Dzm = k > 0 ? dnew[c - (long)g.Nx * g.Ny] : 0.0;
        Dzp = k < g.Nz - 1 ? dold[c + (long)g.Nx * g.Ny] : 0.0;
    }
    const double *xc = chi + c;
    const double x0 = xc[0];
    const double xxm = xc[S.ox[1]], xxp = xc[S.ox[3]];
    const double xym = xc[S.oy[1]], xyp = xc[S.oy[3]];
    const double xzm = zin ? xc[S.oz[1]] : cellv(chi, g, i, j, k - 1);
    const double xzp = zin ? xc[S.oz[3]] : cellv(chi, g, i, j, k + 1);
    const double ix = 1 / g.dx, iy = 1 / g.dy, iz = 1 / g.dz;
    const double ddx = .5 * ix * (Dxp - Dxm) * (.5 * ix * (xxp - xxm)) + Dc * (ix * ix * (xxp - 2 * x0 + xxm));
    const double ddy = .5 * iy * (Dyp - Dym) * (.5 * iy * (xyp - xym)) + Dc * (iy * iy * (xyp - 2 * x0 + xym));
    const double ddz = .5 * iz * (Dzp - Dzm) * (.5 * iz * (xzp - xzm)) + Dc * (iz * iz * (xzp - 2 * x0 + xzm));
    const double diff = M.diff_l * M.diff_l * (ddx + ddy + ddz);
    double cchi = dchi[c];
    cchi -= ug[6];
    cchi += dt * diff / M.c0;
    chi_out[c] = x0 + cchi;
}

static __global__ __launch_bounds__(kT) void k_direct_vertex(Grid g, Dyn D, double dt, const double *__restrict__ vel,
                                                             const double *__restrict__ sig,
                                                             const double *__restrict__ xi,
                                                             double *__restrict__ vel_out,
                                                             double *__restrict__ xi_out, BrickGrid B) {
    int i, j, k;
    if (!brick_point(B, g.Nx, g.Ny, i, j, k)) return;
    const long p = i + (long)g.Nx * (j + (long)g.Ny * k);
    if (k == 0 || k >= g.Nz) {
        // the plane zz = 0 and the top plate are not updated here
#pragma unroll
        for (int q = 0; q < 3; q++) {
            vel_out[q * g.Nv + p] = vel[q * g.Nv + p];
            xi_out[q * g.Nv + p] = xi[q * g.Nv + p];
        }
        return;
    }
    const double dx_inv = 1. / g.dx, dy_inv = 1. / g.dy, dz_inv = 1. / g.dz;
    const double dx_inv_sq = dx_inv * dx_inv, dy_inv_sq = dy_inv * dy_inv, dz_inv_sq = dz_inv * dz_inv;
    const double rho_inv = 1. / D.rho;
    const double dtrho_inv = dt * rho_inv, dtkaprho_inv = D.kappa * dtrho_inv;
    const double dxdt = g.dx / dt, dydt = g.dy / dt, dzdt = g.dz / dt;
    // backward staggered stress derivatives (calculate_first_order_stresses,
    // src/shear_sim_3d.cpp:971-987): cells (i-1..i, j-1..j, k-1..k)
    // (0 < k < Nz here, so the cell planes k - 1 and k exist and the cell
    // index of (i, j, k) is p)
    const Stencil5 S = stencil5(g, i, j, k, g.Nzv);
    const bool zin = g.pz || (k >= 2 && k <= g.Nzv - 3);  // no wall ghost in the z stencils
    const int mx = S.ox[1], my = S.oy[1], mz = S.oz[1];
    const long c000 = p, c100 = p + mx, c010 = p + my, c110 = p + mx + my, c001 = p + mz, c101 = p + mx + mz,
               c011 = p + my + mz, c111 = p + mx + my + mz;
    auto DX = [&](const double *a) {
        return .25 * dx_inv * (a[c000] - a[c100] + a[c010] - a[c110] + a[c001] - a[c101] + a[c011] - a[c111]);
    };
    auto DY = [&](const double *a) {
        return .25 * dy_inv * (a[c000] - a[c010] + a[c100] - a[c110] + a[c001] - a[c011] + a[c101] - a[c111]);
    };
    auto DZ = [&](const double *a) {
        return .25 * dz_inv * (a[c000] - a[c001] + a[c100] - a[c101] + a[c010] - a[c011] + a[c110] - a[c111]);
    };
    const double *S11 = sig, *S12 = sig + g.Nc, *S13 = sig + 2 * g.Nc, *S22 = sig + 3 * g.Nc, *S23 = sig + 4 * g.Nc,
                 *S33 = sig + 5 * g.Nc;
    const double div[3] = {DX(S11) + DY(S12) + DZ(S13), DX(S12) + DY(S22) + DZ(S23), DX(S13) + DY(S23) + DZ(S33)};
    const double u = vel[p], v = vel[g.Nv + p], w = vel[2 * g.Nv + p];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const double *f = vel + q * g.Nv;
        const double *fp = f + p;
        const double f0 = fp[0];
        double z2p, z1p, z1m, z2m;
        if (zin) {
            z2p = fp[S.oz[4]];
            z1p = fp[S.oz[3]];
            z1m = fp[S.oz[1]];
            z2m = fp[S.oz[0]];
        } else {
            z2p = vel_g(f, g, i, j, k + 2);
            z1p = vel_g(f, g, i, j, k + 1);
            z1m = vel_g(f, g, i, j, k - 1);
            z2m = vel_g(f, g, i, j, k - 2);
        }
        const double gsq = dx_inv_sq * (fp[S.ox[3]] - 2 * f0 + fp[S.ox[1]]) +
                           dy_inv_sq * (fp[S.oy[3]] - 2 * f0 + fp[S.oy[1]]) + dz_inv_sq * (z1p - 2 * f0 + z1m);
        const double ddx = eno2(dxdt, u, fp[S.ox[4]], fp[S.ox[3]], f0, fp[S.ox[1]], fp[S.ox[0]]);
        const double ddy = eno2(dydt, v, fp[S.oy[4]], fp[S.oy[3]], f0, fp[S.oy[1]], fp[S.oy[0]]);
        const double ddz = eno2(dzdt, w, z2p, z1p, f0, z1m, z2m);
        const double ugq = u * ddx + v * ddy + w * ddz;
        vel_out[q * g.Nv + p] = f0 + (dtrho_inv * div[q] + dtkaprho_inv * gsq - ugq);
        // reference map: the periodic image of component q along axis q is
        // shifted by the domain length, as xi_g does it
        const double *xp = xi + q * g.Nv + p;
        const double x0 = xp[0];
        auto img = [&](const int(&o)[5], int n, int a, double Lq, bool axis) {
            const double val = xp[o[a]];
            if (!axis || o[a] == (a - 2) * n) return val;
            const double off = a > 2 ? Lq : -Lq;
            return off == 0.0 ? val : val + off;
        };
        const int nxy = g.Nx * g.Ny;
        const double ex = eno2(dxdt, u, img(S.ox, 1, 4, D.Lx, q == 0), img(S.ox, 1, 3, D.Lx, q == 0), x0,
                               img(S.ox, 1, 1, D.Lx, q == 0), img(S.ox, 1, 0, D.Lx, q == 0));
        const double ey = eno2(dydt, v, img(S.oy, g.Nx, 4, D.Ly, q == 1), img(S.oy, g.Nx, 3, D.Ly, q == 1), x0,
                               img(S.oy, g.Nx, 1, D.Ly, q == 1), img(S.oy, g.Nx, 0, D.Ly, q == 1));
        double ez;
        if (zin)
            ez = eno2(dzdt, w, img(S.oz, nxy, 4, D.Lz, q == 2), img(S.oz, nxy, 3, D.Lz, q == 2), x0,
                      img(S.oz, nxy, 1, D.Lz, q == 2), img(S.oz, nxy, 0, D.Lz, q == 2));
        else
            ez = eno2(dzdt, w, xi_g(xi, g, D, q, i, j, k + 2), xi_g(xi, g, D, q, i, j, k + 1), x0,
                      xi_g(xi, g, D, q, i, j, k - 1), xi_g(xi, g, D, q, i, j, k - 2));
        xi_out[q * g.Nv + p] = x0 + (0.0 - (u * ex + v * ey + w * ez));
    }
}

// Top plate at time t: velocity (U t, 0, 0) (U for t >= 1) and reference map
// (a_x + i dx + U t^2 / 2 or + U (t - 1/2), a_y + j dy, 0).
static __global__ __launch_bounds__(kT) void k_direct_plate(Grid g, Dyn D, double t, double *__restrict__ vel,
                                                            double *__restrict__ xi) {
    const long q = xcd_block() * kT + threadIdx.x;
    const long nxy = (long)g.Nx * g.Ny;
    if (g.pz || q >= nxy) return;
    const int i = (int)(q % g.Nx), j = (int)(q / g.Nx);
    const long p = q + nxy * g.Nz;
    vel[p] = t < 1 ? D.U * t : D.U;
    vel[g.Nv + p] = 0;
    vel[2 * g.Nv + p] = 0;
    xi[p] = D.ax + i * g.dx + (t < 1 ? D.U * t * t * .5 : D.U * (t - 0.5));
    xi[g.Nv + p] = D.ay + j * g.dy;
    xi[2 * g.Nv + p] = 0;
}

// The reference map at the grid positions (init_ref, src/shear_sim_3d.cpp:
// 1036-1057) below the top plate.
static __global__ __launch_bounds__(kT) void k_init_ref(Grid g, Dyn D, double *__restrict__ xi) {
    const long p = xcd_block() * kT + threadIdx.x;
    if (p >= g.Nv) return;
    const int i = (int)(p % g.Nx);
    const int j = (int)((p / g.Nx) % g.Ny);
    const int k = (int)(p / ((long)g.Nx * g.Ny));
    if (k >= g.Nz) return;
    xi[p] = D.ax + g.dx * i;
    xi[g.Nv + p] = D.ay + g.dy * j;
    xi[2 * g.Nv + p] = D.az + g.dz * k;
}

// Block partials of l2_comparison (src/shear_sim_3d.cpp:3122-3158) over the
// points p < Nx Ny (Nz + 1): vertex velocity with the trapezoid weight in z
// (1/2 on plane 0 and on plane Nz, the top plate or the periodic image of
// plane 0), and for p < Nc the cell's six stress and chi differences.
// part[q nblocks + block], q = velocity, stress, chi (unscaled sums).
static __global__ __launch_bounds__(kT) void k_l2_parts(Grid g, const double *__restrict__ ua,
                                                        const double *__restrict__ sa, const double *__restrict__ xa,
                                                        const double *__restrict__ ub, const double *__restrict__ sb,
                                                        const double *__restrict__ xb, double *__restrict__ part) {
    const long p = xcd_block() * kT + threadIdx.x;
    const long nxy = (long)g.Nx * g.Ny;
    double v[3] = {0, 0, 0};
    if (p < nxy * (g.Nz + 1)) {
        const int k = (int)(p / nxy);
        const long pv = (g.pz && k == g.Nz) ? p - nxy * g.Nz : p;
        const double du = ua[pv] - ub[pv], dv = ua[g.Nv + pv] - ub[g.Nv + pv];
        const double dw = ua[2 * g.Nv + pv] - ub[2 * g.Nv + pv];
        if (k < g.Nz) {
            v[0] = (du * du + dv * dv + dw * dw) * (k == 0 ? .5 : 1);
            double t = 0;
#pragma unroll
            for (int q = 0; q < 6; q++) {
                const double d = sa[q * g.Nc + p] - sb[q * g.Nc + p];
                t += d * d;
            }
            v[1] = t;
            const double dc = xa[p] - xb[p];
            v[2] = dc * dc;
        } else {
            v[0] = .5 * (du * du + dv * dv + dw * dw);
        }
    }
    __shared__ double red[3][kT / 64];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        double t = v[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
        if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = t;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0;
        for (int w = 0; w < kT / 64; w++) t += red[threadIdx.x][w];
        part[(long)threadIdx.x * gridDim.x + blockIdx.x] = t;
    }
}

#define DHIP(s, call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return stz_detail::fail((s), STZMG_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

Dyn dyn_of(const stz_t s) {
    const stz_params &p = s->p;
    return Dyn{p.rho, p.kappa, p.U, p.ax, p.ay, p.az, p.bx - p.ax, p.by - p.ay, p.bz - p.az};
}

int ensure_ref_map(stz_t s) {
    if (s->xi) return 0;
    const Grid &g = s->g;
    DHIP(s, hipMalloc((void **)&s->xi, sizeof(double) * 3 * g.Nv));
    DHIP(s, hipMalloc((void **)&s->xi2, sizeof(double) * 3 * g.Nv));
    DHIP(s, hipMalloc((void **)&s->vel2, sizeof(double) * 3 * g.Nv));
    DHIP(s, hipMemsetAsync(s->xi, 0, sizeof(double) * 3 * g.Nv, s->stream));
    return 0;
}

int check_direct(stz_t s) {
    if (s->p.transform != STZ_T_NONE)
        return stz_detail::fail(s, STZMG_ERR_UNSUPPORTED, "the direct step is the untransformed shear_sim_3d one");
    if (!(s->p.rho > 0) || !(s->p.kappa >= 0))
        return stz_detail::fail(s, STZMG_ERR_ARG, "direct step needs rho > 0 and kappa >= 0");
    return 0;
}

}  // namespace

extern "C" {

double stz_direct_dt(stz_t s) {
    if (!s) return 0;
    // solve_direct (src/shear_sim_3d.cpp:166-172)
    const double dx = s->g.dx, kappa = s->p.kappa, tmult = s->p.tmult;
    const double cfl = dx / 4.;
    if (!(kappa > 0)) return cfl;
    const double visc = dx * dx / 3. / kappa / 2.;
    return (tmult * visc < cfl) ? tmult * visc : cfl;
}

int stz_init_ref(stz_t s) {
    if (!s) return STZMG_ERR_ARG;
    if (int rc = ensure_ref_map(s)) return rc;
    const Grid &g = s->g;
    const Dyn D = dyn_of(s);
    hipLaunchKernelGGL(k_init_ref, dim3(nbk(g.Nv)), dim3(kT), 0, s->stream, g, D, s->xi);
    hipLaunchKernelGGL(k_direct_plate, dim3(nbk((long)g.Nx * g.Ny)), dim3(kT), 0, s->stream, g, D, s->t, s->vel,
                       s->xi);
    DHIP(s, hipGetLastError());
    DHIP(s, hipStreamSynchronize(s->stream));
    return 0;
}

int stz_set_ref_map(stz_t s, const double *xi) {
    if (!s || !xi) return STZMG_ERR_ARG;
    if (int rc = ensure_ref_map(s)) return rc;
    const Grid &g = s->g;
    DHIP(s, hipMemcpyAsync(s->stage, xi, sizeof(double) * 3 * g.Nv, hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(k_aos2soa, dim3(nbk(g.Nv)), dim3(kT), 0, s->stream, (const double *)s->stage, s->xi, g.Nv, 3);
    DHIP(s, hipGetLastError());
    DHIP(s, hipStreamSynchronize(s->stream));
    return 0;
}

int stz_get_ref_map(stz_t s, double *xi) {
    if (!s || !xi) return STZMG_ERR_ARG;
    if (!s->xi) return stz_detail::fail(s, STZMG_ERR_STATE, "no reference map (stz_init_ref or stz_set_ref_map)");
    const Grid &g = s->g;
    hipLaunchKernelGGL(k_soa2aos, dim3(nbk(g.Nv)), dim3(kT), 0, s->stream, (const double *)s->xi, s->stage, g.Nv, 3);
    DHIP(s, hipGetLastError());
    DHIP(s, hipMemcpyAsync(xi, s->stage, sizeof(double) * 3 * g.Nv, hipMemcpyDeviceToHost, s->stream));
    DHIP(s, hipStreamSynchronize(s->stream));
    return 0;
}

int stz_step_direct(stz_t s, double dt, int nsteps, stz_step_stats *st) {
    if (!s || !(dt > 0) || nsteps < 0) return STZMG_ERR_ARG;
    if (int rc = check_direct(s)) return rc;
    if (int rc = ensure_ref_map(s)) return rc;
    const Grid &g = s->g;
    const Dyn D = dyn_of(s);
    const BrickGrid bc = brick_grid(g.Nx, g.Ny, g.Nz), bv = brick_grid(g.Nx, g.Ny, g.Nzv);
    stz_step_stats S{};
    for (int q = 0; q < nsteps; q++) {
        DHIP(s, hipMemsetAsync(s->counters, 0, sizeof(int) * 4, s->stream));
        hipLaunchKernelGGL(k_direct_plastic, dim3(nbk(g.Nc)), dim3(kT), 0, s->stream, g, s->M, dt,
                           (const double *)s->sig, (const double *)s->chi, s->adapt, s->dchi, s->dpl2, s->counters);
        hipLaunchKernelGGL(k_direct_cell, dim3(brick_blocks(bc)), dim3(kT), 0, s->stream, g, s->M, dt,
                           (const double *)s->vel, (const double *)s->sig, (const double *)s->chi,
                           (const double *)s->adapt, (const double *)s->dchi, (const double *)s->dpl2,
                           (const double *)s->dpl, s->sig2, s->chi2, bc);
        hipLaunchKernelGGL(k_direct_vertex, dim3(brick_blocks(bv)), dim3(kT), 0, s->stream, g, D, dt,
                           (const double *)s->vel, (const double *)s->sig, (const double *)s->xi, s->vel2, s->xi2,
                           bv);
        DHIP(s, hipGetLastError());
        std::swap(s->sig, s->sig2);
        std::swap(s->chi, s->chi2);
        std::swap(s->dpl, s->dpl2);
        std::swap(s->vel, s->vel2);
        std::swap(s->xi, s->xi2);
        s->t += dt;
        hipLaunchKernelGGL(k_direct_plate, dim3(nbk((long)g.Nx * g.Ny)), dim3(kT), 0, s->stream, g, D, s->t, s->vel,
                           s->xi);
        DHIP(s, hipGetLastError());
        S.steps++;
    }
    // one synchronisation per call: the counters hold the last step's count
    int cnt[4] = {0, 0, 0, 0};
    DHIP(s, hipMemcpyAsync(cnt, s->counters, sizeof(cnt), hipMemcpyDeviceToHost, s->stream));
    DHIP(s, hipStreamSynchronize(s->stream));
    S.yielded_cells = nsteps ? cnt[0] : 0;
    if (st) *st = S;
    return 0;
}

int stz_l2_comparison(stz_t a, stz_t b, double l2[3]) {
    if (!a || !b || !l2) return STZMG_ERR_ARG;
    const Grid &g = a->g;
    if (b->g.Nx != g.Nx || b->g.Ny != g.Ny || b->g.Nz != g.Nz || b->g.pz != g.pz)
        return stz_detail::fail(a, STZMG_ERR_ARG, "l2_comparison: the two simulations have different grids");
    // b's fields must be complete before a's stream reads them
    DHIP(a, hipStreamSynchronize(b->stream));
    const long np = (long)g.Nx * g.Ny * (g.Nz + 1);
    const unsigned nb = nbk(np);
    double *part = a->stage;  // 3 nb partials + 3 sums: within 6 max(Nc, Nv) doubles
    double *sums = part + 3 * (long)nb;
    hipLaunchKernelGGL(k_l2_parts, dim3(nb), dim3(kT), 0, a->stream, g, (const double *)a->vel,
                       (const double *)a->sig, (const double *)a->chi, (const double *)b->vel,
                       (const double *)b->sig, (const double *)b->chi, part);
    for (int q = 0; q < 3; q++)
        hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kT), 0, a->stream, (const double *)(part + (long)q * nb),
                           (long)nb, sums + q);
    DHIP(a, hipGetLastError());
    double h[3];
    DHIP(a, hipMemcpyAsync(h, sums, sizeof(h), hipMemcpyDeviceToHost, a->stream));
    DHIP(a, hipStreamSynchronize(a->stream));
    // normalisation (:3152-3155) with a's grid, plate speed and chi_inf
    const stz_params &p = a->p;
    const double norm_val = g.dx * g.dy * g.dz / ((p.bx - p.ax) * (p.by - p.ay) * (p.bz - p.az));
    l2[0] = h[0] * norm_val / (p.U * p.U);
    l2[1] = h[1] * norm_val;
    l2[2] = h[2] * norm_val / (p.chi_inf * p.chi_inf);
    return 0;
}

int stz_direct_time_kernels(stz_t s, double dt, int reps, double ms[4]) {
    if (!s || !(dt > 0) || reps < 1 || !ms) return STZMG_ERR_ARG;
    if (int rc = check_direct(s)) return rc;
    if (int rc = ensure_ref_map(s)) return rc;
    const Grid &g = s->g;
    const Dyn D = dyn_of(s);
    const BrickGrid bc = brick_grid(g.Nx, g.Ny, g.Nz), bv = brick_grid(g.Nx, g.Ny, g.Nzv);
    hipEvent_t e0, e1;
    DHIP(s, hipEventCreate(&e0));
    DHIP(s, hipEventCreate(&e1));
    // Every kernel reads the current state and writes the second buffers (the
    // plate kernel rewrites the plate's values at the current time), so the
    // repeated launches leave the state as it was.
    auto launch = [&](int which) {
        if (which == 0)
            hipLaunchKernelGGL(k_direct_plastic, dim3(nbk(g.Nc)), dim3(kT), 0, s->stream, g, s->M, dt,
                               (const double *)s->sig, (const double *)s->chi, s->adapt, s->dchi, s->dpl2,
                               s->counters);
        else if (which == 1)
            hipLaunchKernelGGL(k_direct_cell, dim3(brick_blocks(bc)), dim3(kT), 0, s->stream, g, s->M, dt,
                               (const double *)s->vel, (const double *)s->sig, (const double *)s->chi,
                               (const double *)s->adapt, (const double *)s->dchi, (const double *)s->dpl2,
                               (const double *)s->dpl, s->sig2, s->chi2, bc);
        else if (which == 2)
            hipLaunchKernelGGL(k_direct_vertex, dim3(brick_blocks(bv)), dim3(kT), 0, s->stream, g, D, dt,
                               (const double *)s->vel, (const double *)s->sig, (const double *)s->xi, s->vel2,
                               s->xi2, bv);
        else
            hipLaunchKernelGGL(k_direct_plate, dim3(nbk((long)g.Nx * g.Ny)), dim3(kT), 0, s->stream, g, D, s->t,
                               s->vel, s->xi);
    };
    int rc = 0;
    for (int w = 0; w < 4 && !rc; w++) {
        launch(w);  // warm
        hipEventRecord(e0, s->stream);
        for (int r = 0; r < reps; r++) launch(w);
        hipEventRecord(e1, s->stream);
        hipError_t e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipGetLastError();
        float t = 0;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        if (e != hipSuccess) rc = stz_detail::fail(s, STZMG_ERR_HIP, "direct kernel timing: %s", hipGetErrorString(e));
        ms[w] = (double)t / reps;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return rc;
}

}  // extern "C"
