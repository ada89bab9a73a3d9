// Reverse-mode (adjoint) kernels of the quasi-static step of shear_sim_3d
// (the forward kernels are in stz_kernels.hpp).
//
// The forward step, per cell c and vertex p:
//   plastic   (a, dchi) = adaptive(sbar(sigma_c), chi_c; s_y, eps0, chi_inf),
//             Dn_c = a / (2 mu) sbar(sigma_c)
//   advect    sigma*_c = sigma_c + P(a, sigma_c) - u.grad(sigma)_c + dt Tr(G_c, sigma_c)
//             chi*_c   = chi_c + dchi_c - u.grad(chi)_c + diff(Dn, Do, chi)_c / c0
//   rhs       r_p = div sigma* (interior rows; wall rows are constants)
//   solve     u'_I = A_II^-1 (r_I - A_IW u_W), wall rows fixed
//   stress    sigma'_c = sigma*_c + dt C:D(u')_c
// is reversed one stage at a time.  The ENO2 and chi-diffusion stencils are
// piecewise linear in the fields for the branch the data select, and are
// transposed with that branch.  Scatters go through global atomics (FP64
// vector atomics), so sums are formed in a run-dependent order.  The
// plastic update is differentiated in forward mode (five directions:
// sbar_0, chi_0, s_y, eps0, chi_inf) through the adaptive sub-stepping loop.
#pragma once

#include "stz_kernels.hpp"

namespace stz {

// ---- forward-mode numbers for the plastic update ---------------------------
struct D5 {
    double v, d[5];
};
__device__ __forceinline__ D5 dcst(double v) {
    D5 r;
    r.v = v;
#pragma unroll
    for (int q = 0; q < 5; q++) r.d[q] = 0;
    return r;
}
__device__ __forceinline__ D5 dvar(double v, int k) {
    D5 r = dcst(v);
    r.d[k] = 1;
    return r;
}
__device__ __forceinline__ D5 operator+(const D5 &a, const D5 &b) {
    D5 r;
    r.v = a.v + b.v;
#pragma unroll
    for (int q = 0; q < 5; q++) r.d[q] = a.d[q] + b.d[q];
    return r;
}
__device__ __forceinline__ D5 operator-(const D5 &a, const D5 &b) {
    D5 r;
    r.v = a.v - b.v;
#pragma unroll
    for (int q = 0; q < 5; q++) r.d[q] = a.d[q] - b.d[q];
    return r;
}
__device__ __forceinline__ D5 operator*(const D5 &a, const D5 &b) {
    D5 r;
    r.v = a.v * b.v;
#pragma unroll
    for (int q = 0; q < 5; q++) r.d[q] = a.d[q] * b.v + a.v * b.d[q];
    return r;
}
__device__ __forceinline__ D5 operator/(const D5 &a, const D5 &b) {
    D5 r;
    r.v = a.v / b.v;
#pragma unroll
    for (int q = 0; q < 5; q++) r.d[q] = (a.d[q] - r.v * b.d[q]) / b.v;
    return r;
}
__device__ __forceinline__ D5 operator*(double a, const D5 &b) {
    D5 r;
    r.v = a * b.v;
#pragma unroll
    for (int q = 0; q < 5; q++) r.d[q] = a * b.d[q];
    return r;
}
__device__ __forceinline__ D5 operator*(const D5 &a, double b) { return b * a; }
__device__ __forceinline__ D5 operator/(const D5 &a, double b) {
    D5 r;
    r.v = a.v / b;
#pragma unroll
    for (int q = 0; q < 5; q++) r.d[q] = a.d[q] / b;
    return r;
}
__device__ __forceinline__ D5 operator/(double a, const D5 &b) { return dcst(a) / b; }
__device__ __forceinline__ D5 operator-(double a, const D5 &b) { return dcst(a) - b; }
__device__ __forceinline__ D5 operator-(const D5 &a, double b) { return a - dcst(b); }
__device__ __forceinline__ D5 operator+(const D5 &a, double b) { return a + dcst(b); }
__device__ __forceinline__ D5 dexp(const D5 &a) {
    D5 r;
    r.v = exp(a.v);
#pragma unroll
    for (int q = 0; q < 5; q++) r.d[q] = r.v * a.d[q];
    return r;
}
__device__ __forceinline__ D5 dabs(const D5 &a) { return a.v < 0 ? (-1.0) * a : a; }

// adaptive_plastic_term (src/shear_sim_3d.cpp:109-149) with its derivatives:
// a = 1 - sbar_end/sbar_0 and dchi = chi_end - chi_0, each with d/d(sbar_0,
// chi_0, s_y, eps0, chi_inf) in Ja / Jc.  The value parts repeat the forward
// arithmetic of adaptive() operation for operation.
__device__ __forceinline__ void adaptive_jac(const Mat &M, double sb0_, double chi0_, double dt, double &a,
                                             double &dchi, double Ja[5], double Jc[5]) {
    const D5 sy = dvar(M.sy, 2), eps0 = dvar(M.eps0, 3), chinf = dvar(M.chi_inf, 4);
    const D5 sb0 = dvar(sb0_, 0), chi0 = dvar(chi0_, 1);
    D5 sb = sb0, chi = chi0, tr = dcst(dt), ts;
    bool keep = true;
    while (keep) {
        D5 dpl = dcst(0.0);
        if (sb.v >= sy.v) {
            const D5 ea = dexp((M.omega * eps0 * sb - M.delta) / M.bath_temp);
            const D5 eb = dexp((-1.0) * (M.omega * eps0 * sb + M.delta) / M.bath_temp);
            const D5 cc = .5 * (ea + eb);
            dpl = dexp((-M.ez) / chi) * cc * (1.0 - sy / sb) / M.tau0;
        }
        const D5 dprime = 2 * M.mu * dpl / sy;
        const D5 F = 2 * dpl * sb * (chinf - chi) / (sy * M.c0);
        if (dprime.v * tr.v > M.eta) {
            ts = M.eta / dabs(dprime);
            tr = tr - ts;
        } else {
            ts = tr;
            keep = false;
        }
        sb = sb - ts * dprime * sy;
        chi = chi + ts * F;
    }
    const D5 A = 1.0 - sb / sb0, C = chi - chi0;
    a = A.v;
    dchi = C.v;
#pragma unroll
    for (int q = 0; q < 5; q++) {
        Ja[q] = A.d[q];
        Jc[q] = C.d[q];
    }
}

// ---- transposes of the ghost rules ------------------------------------------

// Adds val to the cell value read by cellv(f, g, i, j, k): periodic wrap, or
// across a wall the transpose of the linear extrapolation (f(-1) = 2 f0 - f1,
// f(-2) = 3 f0 - 2 f1, likewise at the top).
__device__ __forceinline__ void cell_add(double *f, const Grid &g, int i, int j, int k, double val) {
    if (g.pz) {
        atomicAdd(&f[cidx(g, i, j, wrapc(k, g.Nz))], val);
    } else if (k >= 0 && k < g.Nz) {
        atomicAdd(&f[cidx(g, i, j, k)], val);
    } else if (k < 0) {
        const double w0 = k == -1 ? 2.0 : 3.0, w1 = k == -1 ? -1.0 : -2.0;
        atomicAdd(&f[cidx(g, i, j, 0)], w0 * val);
        atomicAdd(&f[cidx(g, i, j, 1)], w1 * val);
    } else {
        const double w0 = k == g.Nz ? 2.0 : 3.0, w1 = k == g.Nz ? -1.0 : -2.0;
        atomicAdd(&f[cidx(g, i, j, g.Nz - 1)], w0 * val);
        atomicAdd(&f[cidx(g, i, j, g.Nz - 2)], w1 * val);
    }
}

__device__ __forceinline__ long vidx(const Grid &g, int i, int j, int k) {
    if (g.pz) k = wrapc(k, g.Nz);
    return wrapc(i, g.Nx) + (long)g.Nx * (wrapc(j, g.Ny) + (long)g.Ny * k);
}

// Transpose of vel_grad: G-bar (3x3) and the cell-velocity bar onto the 8
// vertices of cell (i,j,k).
__device__ __forceinline__ void grad_add(double *ub, const Grid &g, int i, int j, int k, const double Gb[3][3],
                                         const double vb[3]) {
    const double ix = .25 * (1.0 / g.dx), iy = .25 * (1.0 / g.dy), iz = .25 * (1.0 / g.dz);
#pragma unroll
    for (int n = 0; n < 8; n++) {
        const int di = n & 1, dj = (n >> 1) & 1, dk = n >> 2;
        const long p = vidx(g, i + di, j + dj, k + dk);
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const double val = ix * (di ? Gb[q][0] : -Gb[q][0]) + iy * (dj ? Gb[q][1] : -Gb[q][1]) +
                               iz * (dk ? Gb[q][2] : -Gb[q][2]) + 0.125 * vb[q];
            atomicAdd(&ub[q * g.Nv + p], val);
        }
    }
}

// Weights of eno2 on (f(+2), f(+1), f, f(-1), f(-2)) for the branch the data
// select (the forward picks by the sign of v and two curvature tests).
__device__ __forceinline__ void eno2_w(double h, double v, double f2p, double fp, double f, double fm, double f2m,
                                       double w[5]) {
    const double c2 = fabs(fp - 2 * f + fm);
    const double fw = fabs(f2p - 2 * fp + f);
    const double bw = fabs(f - 2 * fm + f2m);
    const double s = .5 / h;
#pragma unroll
    for (int q = 0; q < 5; q++) w[q] = 0;
    if (v < 0) {
        if (c2 > fw) { w[0] = -s; w[1] = 4 * s; w[2] = -3 * s; }
        else { w[1] = s; w[3] = -s; }
    } else {
        if (c2 > bw) { w[2] = 3 * s; w[3] = -4 * s; w[4] = s; }
        else { w[1] = s; w[3] = -s; }
    }
}

// ---- stages, last to first ---------------------------------------------------

// sigma' = sigma* + dt C:D(u'): u'-bar += (dt C:D)^T sigma'-bar (cell scatter).
static __global__ __launch_bounds__(kT) void k_adj_stress(Grid g, Mat M, double dt, const double *__restrict__ sb,
                                                   double *__restrict__ ub) {
    const long c = (long)blockIdx.x * kT + threadIdx.x;
    if (c >= g.Nc) return;
    const int i = (int)(c % g.Nx), j = (int)((c / g.Nx) % g.Ny), k = (int)(c / ((long)g.Nx * g.Ny));
    const double b11 = sb[c], b12 = sb[g.Nc + c], b13 = sb[2 * g.Nc + c];
    const double b22 = sb[3 * g.Nc + c], b23 = sb[4 * g.Nc + c], b33 = sb[5 * g.Nc + c];
    const double lam = M.lamb, mu = M.mu, L2 = lam + 2 * mu;
    double Gb[3][3];
    Gb[0][0] = dt * (L2 * b11 + lam * b22 + lam * b33);
    Gb[1][1] = dt * (lam * b11 + L2 * b22 + lam * b33);
    Gb[2][2] = dt * (lam * b11 + lam * b22 + L2 * b33);
    Gb[0][1] = Gb[1][0] = dt * mu * b12;
    Gb[0][2] = Gb[2][0] = dt * mu * b13;
    Gb[1][2] = Gb[2][1] = dt * mu * b23;
    const double vb[3] = {0, 0, 0};
    grad_add(ub, g, i, j, k, Gb, vb);
}

// Right-hand side of the adjoint projection: u'-bar with the wall rows zeroed
// (the walls are fixed values, not unknowns).
static __global__ __launch_bounds__(kT) void k_adj_wall_zero(Grid g, const double *__restrict__ ub,
                                                      double *__restrict__ rhs) {
    const long p = (long)blockIdx.x * kT + threadIdx.x;
    if (p >= g.Nv) return;
    const int k = (int)(p / ((long)g.Nx * g.Ny));
    const bool wall = !g.pz && (k == 0 || k == g.Nz);
#pragma unroll
    for (int q = 0; q < 3; q++) rhs[q * g.Nv + p] = wall ? 0.0 : ub[q * g.Nv + p];
}

// r = div sigma* at interior vertices: sigma*-bar = sigma'-bar + div^T lambda
// (gather over the 8 vertices of each cell; lambda = A^-T u'-bar).
static __global__ __launch_bounds__(kT) void k_adj_div(Grid g, const double *__restrict__ lam,
                                                const double *__restrict__ sb, double *__restrict__ ssb) {
    const long c = (long)blockIdx.x * kT + threadIdx.x;
    if (c >= g.Nc) return;
    const int i = (int)(c % g.Nx), j = (int)((c / g.Nx) % g.Ny), k = (int)(c / ((long)g.Nx * g.Ny));
    const double ix = .25 * (1.0 / g.dx), iy = .25 * (1.0 / g.dy), iz = .25 * (1.0 / g.dz);
    double a11 = 0, a12 = 0, a13 = 0, a22 = 0, a23 = 0, a33 = 0;
#pragma unroll
    for (int n = 0; n < 8; n++) {
        const int di = n & 1, dj = (n >> 1) & 1, dk = n >> 2;
        const int kv = k + dk;
        if (!g.pz && (kv == 0 || kv == g.Nz)) continue;
        const long p = vidx(g, i + di, j + dj, kv);
        const double l0 = lam[p], l1 = lam[g.Nv + p], l2 = lam[2 * g.Nv + p];
        const double sx = di ? -ix : ix, sy = dj ? -iy : iy, sz = dk ? -iz : iz;
        a11 += sx * l0;
        a12 += sy * l0 + sx * l1;
        a13 += sz * l0 + sx * l2;
        a22 += sy * l1;
        a23 += sz * l1 + sy * l2;
        a33 += sz * l2;
    }
    ssb[c] = sb[c] + a11;
    ssb[g.Nc + c] = sb[g.Nc + c] + a12;
    ssb[2 * g.Nc + c] = sb[2 * g.Nc + c] + a13;
    ssb[3 * g.Nc + c] = sb[3 * g.Nc + c] + a22;
    ssb[4 * g.Nc + c] = sb[4 * g.Nc + c] + a23;
    ssb[5 * g.Nc + c] = sb[5 * g.Nc + c] + a33;
}

// Transpose of k_advect.  In: the pre-step state (vel, sig, chi, the old
// ad_Dpl dold), this step's plastic quantities (adapt, dnew) and the adjoint
// of sigma* (ssb) and chi* (xsb).  Out (accumulated, the identity parts
// already in sb_out / xb_out): sigma-bar, chi-bar, u-bar, dnew-bar, dold-bar;
// written: adapt-bar and dchi-bar of the cell.
static __global__ __launch_bounds__(kT) void k_adj_advect(Grid g, Mat M, double dt, const double *__restrict__ vel,
                                                   const double *__restrict__ sig, const double *__restrict__ chi,
                                                   const double *__restrict__ adapt, const double *__restrict__ dnew,
                                                   const double *__restrict__ dold, const double *__restrict__ ssb,
                                                   const double *__restrict__ xsb, double *__restrict__ sb_out,
                                                   double *__restrict__ xb_out, double *__restrict__ ub_out,
                                                   double *__restrict__ dnb, double *__restrict__ dob,
                                                   double *__restrict__ ab, double *__restrict__ dchib) {
    const long c = (long)blockIdx.x * kT + threadIdx.x;
    if (c >= g.Nc) return;
    const int i = (int)(c % g.Nx), j = (int)((c / g.Nx) % g.Ny), k = (int)(c / ((long)g.Nx * g.Ny));
    double G[3][3], v[3];
    vel_grad(vel, g, i, j, k, G, v);
    const double h[3] = {g.dx / dt, g.dy / dt, g.dz / dt};
    double gbar[7];
#pragma unroll
    for (int q = 0; q < 6; q++) gbar[q] = -ssb[q * g.Nc + c];
    gbar[6] = -xsb[c];
    // u.grad f, f = sigma components and chi
    double vb[3] = {0, 0, 0};
    for (int q = 0; q < 7; q++) {
        const double *f = q < 6 ? sig + q * g.Nc : chi;
        double *fb = q < 6 ? sb_out + q * g.Nc : xb_out;
        const double gq = gbar[q];
        if (gq == 0.0) continue;
#pragma unroll
        for (int ax = 0; ax < 3; ax++) {
            const int ex = ax == 0, ey = ax == 1, ez = ax == 2;
            double fv[5];
#pragma unroll
            for (int o = 0; o < 5; o++) {
                const int s = 2 - o;  // offsets +2 +1 0 -1 -2
                fv[o] = cellv(f, g, i + s * ex, j + s * ey, k + s * ez);
            }
            double w[5];
            eno2_w(h[ax], v[ax], fv[0], fv[1], fv[2], fv[3], fv[4], w);
            double E = 0;
#pragma unroll
            for (int o = 0; o < 5; o++) E += w[o] * fv[o];
            vb[ax] += gq * E;
#pragma unroll
            for (int o = 0; o < 5; o++)
                if (w[o] != 0.0) {
                    const int s = 2 - o;
                    cell_add(fb, g, i + s * ex, j + s * ey, k + s * ez, gq * v[ax] * w[o]);
                }
        }
    }
    // Truesdell terms: sigma* += dt t(G, sigma)
    const double s11 = sig[c], s12 = sig[g.Nc + c], s13 = sig[2 * g.Nc + c];
    const double s22 = sig[3 * g.Nc + c], s23 = sig[4 * g.Nc + c], s33 = sig[5 * g.Nc + c];
    const double du_dx = G[0][0], du_dy = G[0][1], du_dz = G[0][2];
    const double dv_dx = G[1][0], dv_dy = G[1][1], dv_dz = G[1][2];
    const double dw_dx = G[2][0], dw_dy = G[2][1], dw_dz = G[2][2];
    const double t11 = dt * ssb[c], t12 = dt * ssb[g.Nc + c], t13 = dt * ssb[2 * g.Nc + c];
    const double t22 = dt * ssb[3 * g.Nc + c], t23 = dt * ssb[4 * g.Nc + c], t33 = dt * ssb[5 * g.Nc + c];
    double sbar[6];
    sbar[0] = t11 * (du_dx - dv_dy - dw_dz) + t12 * dv_dx + t13 * dw_dx;
    sbar[1] = t11 * 2 * du_dy - t12 * dw_dz + t13 * dw_dy + t22 * 2 * dv_dx + t23 * dw_dx;
    sbar[2] = t11 * 2 * du_dz + t12 * dv_dz - t13 * dv_dy + t23 * dv_dx + t33 * 2 * dw_dx;
    sbar[3] = t12 * du_dy + t22 * (dv_dy - du_dx - dw_dz) + t23 * dw_dy;
    sbar[4] = t12 * du_dz + t13 * du_dy + t22 * 2 * dv_dz - t23 * du_dx + t33 * 2 * dw_dy;
    sbar[5] = t13 * du_dz + t23 * dv_dz + t33 * (dw_dz - dv_dy - du_dx);
    double Gb[3][3];
    Gb[0][0] = t11 * s11 - t22 * s22 - t23 * s23 - t33 * s33;
    Gb[0][1] = t11 * 2 * s12 + t12 * s22 + t13 * s23;
    Gb[0][2] = t11 * 2 * s13 + t12 * s23 + t13 * s33;
    Gb[1][0] = t12 * s11 + t22 * 2 * s12 + t23 * s13;
    Gb[1][1] = -t11 * s11 - t13 * s13 + t22 * s22 - t33 * s33;
    Gb[1][2] = t12 * s13 + t22 * 2 * s23 + t23 * s33;
    Gb[2][0] = t13 * s11 + t23 * s12 + t33 * 2 * s13;
    Gb[2][1] = t13 * s12 + t23 * s22 + t33 * 2 * s23;
    Gb[2][2] = -t11 * s11 - t12 * s12 - t22 * s22 + t33 * s33;
    // plastic part P(a, sigma)
    const double a = adapt[c], third = 1. / 3.;
    const double p11 = ssb[c], p12 = ssb[g.Nc + c], p13 = ssb[2 * g.Nc + c];
    const double p22 = ssb[3 * g.Nc + c], p23 = ssb[4 * g.Nc + c], p33 = ssb[5 * g.Nc + c];
    ab[c] = p11 * third * (s33 + s22 - 2 * s11) - p12 * s12 - p13 * s13 + p22 * third * (s11 - 2 * s22 + s33) -
            p23 * s23 + p33 * third * (s11 + s22 - 2 * s33);
    sbar[0] += a * third * (-2 * p11 + p22 + p33);
    sbar[1] -= a * p12;
    sbar[2] -= a * p13;
    sbar[3] += a * third * (p11 - 2 * p22 + p33);
    sbar[4] -= a * p23;
    sbar[5] += a * third * (p11 + p22 - 2 * p33);
#pragma unroll
    for (int q = 0; q < 6; q++) atomicAdd(&sb_out[q * g.Nc + c], sbar[q]);
    // chi: chi* = chi + dchi - u.grad chi + diff / c0
    dchib[c] = xsb[c];
    const double wdf = M.diff_l * M.diff_l * xsb[c] / M.c0;
    if (wdf != 0.0) {
        const double Dc = dnew[c];
        const bool xin = i > 0, yin = j > 0;
        const double Dxm = xin ? dnew[c - 1] : dold[cidx(g, i - 1, j, k)];
        const double Dxp = dold[cidx(g, i + 1, j, k)];
        const double Dym = yin ? dnew[c - g.Nx] : dold[cidx(g, i, j - 1, k)];
        const double Dyp = dold[cidx(g, i, j + 1, k)];
        double Dzm, Dzp;
        const long nxy = (long)g.Nx * g.Ny;
        if (g.pz) {
            Dzm = k > 0 ? dnew[c - nxy] : dold[cidx(g, i, j, g.Nz - 1)];
            Dzp = dold[cidx(g, i, j, wrapc(k + 1, g.Nz))];
        } else {
            Dzm = k > 0 ? dnew[c - nxy] : 0.0;
            Dzp = k < g.Nz - 1 ? dold[c + nxy] : 0.0;
        }
        const double x0 = chi[c];
        const double xxm = cellv(chi, g, i - 1, j, k), xxp = cellv(chi, g, i + 1, j, k);
        const double xym = cellv(chi, g, i, j - 1, k), xyp = cellv(chi, g, i, j + 1, k);
        const double xzm = cellv(chi, g, i, j, k - 1), xzp = cellv(chi, g, i, j, k + 1);
        const double ix = 1 / g.dx, iy = 1 / g.dy, iz = 1 / g.dz;
        // d(diff)/d(D neighbours), d/d(Dc), d/d(chi values)
        const double gx = .5 * ix * (xxp - xxm), gy = .5 * iy * (xyp - xym), gz = .5 * iz * (xzp - xzm);
        const double lx = ix * ix * (xxp - 2 * x0 + xxm), ly = iy * iy * (xyp - 2 * x0 + xym),
                     lz = iz * iz * (xzp - 2 * x0 + xzm);
        const double Dxb = wdf * .5 * ix * gx, Dyb = wdf * .5 * iy * gy, Dzb = wdf * .5 * iz * gz;
        atomicAdd(&dob[cidx(g, i + 1, j, k)], Dxb);
        if (xin) atomicAdd(&dnb[c - 1], -Dxb);
        else atomicAdd(&dob[cidx(g, i - 1, j, k)], -Dxb);
        atomicAdd(&dob[cidx(g, i, j + 1, k)], Dyb);
        if (yin) atomicAdd(&dnb[c - g.Nx], -Dyb);
        else atomicAdd(&dob[cidx(g, i, j - 1, k)], -Dyb);
        if (g.pz) {
            atomicAdd(&dob[cidx(g, i, j, wrapc(k + 1, g.Nz))], Dzb);
            if (k > 0) atomicAdd(&dnb[c - nxy], -Dzb);
            else atomicAdd(&dob[cidx(g, i, j, g.Nz - 1)], -Dzb);
        } else {
            if (k < g.Nz - 1) atomicAdd(&dob[c + nxy], Dzb);
            if (k > 0) atomicAdd(&dnb[c - nxy], -Dzb);
        }
        atomicAdd(&dnb[c], wdf * (lx + ly + lz));
        const double ax = .5 * ix * (Dxp - Dxm) * .5 * ix, ay = .5 * iy * (Dyp - Dym) * .5 * iy,
                     az = .5 * iz * (Dzp - Dzm) * .5 * iz;
        cell_add(xb_out, g, i + 1, j, k, wdf * (ax + Dc * ix * ix));
        cell_add(xb_out, g, i - 1, j, k, wdf * (-ax + Dc * ix * ix));
        cell_add(xb_out, g, i, j + 1, k, wdf * (ay + Dc * iy * iy));
        cell_add(xb_out, g, i, j - 1, k, wdf * (-ay + Dc * iy * iy));
        cell_add(xb_out, g, i, j, k + 1, wdf * (az + Dc * iz * iz));
        cell_add(xb_out, g, i, j, k - 1, wdf * (-az + Dc * iz * iz));
        atomicAdd(&xb_out[c], wdf * (-2 * Dc * (ix * ix + iy * iy + iz * iz)));
    }
    grad_add(ub_out, g, i, j, k, Gb, vb);
}

// Transpose of k_plastic: adapt-bar, dchi-bar and dnew-bar of the cell onto
// sigma-bar, chi-bar and the three parameters (per-block partials of
// d/d s_y, d/d eps0, d/d chi_inf in part[3 * block + q]).
static __global__ __launch_bounds__(kT) void k_adj_plastic(Grid g, Mat M, double dt, const double *__restrict__ sig,
                                                    const double *__restrict__ chi, const double *__restrict__ ab,
                                                    const double *__restrict__ dchib,
                                                    const double *__restrict__ dnb, double *__restrict__ sb_out,
                                                    double *__restrict__ xb_out, double *__restrict__ part) {
    const long c = (long)blockIdx.x * kT + threadIdx.x;
    double pg[3] = {0, 0, 0};
    if (c < g.Nc) {
        double s[6];
#pragma unroll
        for (int q = 0; q < 6; q++) s[q] = sig[q * g.Nc + c];
        const double sb = sbar_of(s);
        if (sb > M.sy) {
            double a, dc, Ja[5], Jc[5];
            adaptive_jac(M, sb, chi[c], dt, a, dc, Ja, Jc);
            const double Dbar = dnb[c];
            const double abar = ab[c] + Dbar * (1. / 2. / M.mu * sb);
            const double cbar = dchib[c];
            double sbb = Dbar * (a / 2. / M.mu) + abar * Ja[0] + cbar * Jc[0];
            xb_out[c] += abar * Ja[1] + cbar * Jc[1];
#pragma unroll
            for (int q = 0; q < 3; q++) pg[q] = abar * Ja[2 + q] + cbar * Jc[2 + q];
            // d sbar / d sigma (Field::dev): sbar = sqrt(Q) with dQ/ds11 = t11
            // (the deviatoric part) and dQ/ds12 = 2 s12, so dsbar/ds11 =
            // t11 / (2 sbar) and dsbar/ds12 = s12 / sbar
            const double p = -(s[0] + s[3] + s[5]) * (1 / 3.);
            const double t11 = s[0] + p, t22 = s[3] + p, t33 = s[5] + p;
            const double f = sbb / sb;
            sb_out[c] += f * t11 * .5;
            sb_out[g.Nc + c] += f * s[1];
            sb_out[2 * g.Nc + c] += f * s[2];
            sb_out[3 * g.Nc + c] += f * t22 * .5;
            sb_out[4 * g.Nc + c] += f * s[4];
            sb_out[5 * g.Nc + c] += f * t33 * .5;
        }
    }
    // block sums of the parameter gradients
    __shared__ double red[3][kT / 64];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        double v = pg[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0;
        for (int w = 0; w < kT / 64; w++) t += red[threadIdx.x][w];
        part[3 * (long)blockIdx.x + threadIdx.x] = t;
    }
}

// Adjoint of the loss J = mean_c (ad_Dpl_c / dt)^2 at the final step.
static __global__ __launch_bounds__(kT) void k_adj_loss(Grid g, double dt, const double *__restrict__ dpl,
                                                 double *__restrict__ db) {
    const long c = (long)blockIdx.x * kT + threadIdx.x;
    if (c >= g.Nc) return;
    db[c] = 2.0 * dpl[c] / ((double)g.Nc * dt * dt);
}

}  // namespace stz
