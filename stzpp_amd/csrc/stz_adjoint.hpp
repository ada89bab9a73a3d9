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
//
// The transformed step (trans_sim_3d) has the same stages with the velocity
// gradient L = T^-t grad(dT X + T U), the L and T terms in the advection,
// sbar of the physical stress T sigma' T^t, the right-hand side T div sigma'
// and the stress update dt T^-1 C:D(L) T^-t; its kernels carry the _trans
// suffix.  The metric chi diffusion is transposed term by term, including
// the reference's y-z difference whose last two cells cancel.
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

// The cell-velocity part of grad_add alone (v = mean of the 8 vertices).
__device__ __forceinline__ void grad_add_v(double *ub, const Grid &g, int i, int j, int k, const double vb[3]) {
#pragma unroll
    for (int n = 0; n < 8; n++) {
        const long p = vidx(g, i + (n & 1), j + ((n >> 1) & 1), k + (n >> 2));
#pragma unroll
        for (int q = 0; q < 3; q++) atomicAdd(&ub[q * g.Nv + p], 0.125 * vb[q]);
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
    const long c = xcd_block() * kT + threadIdx.x;
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
    const long p = xcd_block() * kT + threadIdx.x;
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
    const long c = xcd_block() * kT + threadIdx.x;
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

// Transpose of pass 2 of the plain step (k_advect_march<false, PZ>).  In: the pre-step state (vel, sig, chi, the old
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
    const long c = xcd_block() * kT + threadIdx.x;
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
    const long c = xcd_block() * kT + threadIdx.x;
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

// ---- transformed-coordinate step (trans_sim_3d) ----------------------------
//
// 3x3 adjoints.  For C = A B: A-bar += C-bar B^T, B-bar += A^T C-bar.  A
// symmetric matrix built from six components (upper triangle) collects
// both triangles of its adjoint.

__device__ __forceinline__ void mat_to_sym6_adj(const double *Mb, double *s6) {
    s6[0] += Mb[0];
    s6[1] += Mb[1] + Mb[3];
    s6[2] += Mb[2] + Mb[6];
    s6[3] += Mb[4];
    s6[4] += Mb[5] + Mb[7];
    s6[5] += Mb[8];
}

// Transpose of compute_L for one cell: L = T^-t G with
// G[a][d] = i_a sum_n s_a(n) (dT P_n + T U_n)[d]; Lbar -> U-bar at the 8
// corner vertices (the positions are constants).
__device__ __forceinline__ void L_adj(double *ub, const Grid &g, const Xform &X, int i, int j, int k,
                                      const double Lb[9]) {
    // Gbar = T^-1 Lbar
    double Gb[9];
    m3mul(X.Ti, Lb, Gb);
    const double ia[3] = {.25 * (1. / g.dx), .25 * (1. / g.dy), .25 * (1. / g.dz)};
    double Tt[9];
    m3t(X.T, Tt);
    const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
#pragma unroll
    for (int n = 0; n < 8; n++) {
        double vb[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double t = 0;
#pragma unroll
            for (int a = 0; a < 3; a++) t += ia[a] * (off[n][a] ? Gb[3 * a + d] : -Gb[3 * a + d]);
            vb[d] = t;
        }
        double u[3];
        m3vec(Tt, vb, u);
        const long p = vidx(g, i + off[n][0], j + off[n][1], k + off[n][2]);
#pragma unroll
        for (int d = 0; d < 3; d++) atomicAdd(&ub[d * g.Nv + p], u[d]);
    }
}

// sigma' += dt T^-1 C:D(L) T^-t: U-bar += its transpose applied to sigma'-bar.
static __global__ __launch_bounds__(kT) void k_adj_stress_trans(Grid g, Mat M, double dt, Xform X,
                                                                const double *__restrict__ sb,
                                                                double *__restrict__ ub) {
    const long c = xcd_block() * kT + threadIdx.x;
    if (c >= g.Nc) return;
    const int i = (int)(c % g.Nx), j = (int)((c / g.Nx) % g.Ny), k = (int)(c / ((long)g.Nx * g.Ny));
    const int comp[6] = {0, 1, 2, 4, 5, 8};
    double Ab[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 6; q++) Ab[comp[q]] = sb[q * g.Nc + c];
    // A = (dt T^-1) CD T^-t  ->  CDbar = dt T^-t Abar T^-1
    double Tti[9], tmp[9], CDb[9];
    m3t(X.Ti, Tti);
    m3mul(Tti, Ab, tmp);
    m3mul(tmp, X.Ti, CDb);
#pragma unroll
    for (int e = 0; e < 9; e++) CDb[e] *= dt;
    // CD = lamb tr(L) I + mu (L + L^t)
    const double trb = CDb[0] + CDb[4] + CDb[8];
    double Lb[9];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) Lb[3 * a + b] = M.mu * (CDb[3 * a + b] + CDb[3 * b + a]) + (a == b ? M.lamb * trb : 0.0);
    L_adj(ub, g, X, i, j, k, Lb);
}

// r = T div sigma*' at interior vertices: sigma*'-bar = sigma'-bar + div^T (T^t lambda).
static __global__ __launch_bounds__(kT) void k_adj_div_trans(Grid g, Xform X, const double *__restrict__ lam,
                                                             const double *__restrict__ sb,
                                                             double *__restrict__ ssb) {
    const long c = xcd_block() * kT + threadIdx.x;
    if (c >= g.Nc) return;
    const int i = (int)(c % g.Nx), j = (int)((c / g.Nx) % g.Ny), k = (int)(c / ((long)g.Nx * g.Ny));
    const double ix = .25 * (1.0 / g.dx), iy = .25 * (1.0 / g.dy), iz = .25 * (1.0 / g.dz);
    double Tt[9];
    m3t(X.T, Tt);
    double a11 = 0, a12 = 0, a13 = 0, a22 = 0, a23 = 0, a33 = 0;
#pragma unroll
    for (int n = 0; n < 8; n++) {
        const int di = n & 1, dj = (n >> 1) & 1, dk = n >> 2;
        const int kv = k + dk;
        if (!g.pz && (kv == 0 || kv == g.Nz)) continue;
        const long p = vidx(g, i + di, j + dj, kv);
        const double lv[3] = {lam[p], lam[g.Nv + p], lam[2 * g.Nv + p]};
        double l[3];
        m3vec(Tt, lv, l);
        const double sx = di ? -ix : ix, sy = dj ? -iy : iy, sz = dk ? -iz : iz;
        a11 += sx * l[0];
        a12 += sy * l[0] + sx * l[1];
        a13 += sz * l[0] + sx * l[2];
        a22 += sy * l[1];
        a23 += sz * l[1] + sy * l[2];
        a33 += sz * l[2];
    }
    ssb[c] = sb[c] + a11;
    ssb[g.Nc + c] = sb[g.Nc + c] + a12;
    ssb[2 * g.Nc + c] = sb[2 * g.Nc + c] + a13;
    ssb[3 * g.Nc + c] = sb[3 * g.Nc + c] + a22;
    ssb[4 * g.Nc + c] = sb[4 * g.Nc + c] + a23;
    ssb[5 * g.Nc + c] = sb[5 * g.Nc + c] + a33;
}

// Transpose of pass 2 of the transformed step (k_advect_march<true, PZ>,
// see there for the forward terms).  Writes
// adapt-bar and dchi-bar of the cell; accumulates sigma'-bar (including the
// physical-stress part of the plastic update), chi-bar, U-bar, dnew-bar and
// dold-bar.
static __global__ __launch_bounds__(kT) void k_adj_advect_trans(
    Grid g, Mat M, double dt, Xform X, Box B, const double *__restrict__ vel, const double *__restrict__ sig,
    const double *__restrict__ chi, const double *__restrict__ adapt, const double *__restrict__ dnew,
    const double *__restrict__ dold, const double *__restrict__ ssb, const double *__restrict__ xsb,
    double *__restrict__ sb_out, double *__restrict__ xb_out, double *__restrict__ ub_out, double *__restrict__ dnb,
    double *__restrict__ dob, double *__restrict__ ab, double *__restrict__ dchib) {
    const long c = xcd_block() * kT + threadIdx.x;
    if (c >= g.Nc) return;
    const int i = (int)(c % g.Nx), j = (int)((c / g.Nx) % g.Ny), k = (int)(c / ((long)g.Nx * g.Ny));
    const double P0[3] = {B.ax + ((double)i + .5) * g.dx, B.ay + ((double)j + .5) * g.dy,
                          B.az + ((double)k + .5) * g.dz};
    double L[9];
    compute_L(vel, g, B, X, i, j, k, P0, L);
    double v[3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const double *u = vel + q * g.Nv;
        v[q] = 0.125 * (vertv(u, g, i, j, k) + vertv(u, g, i + 1, j, k) + vertv(u, g, i, j, k + 1) +
                        vertv(u, g, i + 1, j, k + 1) + vertv(u, g, i, j + 1, k) + vertv(u, g, i + 1, j + 1, k) +
                        vertv(u, g, i, j + 1, k + 1) + vertv(u, g, i + 1, j + 1, k + 1));
    }
    double sp[9], os[9], Tti[9];
    old_sig(sig, g.Nc, c, X, sp, os);
    m3t(X.Ti, Tti);
    const double h[3] = {g.dx / dt, g.dy / dt, g.dz / dt};
    const int comp[6] = {0, 1, 2, 4, 5, 8};
    double gbar[7];
#pragma unroll
    for (int q = 0; q < 6; q++) gbar[q] = -ssb[q * g.Nc + c];
    gbar[6] = -xsb[c];
    // ENO advection of sigma' and chi (x derivative of chi: f(+1) as f(+2))
    double vb[3] = {0, 0, 0};
    for (int q = 0; q < 7; q++) {
        const double *f = q < 6 ? sig + q * g.Nc : chi;
        double *fb = q < 6 ? sb_out + q * g.Nc : xb_out;
        const double gq = gbar[q];
        if (gq == 0.0) continue;
#pragma unroll
        for (int ax = 0; ax < 3; ax++) {
            const int ex = ax == 0, ey = ax == 1, ez = ax == 2;
            const bool quirk = q == 6 && ax == 0;
            double fv[5];
#pragma unroll
            for (int o = 0; o < 5; o++) {
                int s = 2 - o;
                if (quirk && s == 2) s = 1;
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
                    int s = 2 - o;
                    if (quirk && s == 2) s = 1;
                    cell_add(fb, g, i + s * ex, j + s * ey, k + s * ez, gq * v[ax] * w[o]);
                }
        }
    }
    // dt (L term + T term) and the plastic update of sigma*'
    double LTb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Pb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 6; q++) {
        LTb[comp[q]] = dt * ssb[q * g.Nc + c];
        Pb[comp[q]] = -ssb[q * g.Nc + c];
    }
    double spb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, osb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Lb[9];
    double tmp[9], tmp2[9];
    // T term = -(Mt sp + sp Mt^t), Mt = T^-1 dT
    {
        double Mt[9], Mtt[9];
        m3mul(X.Ti, X.dT, Mt);
        m3t(Mt, Mtt);
        m3mul(Mtt, LTb, tmp);   // Mt^T LTb
        m3mul(LTb, Mt, tmp2);   // LTb Mt
#pragma unroll
        for (int e = 0; e < 9; e++) spb[e] -= tmp[e] + tmp2[e];
    }
    // L term = T^-1 K T^-t, K = -tr(L) os + os L + L^t os
    {
        double Kb[9], Lt[9];
        m3mul(Tti, LTb, tmp);
        m3mul(tmp, X.Ti, Kb);   // Kbar = T^-t LTb T^-1
        m3t(L, Lt);
        const double trL = L[0] + L[4] + L[8];
        m3mul(Kb, Lt, tmp);     // Kb L^t
        m3mul(L, Kb, tmp2);     // L Kb
#pragma unroll
        for (int e = 0; e < 9; e++) osb[e] += -trL * Kb[e] + tmp[e] + tmp2[e];
        double KbO = 0;
#pragma unroll
        for (int e = 0; e < 9; e++) KbO += Kb[e] * os[e];
        double osT[9], Kbt[9];
        m3t(os, osT);
        m3t(Kb, Kbt);
        m3mul(osT, Kb, tmp);    // os^T Kb   (from os L)
        m3mul(os, Kbt, tmp2);   // os Kb^T   (from L^t os)
#pragma unroll
        for (int e = 0; e < 9; e++) Lb[e] = tmp[e] + tmp2[e] - (e % 4 == 0 ? KbO : 0.0);
    }
    // plastic update T^-1 (a dev(os)) T^-t, dev from the upper triangle
    const double a = adapt[c];
    double abar = 0;
    {
        double Qb[9];
        m3mul(Tti, Pb, tmp);
        m3mul(tmp, X.Ti, Qb);   // Qbar = T^-t Pbar T^-1, Q = (T^-1 a) dev T^-t -> same
        const double third = 1. / 3., tr = os[0] + os[4] + os[8];
        const double dv[9] = {os[0] - third * tr, os[1], os[2], os[1], os[4] - third * tr, os[5],
                              os[2], os[5], os[8] - third * tr};
#pragma unroll
        for (int e = 0; e < 9; e++) abar += Qb[e] * dv[e];
        // dev-bar -> os-bar (upper triangle and diagonal)
        const double dtr = a * (Qb[0] + Qb[4] + Qb[8]) * third;
        osb[0] += a * Qb[0] - dtr;
        osb[4] += a * Qb[4] - dtr;
        osb[8] += a * Qb[8] - dtr;
        osb[1] += a * (Qb[1] + Qb[3]);
        osb[2] += a * (Qb[2] + Qb[6]);
        osb[5] += a * (Qb[5] + Qb[7]);
    }
    ab[c] = abar;
    // os = T sp T^t  ->  spbar += T^t osbar T
    {
        double Tt[9];
        m3t(X.T, Tt);
        m3mul(Tt, osb, tmp);
        m3mul(tmp, X.T, tmp2);
#pragma unroll
        for (int e = 0; e < 9; e++) spb[e] += tmp2[e];
    }
    double s6[6] = {0, 0, 0, 0, 0, 0};
    mat_to_sym6_adj(spb, s6);
#pragma unroll
    for (int q = 0; q < 6; q++) atomicAdd(&sb_out[q * g.Nc + c], s6[q]);
    L_adj(ub_out, g, X, i, j, k, Lb);
    // chi: chi* = chi + dchi - u.grad chi + diff_J / c0
    dchib[c] = xsb[c];
    const double wdf = M.diff_l * M.diff_l * xsb[c] / M.c0;
    if (wdf != 0.0) {
        const double Dc = dnew[c];
        const bool xin = i > 0, yin = j > 0;
        const long nxy = (long)g.Nx * g.Ny;
        const double Dxm = xin ? dnew[c - 1] : dold[cidx(g, i - 1, j, k)];
        const double Dxp = dold[cidx(g, i + 1, j, k)];
        const double Dym = yin ? dnew[c - g.Nx] : dold[cidx(g, i, j - 1, k)];
        const double Dyp = dold[cidx(g, i, j + 1, k)];
        double Dzm, Dzp;
        if (g.pz) {
            Dzm = k > 0 ? dnew[c - nxy] : dold[cidx(g, i, j, g.Nz - 1)];
            Dzp = dold[cidx(g, i, j, wrapc(k + 1, g.Nz))];
        } else {
            Dzm = k > 0 ? dnew[c - nxy] : 0.0;
            Dzp = k < g.Nz - 1 ? dold[c + nxy] : 0.0;
        }
        const double ix = 1. / g.dx, iy = 1. / g.dy, iz = 1. / g.dz;
        const double x0 = chi[c];
        const double xxm = cellv(chi, g, i - 1, j, k), xxp = cellv(chi, g, i + 1, j, k);
        const double xym = cellv(chi, g, i, j - 1, k), xyp = cellv(chi, g, i, j + 1, k);
        const double xzm = cellv(chi, g, i, j, k - 1), xzp = cellv(chi, g, i, j, k + 1);
        double J[9], Jt[9];
        m3mul(X.Ti, Tti, J);
        m3t(J, Jt);
        const double gc[3] = {.5 * ix * (xxp - xxm), .5 * iy * (xyp - xym), .5 * iz * (xzp - xzm)};
        const double dD[3] = {.5 * ix * (Dxp - Dxm), .5 * iy * (Dyp - Dym), .5 * iz * (Dzp - Dzm)};
        double Jg[3], JtdD[3];
        m3vec(J, gc, Jg);
        m3vec(Jt, dD, JtdD);
        // term1 = dD . (J gc)
        const double Db0 = wdf * .5 * ix * Jg[0], Db1 = wdf * .5 * iy * Jg[1], Db2 = wdf * .5 * iz * Jg[2];
        atomicAdd(&dob[cidx(g, i + 1, j, k)], Db0);
        if (xin) atomicAdd(&dnb[c - 1], -Db0);
        else atomicAdd(&dob[cidx(g, i - 1, j, k)], -Db0);
        atomicAdd(&dob[cidx(g, i, j + 1, k)], Db1);
        if (yin) atomicAdd(&dnb[c - g.Nx], -Db1);
        else atomicAdd(&dob[cidx(g, i, j - 1, k)], -Db1);
        if (g.pz) {
            atomicAdd(&dob[cidx(g, i, j, wrapc(k + 1, g.Nz))], Db2);
            if (k > 0) atomicAdd(&dnb[c - nxy], -Db2);
            else atomicAdd(&dob[cidx(g, i, j, g.Nz - 1)], -Db2);
        } else {
            if (k < g.Nz - 1) atomicAdd(&dob[c + nxy], Db2);
            if (k > 0) atomicAdd(&dnb[c - nxy], -Db2);
        }
        cell_add(xb_out, g, i + 1, j, k, wdf * .5 * ix * JtdD[0]);
        cell_add(xb_out, g, i - 1, j, k, -wdf * .5 * ix * JtdD[0]);
        cell_add(xb_out, g, i, j + 1, k, wdf * .5 * iy * JtdD[1]);
        cell_add(xb_out, g, i, j - 1, k, -wdf * .5 * iy * JtdD[1]);
        cell_add(xb_out, g, i, j, k + 1, wdf * .5 * iz * JtdD[2]);
        cell_add(xb_out, g, i, j, k - 1, -wdf * .5 * iz * JtdD[2]);
        // term2 = Dc sum J_ab d_ab chi (the y-z difference keeps only its
        // first two cells: the other two are the same cell with opposite signs)
        const double dxx = ix * ix * (xxp - 2 * x0 + xxm), dyy = iy * iy * (xyp - 2 * x0 + xym),
                     dzz = iz * iz * (xzp - 2 * x0 + xzm);
        const double dxy = .25 * ix * iy *
                           (cellv(chi, g, i + 1, j + 1, k) + cellv(chi, g, i - 1, j - 1, k) -
                            cellv(chi, g, i + 1, j - 1, k) - cellv(chi, g, i - 1, j + 1, k));
        const double dxz = .25 * ix * iz *
                           (cellv(chi, g, i + 1, j, k + 1) + cellv(chi, g, i - 1, j, k - 1) -
                            cellv(chi, g, i + 1, j, k - 1) - cellv(chi, g, i - 1, j, k + 1));
        const double ypzm = cellv(chi, g, i, j + 1, k - 1);
        const double dyz = .25 * iy * iz *
                           (cellv(chi, g, i, j + 1, k + 1) + cellv(chi, g, i, j - 1, k - 1) - ypzm + ypzm);
        const double t2 = J[0] * dxx + J[1] * dxy + J[2] * dxz + J[3] * dxy + J[4] * dyy + J[5] * dyz +
                          J[6] * dxz + J[7] * dyz + J[8] * dzz;
        atomicAdd(&dnb[c], wdf * t2);
        const double w2 = wdf * Dc;
        const double cxx = w2 * J[0] * ix * ix, cyy = w2 * J[4] * iy * iy, czz = w2 * J[8] * iz * iz;
        const double cxy = w2 * (J[1] + J[3]) * .25 * ix * iy, cxz = w2 * (J[2] + J[6]) * .25 * ix * iz,
                     cyz = w2 * (J[5] + J[7]) * .25 * iy * iz;
        cell_add(xb_out, g, i + 1, j, k, cxx);
        cell_add(xb_out, g, i - 1, j, k, cxx);
        cell_add(xb_out, g, i, j + 1, k, cyy);
        cell_add(xb_out, g, i, j - 1, k, cyy);
        cell_add(xb_out, g, i, j, k + 1, czz);
        cell_add(xb_out, g, i, j, k - 1, czz);
        atomicAdd(&xb_out[c], -2 * (cxx + cyy + czz));
        cell_add(xb_out, g, i + 1, j + 1, k, cxy);
        cell_add(xb_out, g, i - 1, j - 1, k, cxy);
        cell_add(xb_out, g, i + 1, j - 1, k, -cxy);
        cell_add(xb_out, g, i - 1, j + 1, k, -cxy);
        cell_add(xb_out, g, i + 1, j, k + 1, cxz);
        cell_add(xb_out, g, i - 1, j, k - 1, cxz);
        cell_add(xb_out, g, i + 1, j, k - 1, -cxz);
        cell_add(xb_out, g, i - 1, j, k + 1, -cxz);
        cell_add(xb_out, g, i, j + 1, k + 1, cyz);
        cell_add(xb_out, g, i, j - 1, k - 1, cyz);
    }
    grad_add_v(ub_out, g, i, j, k, vb);
}

// Transpose of k_plastic_trans: as k_adj_plastic with sbar taken from the
// physical stress T sigma' T^t.
static __global__ __launch_bounds__(kT) void k_adj_plastic_trans(
    Grid g, Mat M, double dt, Xform X, const double *__restrict__ sig, const double *__restrict__ chi,
    const double *__restrict__ ab, const double *__restrict__ dchib, const double *__restrict__ dnb,
    double *__restrict__ sb_out, double *__restrict__ xb_out, double *__restrict__ part) {
    const long c = xcd_block() * kT + threadIdx.x;
    double pg[3] = {0, 0, 0};
    if (c < g.Nc) {
        double sp[9], os[9];
        old_sig(sig, g.Nc, c, X, sp, os);
        const double sb = sbar_phys(os);
        if (sb > M.sy) {
            double a, dc, Ja[5], Jc[5];
            adaptive_jac(M, sb, chi[c], dt, a, dc, Ja, Jc);
            const double Dbar = dnb[c];
            const double abar = ab[c] + Dbar * (.5 / M.mu * sb);
            const double cbar = dchib[c];
            const double sbb = Dbar * (.5 * a / M.mu) + abar * Ja[0] + cbar * Jc[0];
            xb_out[c] += abar * Ja[1] + cbar * Jc[1];
#pragma unroll
            for (int q = 0; q < 3; q++) pg[q] = abar * Ja[2 + q] + cbar * Jc[2 + q];
            // d sbar / d os (upper triangle and diagonal), then os = T sp T^t
            const double p = -1. / 3. * (os[0] + os[4] + os[8]);
            const double t11 = os[0] + p, t22 = os[4] + p, t33 = os[8] + p;
            const double f = sbb / sb;
            double osb[9] = {f * .5 * t11, f * os[1], f * os[2], 0, f * .5 * t22, f * os[5], 0, 0, f * .5 * t33};
            double Tt[9], tmp[9], spb[9];
            m3t(X.T, Tt);
            m3mul(Tt, osb, tmp);
            m3mul(tmp, X.T, spb);
            double s6[6] = {0, 0, 0, 0, 0, 0};
            mat_to_sym6_adj(spb, s6);
#pragma unroll
            for (int q = 0; q < 6; q++) sb_out[q * g.Nc + c] += s6[q];
        }
    }
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
    const long c = xcd_block() * kT + threadIdx.x;
    if (c >= g.Nc) return;
    db[c] = 2.0 * dpl[c] / ((double)g.Nc * dt * dt);
}

}  // namespace stz
