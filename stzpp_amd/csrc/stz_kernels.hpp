// Device kernels of the quasi-static STZ step (gfx950, FP64).
//
// Layout in HBM (SoA, x fastest, no ghost padding):
//   vel   3 x Nv   vertex velocity, Nv = Nx*Ny*Nzv (Nzv = Nz+1 walled, Nz periodic);
//                  this is also the multigrid's level-0 layout
//   sig   6 x Nc   cell stress (s11 s12 s13 s22 s23 s33), Nc = Nx*Ny*Nz
//   chi   Nc       effective temperature
//   dpl   Nc       ad_Dpl
// Ghost values are produced on the fly: periodic wrap in x, y (and z when
// periodic); across a wall the reference's linear extrapolation
// f(-1) = 2f(0) - f(1), f(-2) = 2f(-1) - f(0) of the primary fields, with
// ad_Dpl zero there (set_proper_ghost_for_boundaries,
// src/shear_sim_3d.cpp:2662-2789).
#pragma once

#include <hip/hip_runtime.h>

namespace stz {

constexpr int kT = 256;

struct Grid {
    int Nx, Ny, Nz, Nzv;
    int pz;
    long Nc, Nv;
    double dx, dy, dz;
};

struct Mat {
    double mu, lamb, sy, tau0, c0, eps0, delta, omega, bath_temp, chi_inf, ez, diff_l, eta;
};

// Logical block of a 1-D launch.  Workgroups are dealt round-robin over the
// 8 XCDs (block b runs on XCD b % 8), each with its own L2; this gives XCD x
// the contiguous range of logical blocks starting at x q + min(x, r)
// (q = n / 8, r = n % 8), so the neighbour rows and planes a stencil reads
// were mostly fetched into the same L2 by the blocks beside it.  A bijection
// of [0, n) for every n.
__device__ __forceinline__ long xcd_block() {
    const unsigned n = gridDim.x, b = blockIdx.x;
    const unsigned x = b & 7, q = n >> 3, r = n & 7;
    return (long)(x * q + (x < r ? x : r) + (b >> 3));
}

// Brick order of a 3-D stencil launch: each workgroup of kT threads takes a
// kBT x kBT tile of one plane, and consecutive logical workgroups
// (xcd_block) form bricks of kBX x kBX tiles x kBZ planes, bricks ordered
// x, y, z.  The workgroups resident on one XCD then cover a compact block
// whose stencil neighbourhood (two planes and two points around it) stays in
// that XCD's L2, and the y neighbours of a tile sit in its own rows.
#ifndef STZ_BRICK_X
#define STZ_BRICK_X 4
#endif
#ifndef STZ_BRICK_Z
#define STZ_BRICK_Z 2
#endif
constexpr int kBT = 16, kBX = STZ_BRICK_X, kBZ = STZ_BRICK_Z;
static_assert(kBT * kBT == kT, "one thread per tile point");
struct BrickGrid {
    int ntx, nty, nz;    // tiles per plane in x and y, planes
    int nbx, nby;        // bricks in x and y
};
inline BrickGrid brick_grid(int nx, int ny, int nz) {
    BrickGrid b;
    b.ntx = (nx + kBT - 1) / kBT;
    b.nty = (ny + kBT - 1) / kBT;
    b.nz = nz;
    b.nbx = (b.ntx + kBX - 1) / kBX;
    b.nby = (b.nty + kBX - 1) / kBX;
    return b;
}
inline unsigned brick_blocks(const BrickGrid &b) {
    const long nbz = (b.nz + kBZ - 1) / kBZ;
    return (unsigned)((long)b.nbx * b.nby * nbz * (kBX * kBX * kBZ));
}
// (i, j, k) of this thread; false when it lies outside nx x ny x nz.  Lane 0
// of a wave is inside whenever any of its lanes is (a wave covers four rows
// of a tile from its first column).
__device__ __forceinline__ bool brick_point(const BrickGrid &b, int nx, int ny, int &i, int &j, int &k) {
    const long L = xcd_block();
    const int per = kBX * kBX * kBZ;
    const long br = L / per;
    const int w = (int)(L - br * per);
    const int bi = (int)(br % b.nbx), bj = (int)((br / b.nbx) % b.nby), bk = (int)(br / ((long)b.nbx * b.nby));
    const int ti = bi * kBX + w % kBX, tj = bj * kBX + (w / kBX) % kBX;
    k = bk * kBZ + w / (kBX * kBX);
    i = ti * kBT + (int)(threadIdx.x % kBT);
    j = tj * kBT + (int)(threadIdx.x / kBT);
    return ti < b.ntx && tj < b.nty && k < b.nz && i < nx && j < ny;
}

// Periodic index for -n <= i < 2n: every stencil here reaches at most two
// points past an edge and stz_create requires n >= 4, so one correction each
// way suffices (selects, no loop).
__device__ __forceinline__ int wrapc(int i, int n) {
    i = i < 0 ? i + n : i;
    return i >= n ? i - n : i;
}

__device__ __forceinline__ long cidx(const Grid &g, int i, int j, int k) {
    return wrapc(i, g.Nx) + (long)g.Nx * (wrapc(j, g.Ny) + (long)g.Ny * k);
}

// Cell value with ghosts: k in [-2, Nz+2).
__device__ __forceinline__ double cellv(const double *__restrict__ f, const Grid &g, int i, int j, int k) {
    if (g.pz) return f[cidx(g, i, j, wrapc(k, g.Nz))];
    if (k >= 0 && k < g.Nz) return f[cidx(g, i, j, k)];
    if (k < 0) {
        const double f0 = f[cidx(g, i, j, 0)], f1 = f[cidx(g, i, j, 1)];
        const double m1 = (f0 + f0) - f1;
        return k == -1 ? m1 : (m1 + m1) - f0;
    }
    const double a = f[cidx(g, i, j, g.Nz - 1)], b = f[cidx(g, i, j, g.Nz - 2)];
    const double p0 = (a + a) - b;
    return k == g.Nz ? p0 : (p0 + p0) - a;
}

// Vertex velocity component at (i,j,k), k in [0, Nz] (wrapped when periodic).
__device__ __forceinline__ double vertv(const double *__restrict__ u, const Grid &g, int i, int j, int k) {
    if (g.pz) k = wrapc(k, g.Nz);
    return u[wrapc(i, g.Nx) + (long)g.Nx * (wrapc(j, g.Ny) + (long)g.Ny * k)];
}

// Velocity gradient G[q][a] = d u_q / d x_a and cell-centred velocity of
// cell (i,j,k) from its 8 vertices (calculate_first_order_velocities,
// src/shear_sim_3d.cpp:901-916; staggered average, :429-431).
__device__ __forceinline__ void vel_grad(const double *__restrict__ vel, const Grid &g, int i, int j, int k,
                                         double G[3][3], double vc[3]) {
    const double ix = .25 * (1.0 / g.dx), iy = .25 * (1.0 / g.dy), iz = .25 * (1.0 / g.dz);
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const double *u = vel + q * g.Nv;
        const double u000 = vertv(u, g, i, j, k), u100 = vertv(u, g, i + 1, j, k);
        const double u010 = vertv(u, g, i, j + 1, k), u110 = vertv(u, g, i + 1, j + 1, k);
        const double u001 = vertv(u, g, i, j, k + 1), u101 = vertv(u, g, i + 1, j, k + 1);
        const double u011 = vertv(u, g, i, j + 1, k + 1), u111 = vertv(u, g, i + 1, j + 1, k + 1);
        G[q][0] = ix * (u100 - u000 + u110 - u010 + u101 - u001 + u111 - u011);
        G[q][1] = iy * (u010 - u000 + u110 - u100 + u011 - u001 + u111 - u101);
        G[q][2] = iz * (u001 - u000 + u101 - u100 + u011 - u010 + u111 - u110);
        vc[q] = 0.125 * (u000 + u100 + u001 + u101 + u010 + u110 + u011 + u111);
    }
}

// Offsets of the points up to two either side of point (i, j, k) of a
// field of nz planes of Nx Ny points (index a + 2 for a = -2..2): x and y
// periodic, z periodic when pz and plain otherwise -- valid there for
// 2 <= k <= nz - 3, where no wall ghost is involved.  The same offsets
// serve the vertex field at the same (i, j, k) for a = +1.
struct Stencil5 {
    int ox[5], oy[5], oz[5];
};
__device__ __forceinline__ Stencil5 stencil5(const Grid &g, int i, int j, int k, int nz) {
    Stencil5 s;
    const int nxy = g.Nx * g.Ny;
#pragma unroll
    for (int a = -2; a <= 2; a++) {
        s.ox[a + 2] = wrapc(i + a, g.Nx) - i;
        s.oy[a + 2] = (wrapc(j + a, g.Ny) - j) * g.Nx;
        s.oz[a + 2] = (g.pz ? wrapc(k + a, nz) - k : a) * nxy;
    }
    return s;
}

// vel_grad for vertex index c of (i, j, k) with the offsets of (i+1, j+1,
// k+1) (same arithmetic).
__device__ __forceinline__ void vel_grad_o(const double *__restrict__ vel, const Grid &g, long c, int o1x, int o1y,
                                           int o1z, double G[3][3], double vc[3]) {
    const double ix = .25 * (1.0 / g.dx), iy = .25 * (1.0 / g.dy), iz = .25 * (1.0 / g.dz);
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const double *u = vel + q * g.Nv + c;
        const double u000 = u[0], u100 = u[o1x];
        const double u010 = u[o1y], u110 = u[o1x + o1y];
        const double u001 = u[o1z], u101 = u[o1x + o1z];
        const double u011 = u[o1y + o1z], u111 = u[o1x + o1y + o1z];
        G[q][0] = ix * (u100 - u000 + u110 - u010 + u101 - u001 + u111 - u011);
        G[q][1] = iy * (u010 - u000 + u110 - u100 + u011 - u001 + u111 - u101);
        G[q][2] = iz * (u001 - u000 + u101 - u100 + u011 - u010 + u111 - u110);
        vc[q] = 0.125 * (u000 + u100 + u001 + u101 + u010 + u110 + u011 + u111);
    }
}

// Field::dev (src/field.hh:149-184).
__device__ __forceinline__ double sbar_of(const double s[6]) {
    const double p = -(s[0] + s[3] + s[5]) * (1 / 3.);
    const double t11 = s[0] + p, t22 = s[3] + p, t33 = s[5] + p;
    return sqrt(0.5 * (t11 * t11 + t22 * t22 + t33 * t33) + s[1] * s[1] + s[2] * s[2] + s[4] * s[4]);
}

__device__ __forceinline__ double curly_c(const Mat &M, double sb) {
    const double a = exp((M.omega * M.eps0 * sb - M.delta) / M.bath_temp);
    const double b = exp(-(M.omega * M.eps0 * sb + M.delta) / M.bath_temp);
    return .5 * (a + b);
}

__device__ __forceinline__ double calc_dpl(const Mat &M, double sb, double chi) {
    return sb >= M.sy ? exp(-M.ez / chi) * curly_c(M, sb) * (1 - M.sy / sb) / M.tau0 : 0.0;
}

// adaptive_plastic_term (src/shear_sim_3d.cpp:109-149): returns
// 1 - sbar_end/sbar_0, the plastic chi change in dchi, the substep count in n.
__device__ __forceinline__ double adaptive(const Mat &M, double sb0, double chi0, double dt, double &dchi,
                                           int &n) {
    double tr = dt, ts, sb = sb0, chi = chi0;
    bool keep = true;
    n = 0;
    while (keep) {
        const double dpl = calc_dpl(M, sb, chi);
        const double dprime = 2 * M.mu * dpl / M.sy;
        const double F = 2 * dpl * sb * (M.chi_inf - chi) / (M.sy * M.c0);
        if (dprime * tr > M.eta) {
            ts = M.eta / fabs(dprime);
            tr = tr - ts;
        } else {
            ts = tr;
            keep = false;
        }
        sb = sb - ts * dprime * M.sy;
        chi = chi + ts * F;
        n++;
    }
    dchi = chi - chi0;
    return 1 - sb / sb0;
}

// Adds the number of lanes of the wave with `flag` set to *counter, one
// atomic per wave.  Call it outside divergent code: lane 0 of a wave is
// active whenever any of its lanes is (cells are numbered with the lane).
__device__ __forceinline__ void count_wave(int *counter, bool flag) {
    const unsigned long long b = __ballot(flag);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(counter, (int)__popcll(b));
}

// eno2 (src/shear_sim_3d.cpp:1257-1266), h already divided by dt.
__device__ __forceinline__ double eno2(double h, double v, double f2p, double fp, double f, double fm, double f2m) {
    const double c2 = fabs(fp - 2 * f + fm);
    const double fw = fabs(f2p - 2 * fp + f);
    const double bw = fabs(f - 2 * fm + f2m);
    return .5 / h * (v < 0 ? (c2 > fw ? -f2p + 4 * fp - 3 * f : fp - fm)
                           : (c2 > bw ? 3 * f - 4 * fm + f2m : fp - fm));
}

// Pass 1: pointwise STZ plasticity of every cell (the part of
// advection_grid_step before the advective terms, src/shear_sim_3d.cpp:
// 424-434).  Writes adapt = 2 mu dt Dpl/sbar, the plastic chi change and
// this step's ad_Dpl.
static __global__ __launch_bounds__(kT) void k_plastic(Grid g, Mat M, double dt, const double *__restrict__ sig,
                                                const double *__restrict__ chi, double *__restrict__ adapt,
                                                double *__restrict__ dchi, double *__restrict__ dpl_new,
                                                int *__restrict__ counters) {
    const long c = xcd_block() * kT + threadIdx.x;
    if (c >= g.Nc) return;
    double s[6];
#pragma unroll
    for (int q = 0; q < 6; q++) s[q] = sig[q * g.Nc + c];
    const double sb = sbar_of(s);
    double a = 0, dc = 0;
    if (sb > M.sy) {
        int n;
        a = adaptive(M, sb, chi[c], dt, dc, n);
    }
    count_wave(&counters[0], sb > M.sy);
    adapt[c] = a;
    dchi[c] = dc;
    dpl_new[c] = a / 2. / M.mu * sb;
}

// Right-hand side and warm start of the projection solve (setup_fields with
// x_field / r_field, src/shear_sim_3d.hh:276-385): x = current velocity;
// r = div sigma* at interior vertices (backward 8-cell differences,
// calculate_first_order_stresses, src/shear_sim_3d.cpp:980-996), and
// (-+ fix_a11 * top_u, 0, 0) on the bottom/top plates.
static __global__ __launch_bounds__(kT) void k_rhs(Grid g, const double *__restrict__ sig, const double *__restrict__ vel,
                                            double fix_a11, double top_u, double *__restrict__ x,
                                            double *__restrict__ r) {
    const long p = xcd_block() * kT + threadIdx.x;
    if (p >= g.Nv) return;
    const int i = (int)(p % g.Nx);
    const int j = (int)((p / g.Nx) % g.Ny);
    const int k = (int)(p / ((long)g.Nx * g.Ny));
#pragma unroll
    for (int q = 0; q < 3; q++) x[q * g.Nv + p] = vel[q * g.Nv + p];
    double r0, r1, r2;
    if (!g.pz && k == 0) {
        r0 = -fix_a11 * top_u; r1 = 0; r2 = 0;
    } else if (!g.pz && k == g.Nz) {
        r0 = fix_a11 * top_u; r1 = 0; r2 = 0;
    } else {
        const double ix = .25 * (1.0 / g.dx), iy = .25 * (1.0 / g.dy), iz = .25 * (1.0 / g.dz);
        // cells (i-1..i, j-1..j, k-1..k); the cell index of (i, j, k) is p
        const int mx = wrapc(i - 1, g.Nx) - i, my = (wrapc(j - 1, g.Ny) - j) * g.Nx;
        const int mz = ((g.pz ? wrapc(k - 1, g.Nz) : k - 1) - k) * (g.Nx * g.Ny);
        const long c000 = p, c100 = p + mx, c010 = p + my, c110 = p + mx + my, c001 = p + mz, c101 = p + mx + mz,
                   c011 = p + my + mz, c111 = p + mx + my + mz;
        auto DX = [&](const double *a) {
            return ix * (a[c000] - a[c100] + a[c010] - a[c110] + a[c001] - a[c101] + a[c011] - a[c111]);
        };
        auto DY = [&](const double *a) {
            return iy * (a[c000] - a[c010] + a[c100] - a[c110] + a[c001] - a[c011] + a[c101] - a[c111]);
        };
        auto DZ = [&](const double *a) {
            return iz * (a[c000] - a[c001] + a[c100] - a[c101] + a[c010] - a[c011] + a[c110] - a[c111]);
        };
        const double *S11 = sig, *S12 = sig + g.Nc, *S13 = sig + 2 * g.Nc, *S22 = sig + 3 * g.Nc,
                     *S23 = sig + 4 * g.Nc, *S33 = sig + 5 * g.Nc;
        r0 = DX(S11) + DY(S12) + DZ(S13);
        r1 = DX(S12) + DY(S22) + DZ(S23);
        r2 = DX(S13) + DY(S23) + DZ(S33);
    }
    r[p] = r0;
    r[g.Nv + p] = r1;
    r[2 * g.Nv + p] = r2;
}

// Copy the solution back into the velocity field with the plates set
// exactly (projection copy-out and set_boundary_conditions,
// src/shear_sim_3d.cpp:615-681, :2362-2470).
static __global__ __launch_bounds__(kT) void k_copy_vel(Grid g, const double *__restrict__ x, double top_u,
                                                 double *__restrict__ vel, int bottom = 1) {
    const long p = xcd_block() * kT + threadIdx.x;
    if (p >= g.Nv) return;
    const int k = (int)(p / ((long)g.Nx * g.Ny));
    double v0 = x[p], v1 = x[g.Nv + p], v2 = x[2 * g.Nv + p];
    if (!g.pz && k == 0 && bottom) { v0 = -top_u; v1 = 0; v2 = 0; }
    if (!g.pz && k == g.Nz) { v0 = top_u; v1 = 0; v2 = 0; }
    vel[p] = v0;
    vel[g.Nv + p] = v1;
    vel[2 * g.Nv + p] = v2;
}

// Stress update sigma += dt C:D(u) (src/shear_sim_3d.cpp:684-712), and
// per block the sum of sym_mat3::modsq of the increments (the reference's
// proj_mag, :699-724; modsq, src/mat3.hh:338-340) into pm_part[block].
static __global__ __launch_bounds__(kT) void k_stress_update(Grid g, Mat M, double dt, const double *__restrict__ vel,
                                                      double *__restrict__ sig, double *__restrict__ pm_part) {
    const long c = xcd_block() * kT + threadIdx.x;
    double pm = 0;
    if (c < g.Nc) {
        const int i = (int)(c % g.Nx);
        const int j = (int)((c / g.Nx) % g.Ny);
        const int k = (int)(c / ((long)g.Nx * g.Ny));
        double G[3][3], v[3];
        vel_grad_o(vel, g, c, wrapc(i + 1, g.Nx) - i, (wrapc(j + 1, g.Ny) - j) * g.Nx,
                   (g.pz ? wrapc(k + 1, g.Nz) - k : 1) * (g.Nx * g.Ny), G, v);
        const double lam = M.lamb, mu = M.mu, L2 = lam + 2 * mu;
        const double d11 = dt * (lam * (G[2][2] + G[1][1]) + L2 * G[0][0]);
        const double d12 = dt * mu * (G[0][1] + G[1][0]);
        const double d13 = dt * mu * (G[0][2] + G[2][0]);
        const double d22 = dt * (lam * (G[2][2] + G[0][0]) + L2 * G[1][1]);
        const double d23 = dt * mu * (G[1][2] + G[2][1]);
        const double d33 = dt * (lam * (G[1][1] + G[0][0]) + L2 * G[2][2]);
        sig[c] += d11;
        sig[g.Nc + c] += d12;
        sig[2 * g.Nc + c] += d13;
        sig[3 * g.Nc + c] += d22;
        sig[4 * g.Nc + c] += d23;
        sig[5 * g.Nc + c] += d33;
        pm = 0.5 * (d11 * d11 + d12 * d12 + d13 * d13 + d12 * d12 + d22 * d22 + d23 * d23 + d13 * d13 +
                    d23 * d23 + d33 * d33);
    }
    __shared__ double red[kT / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) pm += __shfl_down(pm, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = pm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < kT / 64; w++) t += red[w];
        pm_part[blockIdx.x] = t;
    }
}

// out[0] = sum of part[0 .. n) in a fixed order (one block).
static __global__ __launch_bounds__(kT) void k_sum_parts(const double *__restrict__ part, long n,
                                                  double *__restrict__ out) {
    double t = 0;
    for (long b = threadIdx.x; b < n; b += kT) t += part[b];
    __shared__ double red[kT];
    red[threadIdx.x] = t;
    __syncthreads();
    for (int st = kT / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// ---- transformed-coordinate step (trans_sim_3d) ---------------------------
//
// State is the transformed velocity U (vertices), transformed stress sigma'
// and chi (cells).  Xform holds T, T^-1 and dT/dt at the time a kernel
// works at (trans_sim_3d::update_matrices, src/trans_sim_3d.hh:132-140).
struct Xform {
    double T[9], Ti[9], dT[9];
};

// 3x3 algebra in the summation order of src/mat3.hh:139-187.
__device__ __forceinline__ void m3mul(const double *a, const double *b, double *o) {
    double r[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) r[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
#pragma unroll
    for (int e = 0; e < 9; e++) o[e] = r[e];
}
__device__ __forceinline__ void m3t(const double *a, double *o) {
    const double r[9] = {a[0], a[3], a[6], a[1], a[4], a[7], a[2], a[5], a[8]};
#pragma unroll
    for (int e = 0; e < 9; e++) o[e] = r[e];
}
__device__ __forceinline__ void m3vec(const double *a, const double *v, double *o) {
    const double r0 = a[0] * v[0] + a[1] * v[1] + a[2] * v[2];
    const double r1 = a[3] * v[0] + a[4] * v[1] + a[5] * v[2];
    const double r2 = a[6] * v[0] + a[7] * v[1] + a[8] * v[2];
    o[0] = r0; o[1] = r1; o[2] = r2;
}

struct Box {
    double ax, ay, az, bx, by, bz;
};

// L = T^-t grad_X(dT/dt X + T U) over the cell whose lowest vertex is
// (i,j,k) and whose corner 0 is placed at P0 (compute_L,
// src/trans_sim_3d.cpp:327-401).  Corner n of the reference's numbering
// (counter-clockwise on the bottom face, then the top face).
__device__ __forceinline__ void compute_L(const double *__restrict__ vel, const Grid &g, const Box &b,
                                          const Xform &X, int i, int j, int k, const double P0[3], double L[9]) {
    const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    const double pn[3] = {P0[0] == b.bx ? b.ax : P0[0] + g.dx, P0[1] == b.by ? b.ay : P0[1] + g.dy,
                          (g.pz && P0[2] == b.bz) ? b.az : P0[2] + g.dz};
    double v[8][3];
#pragma unroll
    for (int n = 0; n < 8; n++) {
        double P[3], Vl[3], a[3], c[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            P[d] = off[n][d] ? pn[d] : P0[d];
            Vl[d] = vertv(vel + d * g.Nv, g, i + off[n][0], j + off[n][1], k + off[n][2]);
        }
        m3vec(X.dT, P, a);
        m3vec(X.T, Vl, c);
#pragma unroll
        for (int d = 0; d < 3; d++) v[n][d] = a[d] + c[d];
    }
    double G[9];
    const double ix = .25 * (1. / g.dx), iy = .25 * (1. / g.dy), iz = .25 * (1. / g.dz);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        G[d] = ix * (((((((v[1][d] - v[0][d]) + v[2][d]) - v[3][d]) + v[5][d]) - v[4][d]) + v[6][d]) - v[7][d]);
        G[3 + d] = iy * (((((((v[3][d] - v[0][d]) + v[2][d]) - v[1][d]) + v[7][d]) - v[4][d]) + v[6][d]) - v[5][d]);
        G[6 + d] = iz * (((((((v[4][d] - v[0][d]) + v[5][d]) - v[1][d]) + v[7][d]) - v[3][d]) + v[6][d]) - v[2][d]);
    }
    double Tti[9];
    m3t(X.Ti, Tti);
    m3mul(Tti, G, L);
}

// compute_L for cell c of (i, j, k) with the vertex offsets of (i+1, j+1,
// k+1) (same arithmetic).
__device__ __forceinline__ void compute_L_o(const double *__restrict__ vel, const Grid &g, const Box &b,
                                            const Xform &X, long c, int o1x, int o1y, int o1z, const double P0[3],
                                            double L[9]) {
    const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    const double pn[3] = {P0[0] == b.bx ? b.ax : P0[0] + g.dx, P0[1] == b.by ? b.ay : P0[1] + g.dy,
                          (g.pz && P0[2] == b.bz) ? b.az : P0[2] + g.dz};
    double v[8][3];
#pragma unroll
    for (int n = 0; n < 8; n++) {
        double P[3], Vl[3], a[3], cc[3];
        const long q = c + (off[n][0] ? o1x : 0) + (off[n][1] ? o1y : 0) + (off[n][2] ? o1z : 0);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            P[d] = off[n][d] ? pn[d] : P0[d];
            Vl[d] = vel[d * g.Nv + q];
        }
        m3vec(X.dT, P, a);
        m3vec(X.T, Vl, cc);
#pragma unroll
        for (int d = 0; d < 3; d++) v[n][d] = a[d] + cc[d];
    }
    double G[9];
    const double ix = .25 * (1. / g.dx), iy = .25 * (1. / g.dy), iz = .25 * (1. / g.dz);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        G[d] = ix * (((((((v[1][d] - v[0][d]) + v[2][d]) - v[3][d]) + v[5][d]) - v[4][d]) + v[6][d]) - v[7][d]);
        G[3 + d] = iy * (((((((v[3][d] - v[0][d]) + v[2][d]) - v[1][d]) + v[7][d]) - v[4][d]) + v[6][d]) - v[5][d]);
        G[6 + d] = iz * (((((((v[4][d] - v[0][d]) + v[5][d]) - v[1][d]) + v[7][d]) - v[3][d]) + v[6][d]) - v[2][d]);
    }
    double Tti[9];
    m3t(X.Ti, Tti);
    m3mul(Tti, G, L);
}

// Physical stress T sigma' T^t of cell c (compute_old_sig,
// src/trans_sim_3d.hh:147) and its sbar (calc_devsq, :151-165).
__device__ __forceinline__ void old_sig(const double *__restrict__ sig, long Nc, long c, const Xform &X,
                                        double sp[9], double os[9]) {
    const double s11 = sig[c], s12 = sig[Nc + c], s13 = sig[2 * Nc + c];
    const double s22 = sig[3 * Nc + c], s23 = sig[4 * Nc + c], s33 = sig[5 * Nc + c];
    const double q[9] = {s11, s12, s13, s12, s22, s23, s13, s23, s33};
#pragma unroll
    for (int e = 0; e < 9; e++) sp[e] = q[e];
    double Tt[9], tmp[9];
    m3t(X.T, Tt);
    m3mul(X.T, sp, tmp);
    m3mul(tmp, Tt, os);
}
__device__ __forceinline__ double sbar_phys(const double os[9]) {
    const double p = -1. / 3. * (os[0] + os[4] + os[8]);
    const double t11 = os[0] + p, t22 = os[4] + p, t33 = os[8] + p;
    return sqrt(0.5 * (t11 * t11 + t22 * t22 + t33 * t33) + os[1] * os[1] + os[2] * os[2] + os[5] * os[5]);
}

// Pass 1 of the transformed step: STZ plasticity on the physical stress
// (advection_grid_step, src/trans_sim_3d.cpp:940-969, with
// adaptive_plastic_term :740-775, eta = M.eta).  Writes adapt, the plastic
// chi change and this step's ad_Dpl (0 below yield).
static __global__ __launch_bounds__(kT) void k_plastic_trans(Grid g, Mat M, double dt, Xform X,
                                                      const double *__restrict__ sig,
                                                      const double *__restrict__ chi, double *__restrict__ adapt,
                                                      double *__restrict__ dchi, double *__restrict__ dpl_new,
                                                      int *__restrict__ counters) {
    const long c = xcd_block() * kT + threadIdx.x;
    if (c >= g.Nc) return;
    double sp[9], os[9];
    old_sig(sig, g.Nc, c, X, sp, os);
    const double sb = sbar_phys(os);
    double a = 0, dc = 0, D = 0;
    if (sb > M.sy) {
        int n;
        a = adaptive(M, sb, chi[c], dt, dc, n);
        D = .5 * a / M.mu * sb;
    }
    count_wave(&counters[0], sb > M.sy);
    adapt[c] = a;
    dchi[c] = dc;
    dpl_new[c] = D;
}

// Right-hand side T div sigma' at interior vertices, zero on the walls, and
// the warm start x = U (r_field, src/trans_sim_3d.hh:202-248).
static __global__ __launch_bounds__(kT) void k_rhs_trans(Grid g, Xform X, const double *__restrict__ sig,
                                                  const double *__restrict__ vel, double *__restrict__ x,
                                                  double *__restrict__ r) {
    const long p = xcd_block() * kT + threadIdx.x;
    if (p >= g.Nv) return;
    const int i = (int)(p % g.Nx);
    const int j = (int)((p / g.Nx) % g.Ny);
    const int k = (int)(p / ((long)g.Nx * g.Ny));
#pragma unroll
    for (int q = 0; q < 3; q++) x[q * g.Nv + p] = vel[q * g.Nv + p];
    double o[3] = {0, 0, 0};
    if (g.pz || (k != 0 && k != g.Nz)) {
        const double ix = .25 * (1.0 / g.dx), iy = .25 * (1.0 / g.dy), iz = .25 * (1.0 / g.dz);
        // cells (i-1..i, j-1..j, k-1..k); the cell index of (i, j, k) is p
        const int mx = wrapc(i - 1, g.Nx) - i, my = (wrapc(j - 1, g.Ny) - j) * g.Nx;
        const int mz = ((g.pz ? wrapc(k - 1, g.Nz) : k - 1) - k) * (g.Nx * g.Ny);
        const long c000 = p, c100 = p + mx, c010 = p + my, c110 = p + mx + my, c001 = p + mz, c101 = p + mx + mz,
                   c011 = p + my + mz, c111 = p + mx + my + mz;
        auto DX = [&](const double *a) {
            return ix * (a[c000] - a[c100] + a[c010] - a[c110] + a[c001] - a[c101] + a[c011] - a[c111]);
        };
        auto DY = [&](const double *a) {
            return iy * (a[c000] - a[c010] + a[c100] - a[c110] + a[c001] - a[c011] + a[c101] - a[c111]);
        };
        auto DZ = [&](const double *a) {
            return iz * (a[c000] - a[c001] + a[c100] - a[c101] + a[c010] - a[c011] + a[c110] - a[c111]);
        };
        const double *S11 = sig, *S12 = sig + g.Nc, *S13 = sig + 2 * g.Nc, *S22 = sig + 3 * g.Nc,
                     *S23 = sig + 4 * g.Nc, *S33 = sig + 5 * g.Nc;
        const double d[3] = {DX(S11) + DY(S12) + DZ(S13), DX(S12) + DY(S22) + DZ(S23),
                             DX(S13) + DY(S23) + DZ(S33)};
        m3vec(X.T, d, o);
    }
    r[p] = o[0];
    r[g.Nv + p] = o[1];
    r[2 * g.Nv + p] = o[2];
}

// sigma' += dt T^-1 C:D(L) T^-t with L at the cell's lowest vertex position
// (projection_step, src/trans_sim_3d.cpp:1096-1118).
static __global__ __launch_bounds__(kT) void k_stress_update_trans(Grid g, Mat M, double dt, Xform X, Box B,
                                                            const double *__restrict__ vel,
                                                            double *__restrict__ sig) {
    const long c = xcd_block() * kT + threadIdx.x;
    if (c >= g.Nc) return;
    const int i = (int)(c % g.Nx);
    const int j = (int)((c / g.Nx) % g.Ny);
    const int k = (int)(c / ((long)g.Nx * g.Ny));
    const double P0[3] = {B.ax + (double)i * g.dx, B.ay + (double)j * g.dy, B.az + (double)k * g.dz};
    double L[9], Lt[9], CD[9], A[9], Bm[9], Tti[9];
    compute_L_o(vel, g, B, X, c, wrapc(i + 1, g.Nx) - i, (wrapc(j + 1, g.Ny) - j) * g.Nx,
                (g.pz ? wrapc(k + 1, g.Nz) - k : 1) * (g.Nx * g.Ny), P0, L);
    m3t(L, Lt);
    m3t(X.Ti, Tti);
    const double tr = L[0] + L[4] + L[8];
#pragma unroll
    for (int e = 0; e < 9; e++) CD[e] = ((e % 4 == 0) ? M.lamb * tr : 0.0) + M.mu * (L[e] + Lt[e]);
#pragma unroll
    for (int e = 0; e < 9; e++) A[e] = X.Ti[e] * dt;
    m3mul(A, CD, Bm);
    m3mul(Bm, Tti, A);
    const int comp[6] = {0, 1, 2, 4, 5, 8};
#pragma unroll
    for (int q = 0; q < 6; q++) sig[q * g.Nc + c] += A[comp[q]];
}

// Host AoS <-> device SoA for arbitrary component counts.
static __global__ void k_aos2soa(const double *__restrict__ a, double *__restrict__ s, long n, int nc) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    for (int q = 0; q < nc; q++) s[q * n + p] = a[(long)nc * p + q];
}
static __global__ void k_soa2aos(const double *__restrict__ s, double *__restrict__ a, long n, int nc) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    for (int q = 0; q < nc; q++) a[(long)nc * p + q] = s[q * n + p];
}

}  // namespace stz
