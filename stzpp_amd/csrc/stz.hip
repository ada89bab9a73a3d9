// Host side of the quasi-static STZ step and its C ABI (include/stz.h).
//
// One step follows shear_sim_3d::step_forward_qs (src/shear_sim_3d.cpp:
// 381-389): plasticity + advection (two passes, k_plastic then k_advect_march, so
// that the reference's in-place ad_Dpl ordering is reproduced without a
// sequential loop), time advance, projection (right-hand side, multigrid
// solve, copy-out with plate values, stress update).  The transformed step
// (trans_sim_3d::step_forward_qs, src/trans_sim_3d.cpp:72-80) has the same
// shape with the transform T(t) at the advection and T(t+dt) at the
// projection, whose operator and Galerkin hierarchy are rebuilt every step
// (set_up_stencil_and_matrices, :1187-1489).  Everything runs on the
// multigrid handle's HIP stream; the state never leaves HBM between steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <utility>

#include "../../include/stz.h"
#include "mg_internal.hpp"
#include "stz_internal.hpp"
#include "stz_advect.hpp"
#include "stz_kernels.hpp"

using namespace stz;

namespace {

int sfail(stz_t s, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (s) s->err = buf;
    return code;
}

#define SHIP(s, call)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) return sfail((s), STZMG_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

inline unsigned nb(long n) { return (unsigned)((n + kT - 1) / kT); }

// Plate speed (ramped U t for t < 1, src/shear_sim_3d.hh:375); the
// transformed simulation's walls are at rest (src/trans_sim_3d.cpp:1122-1184).
double top_u(const stz_t s, double t) { return s->p.transform ? 0.0 : (t < 1 ? s->p.U * t : s->p.U); }

void m3set(double *a, double a11, double a12, double a13, double a21, double a22, double a23, double a31,
           double a32, double a33) {
    a[0] = a11; a[1] = a12; a[2] = a13; a[3] = a21; a[4] = a22; a[5] = a23; a[6] = a31; a[7] = a32; a[8] = a33;
}

// Inverse of a 3x3 matrix by cofactors; false when singular.
bool m3inv(const double *a, double *o) {
    const double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
    const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    if (!(det != 0.0) || !std::isfinite(det)) return false;
    const double id = 1.0 / det;
    o[0] = c00 * id;
    o[1] = (a[2] * a[7] - a[1] * a[8]) * id;
    o[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    o[3] = c01 * id;
    o[4] = (a[0] * a[8] - a[2] * a[6]) * id;
    o[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    o[6] = c02 * id;
    o[7] = (a[1] * a[6] - a[0] * a[7]) * id;
    o[8] = (a[0] * a[4] - a[1] * a[3]) * id;
    return true;
}

Xform xform_at(const stz_t s, double t) {
    Xform X;
    if (s->p.transform == STZ_T_USER) {
        for (int e = 0; e < 9; e++) {
            X.T[e] = s->user_T[e] + (t - s->user_t0) * s->user_dT[e];
            X.dT[e] = s->user_dT[e];
        }
        if (!m3inv(X.T, X.Ti))
            for (int e = 0; e < 9; e++) X.Ti[e] = std::nan("");
        return X;
    }
    stz_transform(s->p.transform, s->p.Ut, s->p.bz, t, X.T, X.Ti, X.dT);
    return X;
}

// Transformed operator at T(t+dt): stencil = -dt s_ent(T), fix on the walls
// (set_up_stencil_and_matrices, src/trans_sim_3d.cpp:1187-1489), with the
// Galerkin hierarchy rebuilt on the GPU.  transpose: load the transposed
// stencil (block(o) <- block(-o)^T) instead; with the same diagonal wall rows
// its interior block is the transpose of the forward operator's.
int trans_operator(stz_t s, double dt, const Xform &X, bool transpose = false) {
    const Grid &g = s->g;
    double B[243], F[9] = {0};
    int rc = stzmg_trans_stencil(X.T, s->p.mu, s->p.lamb, g.dx, g.dy, g.dz, dt, B);
    if (rc) return sfail(s, rc, "transformed stencil");
    if (transpose) {
        double Bt[243];
        for (int b = 0; b < 27; b++)
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) Bt[9 * b + 3 * r + c] = B[9 * (26 - b) + 3 * c + r];
        std::copy(Bt, Bt + 243, B);
    }
    const double fv = -dt * 2 * ((s->p.lamb + 2 * s->p.mu) / g.dx / g.dx + s->p.mu / g.dy / g.dy + s->p.mu / g.dz / g.dz);
    F[0] = F[4] = F[8] = fv;
    rc = stzmg_set_operator_const(s->mg, B, g.pz ? nullptr : F);
    if (rc) return sfail(s, rc, "projection operator: %s", stzmg_last_error(s->mg));
    s->a11 = B[9 * 13];
    s->fix_a11 = F[0];
    s->mg_dt = -1;  // the shear operator is no longer loaded
    return 0;
}

}  // namespace

namespace stz_detail {

Xform xform(const stz_t s, double t) { return xform_at(s, t); }

int transposed_trans_operator(stz_t s, double dt, double t) { return trans_operator(s, dt, xform_at(s, t), true); }

int fail(stz_t s, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (s) s->err = buf;
    return code;
}

// (Re)build the projection operator for step size dt
// (set_up_stencil_and_matrices, src/shear_sim_3d.hh:45-55).
int ensure_operator(stz_t s, double dt) {
    if (s->mg && s->mg_dt == dt) return 0;
    const Grid &g = s->g;
    if (!s->mg) {
        stzmg_desc d;
        stzmg_desc_default(&d, g.Nx, g.Ny, g.Nzv, 1, 1, g.pz);
        int rc = stzmg_create(&d, &s->mg);
        if (rc) return sfail(s, rc, "projection solver: %s", stzmg_last_error(nullptr));
        s->stream = (hipStream_t)stzmg_stream(s->mg);
        // the transformed step sets a new operator every step: keep its
        // bottom sweeps rather than tabulate them each time
        if (s->p.transform != STZ_T_NONE) stzmg_set_option(s->mg, STZMG_OPT_BOTTOM_MAP, 0);
    }
    double B[243], F[9];
    stzmg_shear_stencil(s->p.mu, s->p.lamb, g.dx, g.dy, g.dz, dt, 0.0, B, F);
    int rc = stzmg_set_operator_const(s->mg, B, g.pz ? nullptr : F);
    if (rc) return sfail(s, rc, "projection operator: %s", stzmg_last_error(s->mg));
    s->a11 = B[9 * 13];
    s->fix_a11 = F[0];
    s->mg_dt = dt;
    return 0;
}

}  // namespace stz_detail

namespace {

using stz_detail::ensure_operator;

}  // namespace

extern "C" {

int stz_params_default(stz_params *p, int Nx, int Ny, int Nz) {
    if (!p) return STZMG_ERR_ARG;
    memset(p, 0, sizeof(*p));
    p->Nx = Nx; p->Ny = Ny; p->Nz = Nz;
    p->prd[0] = 1; p->prd[1] = 1; p->prd[2] = 0;
    p->ax = -1; p->bx = 1; p->ay = -1; p->by = 1; p->az = -0.5; p->bz = 0.5;
    // Vitreloy 1 parameters of the shear_sim driver, nondimensionalised by
    // the yield stress, the length 1 cm and the shear-wave time
    // (src/shear_sim.cpp:30-86), with TZ = 21000 K.
    const double TZ = 21000, kb = 1.3806503e-23;
    const double E = 101e9, nu = .35, sy = .85e9;
    const double mu = E / (2 * (1 + nu)), K = E / (3 * (1 - 2 * nu)), rho = 6125;
    const double ts = 1e-2 / std::sqrt(mu / rho);
    p->mu = mu / sy;
    p->lamb = K / sy - (2. / 3.) * p->mu;
    p->sy = 1;
    p->tau0 = 1e-13 / ts;
    p->c0 = .4;
    p->eps0 = .3;
    p->delta = 8000 / TZ;
    p->omega = 300 * 26. / 17. * 1e-30 * sy / (TZ * kb);
    p->bath_temp = 400 / TZ;
    p->chi_inf = 900 / TZ;
    p->ez = 1;
    p->diff_l = 1.5 * (p->bx - p->ax) / Nx;
    p->U = 1e-7;
    p->eta = 0;
    p->transform = STZ_T_NONE;
    p->Ut = 0;
    p->rho = p->mu;
    p->kappa = .35;
    p->tmult = .5;
    return 0;
}

int stz_transform(int fam, double Ut, double bz, double t, double *T, double *Ti, double *dT) {
    if (!T || !Ti || !dT || fam < STZ_T_NONE || fam > STZ_T_IDENTITY) return STZMG_ERR_ARG;
    if (fam == STZ_T_PURE_SHEAR) {
        m3set(T, std::exp(Ut * t), 0, 0, 0, 1, 0, 0, 0, std::exp(-Ut * t));
        m3set(Ti, std::exp(-Ut * t), 0, 0, 0, 1, 0, 0, 0, std::exp(Ut * t));
        m3set(dT, Ut * std::exp(Ut * t), 0, 0, 0, 0, 0, 0, 0, -Ut * std::exp(-Ut * t));
    } else if (fam == STZ_T_SIMPLE_SHEAR) {
        m3set(T, 1, 0, Ut * t / bz, 0, 1, 0, 0, 0, 1);
        m3set(Ti, 1, 0, -Ut * t / bz, 0, 1, 0, 0, 0, 1);
        m3set(dT, 0, 0, Ut / bz, 0, 0, 0, 0, 0, 0);
    } else if (fam == STZ_T_HOMOGENEOUS) {
        const double e = std::exp(Ut * t), ei = std::exp(-Ut * t);
        m3set(T, e, 0, 0, 0, e, 0, 0, 0, e);
        m3set(Ti, ei, 0, 0, 0, ei, 0, 0, 0, ei);
        m3set(dT, Ut * e, 0, 0, 0, Ut * e, 0, 0, 0, Ut * e);
    } else {
        m3set(T, 1, 0, 0, 0, 1, 0, 0, 0, 1);
        m3set(Ti, 1, 0, 0, 0, 1, 0, 0, 0, 1);
        m3set(dT, 0, 0, 0, 0, 0, 0, 0, 0, 0);
    }
    return 0;
}

static thread_local std::string g_stz_err;

int stz_set_transform(stz_t s, const double *T, const double *dTdt) {
    if (!s || !T || !dTdt) return STZMG_ERR_ARG;
    if (s->p.transform == STZ_T_NONE)
        return sfail(s, STZMG_ERR_STATE, "stz_set_transform: the handle runs the untransformed step");
    double Ti[9];
    if (!m3inv(T, Ti)) return sfail(s, STZMG_ERR_SINGULAR, "stz_set_transform: T is singular");
    for (int e = 0; e < 9; e++) {
        if (!std::isfinite(T[e]) || !std::isfinite(dTdt[e]))
            return sfail(s, STZMG_ERR_ARG, "stz_set_transform: non-finite entry");
        s->user_T[e] = T[e];
        s->user_dT[e] = dTdt[e];
    }
    s->user_t0 = s->t;
    s->p.transform = STZ_T_USER;
    return 0;
}

int stz_create(const stz_params *p, stz_t *out) {
    if (!p || !out) return STZMG_ERR_ARG;
    *out = nullptr;
    if (!p->prd[0] || !p->prd[1]) {
        g_stz_err = "x and y must be periodic";
        return STZMG_ERR_UNSUPPORTED;
    }
    if (p->Nx < 4 || p->Ny < 4 || p->Nz < 4) {
        g_stz_err = "grid too small";
        return STZMG_ERR_ARG;
    }
    stz_t s = new stz_s;
    s->p = *p;
    Grid &g = s->g;
    g.Nx = p->Nx; g.Ny = p->Ny; g.Nz = p->Nz;
    g.pz = p->prd[2] != 0;
    g.Nzv = g.Nz + (g.pz ? 0 : 1);
    g.Nc = (long)g.Nx * g.Ny * g.Nz;
    g.Nv = (long)g.Nx * g.Ny * g.Nzv;
    if (g.Nv > (long)std::numeric_limits<int>::max()) {
        // the step kernels index cells and vertices with 32-bit integers
        g_stz_err = "grid too large: at most 2^31 - 1 vertices";
        delete s;
        return STZMG_ERR_UNSUPPORTED;
    }
    g.dx = (p->bx - p->ax) / g.Nx;
    g.dy = (p->by - p->ay) / g.Ny;
    g.dz = (p->bz - p->az) / g.Nz;
    if (p->transform < STZ_T_NONE || p->transform > STZ_T_IDENTITY) {
        g_stz_err = "unknown transform family";
        delete s;
        return STZMG_ERR_ARG;
    }
    const double eta = p->eta > 0 ? p->eta : (p->transform ? 0.002 : 0.0005);
    s->M = Mat{p->mu, p->lamb, p->sy, p->tau0, p->c0, p->eps0, p->delta, p->omega, p->bath_temp,
               p->chi_inf, p->ez, p->diff_l, eta};
    // The solver owns the stream; create it now with a placeholder dt.
    if (int rc = ensure_operator(s, 1.0)) {
        g_stz_err = s->err;
        if (s->mg) stzmg_destroy(s->mg);
        delete s;
        return rc;
    }
    const long Nc = g.Nc, Nv = g.Nv;
    hipError_t e = hipSuccess;
    auto A = [&](double **ptr, long n) {
        if (e == hipSuccess) e = hipMalloc((void **)ptr, sizeof(double) * n);
    };
    A(&s->vel, 3 * Nv); A(&s->sig, 6 * Nc); A(&s->sig2, 6 * Nc); A(&s->chi, Nc); A(&s->chi2, Nc);
    A(&s->dpl, Nc); A(&s->dpl2, Nc); A(&s->adapt, Nc); A(&s->dchi, Nc);
    A(&s->stage, 6 * (Nc > Nv ? Nc : Nv));
    if (e == hipSuccess) e = hipMalloc((void **)&s->counters, sizeof(int) * 4);
    A(&s->pm_part, (long)nb(Nc) + 1);
    if (e != hipSuccess) {
        g_stz_err = "device allocation failed";
        stz_destroy(s);
        return STZMG_ERR_NOMEM;
    }
    hipMemsetAsync(s->dpl, 0, sizeof(double) * Nc, s->stream);
    hipMemsetAsync(s->vel, 0, sizeof(double) * 3 * Nv, s->stream);
    hipMemsetAsync(s->sig, 0, sizeof(double) * 6 * Nc, s->stream);
    hipMemsetAsync(s->chi, 0, sizeof(double) * Nc, s->stream);
    hipStreamSynchronize(s->stream);
    *out = s;
    return 0;
}

int stz_destroy(stz_t s) {
    if (!s) return STZMG_ERR_ARG;
    if (s->stream) hipStreamSynchronize(s->stream);
    for (double *q : {s->vel, s->sig, s->sig2, s->chi, s->chi2, s->dpl, s->dpl2, s->adapt, s->dchi, s->stage, s->xi,
                      s->xi2, s->vel2})
        hipFree(q);
    hipFree(s->counters);
    hipFree(s->pm_part);
    if (s->mg) stzmg_destroy(s->mg);
    delete s;
    return 0;
}

const char *stz_last_error(stz_t s) { return s ? s->err.c_str() : g_stz_err.c_str(); }

stzmg_t stz_multigrid(stz_t s) { return s ? s->mg : nullptr; }

int stz_set_state(stz_t s, double t, const double *u, const double *sig, const double *chi) {
    if (!s || !u || !sig || !chi) return STZMG_ERR_ARG;
    const Grid &g = s->g;
    s->t = t;
    SHIP(s, hipMemcpyAsync(s->stage, u, sizeof(double) * 3 * g.Nv, hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(k_aos2soa, dim3(nb(g.Nv)), dim3(kT), 0, s->stream, (const double *)s->stage, s->sig2, g.Nv, 3);
    // plates at time t (set_boundary_conditions).  The shear step sets only
    // the top plate: with one rank in z the bottom branch of its
    // `pz == max_pz ... else if (pz == min_pz)` chain is never taken
    // (src/shear_sim_3d.cpp:2393, :2468), so the bottom plane keeps the
    // given velocity until the first projection; the transformed step sets
    // both walls to rest (src/trans_sim_3d.cpp:1138-1179).
    hipLaunchKernelGGL(k_copy_vel, dim3(nb(g.Nv)), dim3(kT), 0, s->stream, g, (const double *)s->sig2, top_u(s, t),
                       s->vel, s->p.transform != STZ_T_NONE ? 1 : 0);
    SHIP(s, hipMemcpyAsync(s->stage, sig, sizeof(double) * 6 * g.Nc, hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(k_aos2soa, dim3(nb(g.Nc)), dim3(kT), 0, s->stream, (const double *)s->stage, s->sig, g.Nc, 6);
    SHIP(s, hipMemcpyAsync(s->chi, chi, sizeof(double) * g.Nc, hipMemcpyHostToDevice, s->stream));
    SHIP(s, hipMemsetAsync(s->dpl, 0, sizeof(double) * g.Nc, s->stream));
    SHIP(s, hipGetLastError());
    SHIP(s, hipStreamSynchronize(s->stream));
    return 0;
}

int stz_get_state(stz_t s, double *t, double *u, double *sig, double *chi, double *dpl) {
    if (!s) return STZMG_ERR_ARG;
    const Grid &g = s->g;
    if (t) *t = s->t;
    if (u) {
        hipLaunchKernelGGL(k_soa2aos, dim3(nb(g.Nv)), dim3(kT), 0, s->stream, (const double *)s->vel, s->stage, g.Nv, 3);
        SHIP(s, hipMemcpyAsync(u, s->stage, sizeof(double) * 3 * g.Nv, hipMemcpyDeviceToHost, s->stream));
        SHIP(s, hipStreamSynchronize(s->stream));
    }
    if (sig) {
        hipLaunchKernelGGL(k_soa2aos, dim3(nb(g.Nc)), dim3(kT), 0, s->stream, (const double *)s->sig, s->stage, g.Nc, 6);
        SHIP(s, hipMemcpyAsync(sig, s->stage, sizeof(double) * 6 * g.Nc, hipMemcpyDeviceToHost, s->stream));
        SHIP(s, hipStreamSynchronize(s->stream));
    }
    if (chi) SHIP(s, hipMemcpyAsync(chi, s->chi, sizeof(double) * g.Nc, hipMemcpyDeviceToHost, s->stream));
    if (dpl) SHIP(s, hipMemcpyAsync(dpl, s->dpl, sizeof(double) * g.Nc, hipMemcpyDeviceToHost, s->stream));
    SHIP(s, hipStreamSynchronize(s->stream));
    return 0;
}

}  // extern "C"

namespace stz_detail {

bool step_options(stz_t s, const stzmg_solve_opts *opts, stzmg_solve_opts &o) {
    if (opts && !(opts->mode == STZMG_SOLVE_REF && !(opts->tol > 0))) {
        o = *opts;
        return false;
    }
    // the reference's loop; with opts (mode STZMG_SOLVE_REF, tol 0) its
    // use_graph is kept
    stzmg_solve_opts_default(&o);
    if (opts) o.use_graph = opts->use_graph;
    o.mode = STZMG_SOLVE_REF;
    // tol = (eps * stencil[13].a11 * 1e3)^2, src/shear_sim_3d.cpp:551-555
    const double tol = std::numeric_limits<double>::epsilon() * s->a11 * 1e3;
    o.tol = tol * tol;
    o.extra_cycles = 2;
    // shear: norm after every cycle; trans: only on the std_iters schedule
    // (src/trans_sim_3d.cpp:1020-1040)
    o.ref_norm_schedule = s->p.transform != STZ_T_NONE ? 1 : 0;
    o.max_cycles = 100000;
    return true;
}

template <bool TRANS, bool PZ>
static void advect_variant(stz_t s, double dt, const XformD &X, const Box &box) {
    const Grid &g = s->g;
    int &r = s->adv_resident[(TRANS ? 2 : 0) + (PZ ? 1 : 0)];
    if (r <= 0) {
        // workgroups the device keeps resident at once: one round of the launch
        int dev = 0, ncu = 0, per = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
            ncu = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_advect_march<TRANS, PZ>, kT, 0) != hipSuccess ||
            per < 1)
            per = 1;
        r = ncu * per;
    }
    const AdvTiles at = adv_tiles(g, r);
    hipLaunchKernelGGL((k_advect_march<TRANS, PZ>), dim3((unsigned)((long)at.ntx * at.nty * at.nzc)), dim3(kT), 0,
                       s->stream, g, s->M, dt, X, box, at, (const double *)s->vel, (const double *)s->sig,
                       (const double *)s->chi, (const double *)s->adapt, (const double *)s->dchi,
                       (const double *)s->dpl2, (const double *)s->dpl, s->sig2, s->chi2);
}

// Pass 2 of the step: k_advect_march for the plain or transformed step,
// periodic or walled in z.
static void launch_advect(stz_t s, bool trans, double dt, const XformD &X, const Box &box) {
    if (trans) {
        if (s->g.pz) advect_variant<true, true>(s, dt, X, box);
        else advect_variant<true, false>(s, dt, X, box);
    } else {
        if (s->g.pz) advect_variant<false, true>(s, dt, X, box);
        else advect_variant<false, false>(s, dt, X, box);
    }
}

int step_once(stz_t s, double dt, stzmg_solve_opts &o, bool ref_tol, stz_step_stats &S, double *sigstar) {
    const bool trans = s->p.transform != STZ_T_NONE;
    const Grid &g = s->g;
    const Box box{s->p.ax, s->p.ay, s->p.az, s->p.bx, s->p.by, s->p.bz};
    SHIP(s, hipMemsetAsync(s->counters, 0, sizeof(int) * 4, s->stream));
    if (trans) {
        const Xform Xn = xform_at(s, s->t);
        hipLaunchKernelGGL(k_plastic_trans, dim3(nb(g.Nc)), dim3(kT), 0, s->stream, g, s->M, dt, Xn,
                           (const double *)s->sig, (const double *)s->chi, s->adapt, s->dchi, s->dpl2,
                           s->counters);
        launch_advect(s, true, dt, xform_products(Xn), box);
    } else {
        hipLaunchKernelGGL(k_plastic, dim3(nb(g.Nc)), dim3(kT), 0, s->stream, g, s->M, dt,
                           (const double *)s->sig, (const double *)s->chi, s->adapt, s->dchi, s->dpl2,
                           s->counters);
        launch_advect(s, false, dt, XformD{}, box);
    }
    SHIP(s, hipGetLastError());
    std::swap(s->sig, s->sig2);
    std::swap(s->chi, s->chi2);
    if (sigstar)
        SHIP(s, hipMemcpyAsync(sigstar, s->sig, sizeof(double) * 6 * g.Nc, hipMemcpyDeviceToDevice, s->stream));
    std::swap(s->dpl, s->dpl2);
    s->t += dt;
    const double ut = top_u(s, s->t);
    Xform X1{};
    if (trans) {
        X1 = xform_at(s, s->t);
        if (int rc = trans_operator(s, dt, X1)) return rc;
        if (ref_tol) {
            const double tol = std::numeric_limits<double>::epsilon() * s->a11 * 1e3;
            o.tol = tol * tol;
        }
        hipLaunchKernelGGL(k_rhs_trans, dim3(nb(g.Nv)), dim3(kT), 0, s->stream, g, X1, (const double *)s->sig,
                           (const double *)s->vel, stzmg_internal::x0(s->mg), stzmg_internal::r0(s->mg));
    } else {
        hipLaunchKernelGGL(k_rhs, dim3(nb(g.Nv)), dim3(kT), 0, s->stream, g, (const double *)s->sig,
                           (const double *)s->vel, s->fix_a11, ut, stzmg_internal::x0(s->mg),
                           stzmg_internal::r0(s->mg));
    }
    SHIP(s, hipGetLastError());
    stzmg_solve_stats ms{};
    int rc = stzmg_solve(s->mg, &o, &ms);
    S.vcycles += ms.cycles;
    S.solve_ms += ms.ms;
    S.last_res = ms.res;
    if (rc) return sfail(s, rc, "projection solve at t=%g: %s", s->t, stzmg_last_error(s->mg));
    hipLaunchKernelGGL(k_copy_vel, dim3(nb(g.Nv)), dim3(kT), 0, s->stream, g,
                       (const double *)stzmg_internal::x0(s->mg), ut, s->vel);
    if (trans)
        hipLaunchKernelGGL(k_stress_update_trans, dim3(nb(g.Nc)), dim3(kT), 0, s->stream, g, s->M, dt, X1,
                           box, (const double *)s->vel, s->sig);
    else {
        hipLaunchKernelGGL(k_stress_update, dim3(nb(g.Nc)), dim3(kT), 0, s->stream, g, s->M, dt,
                           (const double *)s->vel, s->sig, s->pm_part);
        hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kT), 0, s->stream, (const double *)s->pm_part, (long)nb(g.Nc),
                           s->pm_part + nb(g.Nc));
    }
    SHIP(s, hipGetLastError());
    int cnt[4];
    double pm = 0;
    SHIP(s, hipMemcpyAsync(cnt, s->counters, sizeof(cnt), hipMemcpyDeviceToHost, s->stream));
    if (!trans) SHIP(s, hipMemcpyAsync(&pm, s->pm_part + nb(g.Nc), sizeof(double), hipMemcpyDeviceToHost, s->stream));
    SHIP(s, hipStreamSynchronize(s->stream));
    s->proj_mag = pm;
    S.yielded_cells = cnt[0];
    S.steps++;
    return step_output(s);
}

}  // namespace stz_detail

extern "C" {

int stz_step_qs(stz_t s, double dt, int nsteps, const stzmg_solve_opts *opts, stz_step_stats *st) {
    if (!s || !(dt > 0) || nsteps < 0) return STZMG_ERR_ARG;
    const bool trans = s->p.transform != STZ_T_NONE;
    if (!trans)
        if (int rc = ensure_operator(s, dt)) return rc;
    stz_step_stats S{};
    auto t0 = std::chrono::steady_clock::now();
    stzmg_solve_opts o;
    // the reference's tolerance follows the transformed operator, which is
    // rebuilt every step (the untransformed one is fixed by dt)
    const bool ref_loop = stz_detail::step_options(s, opts, o);
    for (int q = 0; q < nsteps; q++)
        if (int rc = stz_detail::step_once(s, dt, o, trans && ref_loop, S, nullptr)) return rc;
    S.step_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (st) *st = S;
    return 0;
}

}  // extern "C"
