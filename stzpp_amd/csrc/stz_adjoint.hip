// Host side of the adjoint of the quasi-static step (include/stz.h,
// stz_adjoint).  The forward sweep keeps a checkpoint of the state every k
// steps; the reverse sweep restores a checkpoint, recomputes its segment of
// steps keeping each step's input state and advected stress, and reverses
// the steps of the segment last to first with the kernels of stz_adjoint.hpp.
// The adjoint projection solve uses the simulation's multigrid handle: the
// shear operator is symmetric and is reused; the transformed step's operator
// is not (its stencil at T(t) has block(-o) != block(o)^T), so the transposed
// operator of each step is loaded before that step's adjoint solve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/stz.h"
#include "mg_internal.hpp"
#include "stz_adjoint.hpp"
#include "stz_internal.hpp"

using namespace stz;

namespace {

inline unsigned nbk(long n) { return (unsigned)((n + kT - 1) / kT); }

#define AHIP(s, call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return stz_detail::fail((s), STZMG_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

// One device copy of the step state.
struct State {
    double *vel = nullptr, *sig = nullptr, *chi = nullptr, *dpl = nullptr;
    double t = 0;
};

struct Buffers {
    std::vector<double *> all;
    hipError_t err = hipSuccess;
    double *get(long n) {
        double *p = nullptr;
        if (err == hipSuccess) err = hipMalloc((void **)&p, sizeof(double) * (size_t)(n > 0 ? n : 1));
        if (p) all.push_back(p);
        return p;
    }
    ~Buffers() {
        for (double *p : all) hipFree(p);
    }
};

State make_state(Buffers &B, const Grid &g) {
    State st;
    st.vel = B.get(3 * g.Nv);
    st.sig = B.get(6 * g.Nc);
    st.chi = B.get(g.Nc);
    st.dpl = B.get(g.Nc);
    return st;
}

hipError_t copy_state(stz_t s, const State &from_dev_or_null, State *to, bool from_sim) {
    const Grid &g = s->g;
    auto cp = [&](double *d, const double *src, long n) {
        return hipMemcpyAsync(d, src, sizeof(double) * n, hipMemcpyDeviceToDevice, s->stream);
    };
    hipError_t e = hipSuccess;
    if (from_sim) {
        if (e == hipSuccess) e = cp(to->vel, s->vel, 3 * g.Nv);
        if (e == hipSuccess) e = cp(to->sig, s->sig, 6 * g.Nc);
        if (e == hipSuccess) e = cp(to->chi, s->chi, g.Nc);
        if (e == hipSuccess) e = cp(to->dpl, s->dpl, g.Nc);
        to->t = s->t;
    } else {
        if (e == hipSuccess) e = cp(s->vel, from_dev_or_null.vel, 3 * g.Nv);
        if (e == hipSuccess) e = cp(s->sig, from_dev_or_null.sig, 6 * g.Nc);
        if (e == hipSuccess) e = cp(s->chi, from_dev_or_null.chi, g.Nc);
        if (e == hipSuccess) e = cp(s->dpl, from_dev_or_null.dpl, g.Nc);
        s->t = from_dev_or_null.t;
    }
    return e;
}

// The operator the adjoint solve needs is A^T; the shear stencil is
// symmetric (block(-o) = block(o)^T), which is checked here.
bool stencil_symmetric(stz_t s, double dt) {
    double B[243], F[9];
    stzmg_shear_stencil(s->p.mu, s->p.lamb, s->g.dx, s->g.dy, s->g.dz, dt, 0.0, B, F);
    for (int b = 0; b < 27; b++)
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) {
                const double x = B[9 * b + 3 * r + c], y = B[9 * (26 - b) + 3 * c + r];
                if (std::fabs(x - y) > 1e-14 * (std::fabs(x) + std::fabs(y))) return false;
            }
    return true;
}

}  // namespace

extern "C" {

void stz_adjoint_opts_default(stz_adjoint_opts *o) {
    if (!o) return;
    o->checkpoint_every = 5;
    o->rtol = 1e-12;
}

int stz_adjoint(stz_t s, double dt, int nsteps, const stz_adjoint_opts *opts, double *loss, double *grad_chi0,
                double *grad_params) {
    if (!s || !(dt > 0) || nsteps < 1) return STZMG_ERR_ARG;
    if (s->g.pz)
        return stz_detail::fail(s, STZMG_ERR_UNSUPPORTED,
                                "adjoint needs walled z (the periodic projection is singular)");
    stz_adjoint_opts ao;
    stz_adjoint_opts_default(&ao);
    if (opts) ao = *opts;
    if (ao.checkpoint_every < 1 || !(ao.rtol > 0)) return STZMG_ERR_ARG;
    const bool trans = s->p.transform != STZ_T_NONE;
    if (!trans) {
        if (int rc = stz_detail::ensure_operator(s, dt)) return rc;
        if (!stencil_symmetric(s, dt))
            return stz_detail::fail(s, STZMG_ERR_UNSUPPORTED, "projection operator is not symmetric");
    }
    const Grid g = s->g;
    const Box box{s->p.ax, s->p.ay, s->p.az, s->p.bx, s->p.by, s->p.bz};
    const int K = ao.checkpoint_every;
    const int nck = (nsteps + K - 1) / K;

    stzmg_solve_opts so;
    stzmg_solve_opts_default(&so);
    so.mode = STZMG_SOLVE_RELRES;
    so.rtol = ao.rtol;
    so.max_cycles = 10000;

    Buffers B;
    std::vector<State> ck(nck), seg(K);
    for (auto &c : ck) c = make_state(B, g);
    for (auto &c : seg) c = make_state(B, g);
    std::vector<double *> sstar(K);
    for (auto &p : sstar) p = B.get(6 * g.Nc);
    State fin = make_state(B, g);
    double *ub = B.get(3 * g.Nv), *ub2 = B.get(3 * g.Nv), *ubp = B.get(3 * g.Nv), *rhs = B.get(3 * g.Nv);
    double *lam = B.get(3 * g.Nv), *zero = B.get(3 * g.Nv);
    double *sb = B.get(6 * g.Nc), *sb2 = B.get(6 * g.Nc), *ssb = B.get(6 * g.Nc);
    double *xb = B.get(g.Nc), *xb2 = B.get(g.Nc), *db = B.get(g.Nc), *dnb = B.get(g.Nc), *dob = B.get(g.Nc);
    double *ab = B.get(g.Nc), *dchib = B.get(g.Nc), *ta = B.get(g.Nc), *tc = B.get(g.Nc), *td = B.get(g.Nc);
    const unsigned nblk_c = nbk(g.Nc);
    double *part = B.get(3L * nblk_c);
    if (B.err != hipSuccess) return stz_detail::fail(s, STZMG_ERR_NOMEM, "adjoint buffers: %s", hipGetErrorString(B.err));
    AHIP(s, hipMemsetAsync(zero, 0, sizeof(double) * 3 * g.Nv, s->stream));

    // forward sweep with checkpoints
    stz_step_stats S{};
    for (int q = 0; q < nsteps; q++) {
        if (q % K == 0) AHIP(s, copy_state(s, State{}, &ck[q / K], true));
        if (int rc = stz_detail::step_once(s, dt, so, false, S, nullptr)) return rc;
    }
    AHIP(s, copy_state(s, State{}, &fin, true));
    std::vector<double> hd(g.Nc);
    AHIP(s, hipMemcpyAsync(hd.data(), s->dpl, sizeof(double) * g.Nc, hipMemcpyDeviceToHost, s->stream));
    AHIP(s, hipStreamSynchronize(s->stream));
    double J = 0;
    for (long c = 0; c < g.Nc; c++) J += (hd[c] / dt) * (hd[c] / dt);
    J /= (double)g.Nc;
    if (loss) *loss = J;

    // reverse sweep
    hipLaunchKernelGGL(k_adj_loss, dim3(nblk_c), dim3(kT), 0, s->stream, g, dt, (const double *)s->dpl, db);
    AHIP(s, hipMemsetAsync(ub, 0, sizeof(double) * 3 * g.Nv, s->stream));
    AHIP(s, hipMemsetAsync(sb, 0, sizeof(double) * 6 * g.Nc, s->stream));
    AHIP(s, hipMemsetAsync(xb, 0, sizeof(double) * g.Nc, s->stream));
    double pg[3] = {0, 0, 0};
    std::vector<double> hp(3L * nblk_c);
    for (int sg = nck - 1; sg >= 0; sg--) {
        const int s0 = sg * K, s1 = std::min(s0 + K, nsteps);
        AHIP(s, copy_state(s, ck[sg], nullptr, false));
        for (int q = s0; q < s1; q++) {
            AHIP(s, copy_state(s, State{}, &seg[q - s0], true));
            if (int rc = stz_detail::step_once(s, dt, so, false, S, sstar[q - s0])) return rc;
        }
        for (int q = s1 - 1; q >= s0; q--) {
            const State &pre = seg[q - s0];
            // transforms at the advection (t) and at the projection (t + dt)
            const double t1 = pre.t + dt;
            const Xform Xn = trans ? stz_detail::xform(s, pre.t) : Xform{};
            const Xform X1 = trans ? stz_detail::xform(s, t1) : Xform{};
            // sigma' = sigma* + dt C:D(u')  (trans: dt T^-1 C:D(L) T^-t)
            AHIP(s, hipMemcpyAsync(ubp, ub, sizeof(double) * 3 * g.Nv, hipMemcpyDeviceToDevice, s->stream));
            if (trans)
                hipLaunchKernelGGL(k_adj_stress_trans, dim3(nblk_c), dim3(kT), 0, s->stream, g, s->M, dt, X1,
                                   (const double *)sb, ubp);
            else
                hipLaunchKernelGGL(k_adj_stress, dim3(nblk_c), dim3(kT), 0, s->stream, g, s->M, dt,
                                   (const double *)sb, ubp);
            // u' = A^-1 r on the interior rows
            hipLaunchKernelGGL(k_adj_wall_zero, dim3(nbk(g.Nv)), dim3(kT), 0, s->stream, g, (const double *)ubp,
                               rhs);
            AHIP(s, hipGetLastError());
            if (trans)
                if (int rc = stz_detail::transposed_trans_operator(s, dt, t1)) return rc;
            if (int rc = stzmg_set_fields_device(s->mg, zero, rhs))
                return stz_detail::fail(s, rc, "adjoint solve: %s", stzmg_last_error(s->mg));
            stzmg_solve_stats ms{};
            if (int rc = stzmg_solve(s->mg, &so, &ms))
                return stz_detail::fail(s, rc, "adjoint solve: %s", stzmg_last_error(s->mg));
            if (int rc = stzmg_get_x_device(s->mg, lam))
                return stz_detail::fail(s, rc, "adjoint solve: %s", stzmg_last_error(s->mg));
            // r = div sigma*  (trans: T div sigma*')
            if (trans)
                hipLaunchKernelGGL(k_adj_div_trans, dim3(nblk_c), dim3(kT), 0, s->stream, g, X1,
                                   (const double *)lam, (const double *)sb, ssb);
            else
                hipLaunchKernelGGL(k_adj_div, dim3(nblk_c), dim3(kT), 0, s->stream, g, (const double *)lam,
                                   (const double *)sb, ssb);
            // this step's plastic quantities from its input state
            if (trans)
                hipLaunchKernelGGL(k_plastic_trans, dim3(nblk_c), dim3(kT), 0, s->stream, g, s->M, dt, Xn,
                                   (const double *)pre.sig, (const double *)pre.chi, ta, tc, td, s->counters);
            else
                hipLaunchKernelGGL(k_plastic, dim3(nblk_c), dim3(kT), 0, s->stream, g, s->M, dt,
                                   (const double *)pre.sig, (const double *)pre.chi, ta, tc, td, s->counters);
            // advection
            AHIP(s, hipMemcpyAsync(sb2, ssb, sizeof(double) * 6 * g.Nc, hipMemcpyDeviceToDevice, s->stream));
            AHIP(s, hipMemcpyAsync(xb2, xb, sizeof(double) * g.Nc, hipMemcpyDeviceToDevice, s->stream));
            AHIP(s, hipMemcpyAsync(dnb, db, sizeof(double) * g.Nc, hipMemcpyDeviceToDevice, s->stream));
            AHIP(s, hipMemsetAsync(dob, 0, sizeof(double) * g.Nc, s->stream));
            AHIP(s, hipMemsetAsync(ub2, 0, sizeof(double) * 3 * g.Nv, s->stream));
            if (trans)
                hipLaunchKernelGGL(k_adj_advect_trans, dim3(nblk_c), dim3(kT), 0, s->stream, g, s->M, dt, Xn, box,
                                   (const double *)pre.vel, (const double *)pre.sig, (const double *)pre.chi,
                                   (const double *)ta, (const double *)td, (const double *)pre.dpl,
                                   (const double *)ssb, (const double *)xb, sb2, xb2, ub2, dnb, dob, ab, dchib);
            else
                hipLaunchKernelGGL(k_adj_advect, dim3(nblk_c), dim3(kT), 0, s->stream, g, s->M, dt,
                                   (const double *)pre.vel, (const double *)pre.sig, (const double *)pre.chi,
                                   (const double *)ta, (const double *)td, (const double *)pre.dpl,
                                   (const double *)ssb, (const double *)xb, sb2, xb2, ub2, dnb, dob, ab, dchib);
            // plasticity
            if (trans)
                hipLaunchKernelGGL(k_adj_plastic_trans, dim3(nblk_c), dim3(kT), 0, s->stream, g, s->M, dt, Xn,
                                   (const double *)pre.sig, (const double *)pre.chi, (const double *)ab,
                                   (const double *)dchib, (const double *)dnb, sb2, xb2, part);
            else
                hipLaunchKernelGGL(k_adj_plastic, dim3(nblk_c), dim3(kT), 0, s->stream, g, s->M, dt,
                                   (const double *)pre.sig, (const double *)pre.chi, (const double *)ab,
                                   (const double *)dchib, (const double *)dnb, sb2, xb2, part);
            AHIP(s, hipGetLastError());
            AHIP(s, hipMemcpyAsync(hp.data(), part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, s->stream));
            AHIP(s, hipStreamSynchronize(s->stream));
            for (unsigned b = 0; b < nblk_c; b++)
                for (int q3 = 0; q3 < 3; q3++) pg[q3] += hp[3 * b + q3];
            std::swap(ub, ub2);
            std::swap(sb, sb2);
            std::swap(xb, xb2);
            std::swap(db, dob);
        }
    }
    if (grad_chi0)
        AHIP(s, hipMemcpyAsync(grad_chi0, xb, sizeof(double) * g.Nc, hipMemcpyDeviceToHost, s->stream));
    if (grad_params)
        for (int q3 = 0; q3 < 3; q3++) grad_params[q3] = pg[q3];
    AHIP(s, copy_state(s, fin, nullptr, false));
    AHIP(s, hipStreamSynchronize(s->stream));
    return 0;
}

}  // extern "C"
