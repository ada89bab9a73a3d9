// Handle of the quasi-static step (include/stz.h) and the pieces of the
// step shared by the forward driver (stz.hip) and the adjoint (stz_adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/stz.h"
#include "stz_kernels.hpp"

struct stz_s {
    stz_params p{};
    stz::Grid g{};
    stz::Mat M{};
    double t = 0;
    stzmg_t mg = nullptr;
    double mg_dt = -1;          // dt the operator was built for
    double a11 = 0;             // stencil[13].a11 of the current operator
    double fix_a11 = 0;
    hipStream_t stream = nullptr;
    double *vel = nullptr, *sig = nullptr, *sig2 = nullptr, *chi = nullptr, *chi2 = nullptr;
    double *dpl = nullptr, *dpl2 = nullptr, *adapt = nullptr, *dchi = nullptr, *stage = nullptr;
    int *counters = nullptr;
    // direct (explicit dynamics) step: reference map and its double buffer,
    // second velocity buffer; allocated by the first direct-step call
    double *xi = nullptr, *xi2 = nullptr, *vel2 = nullptr;
    double *pm_part = nullptr;  // proj_mag: per-block partial sums, then the total
    double proj_mag = 0;        // sum of modsq of the last projection's stress increments
    std::string out_dir;        // per-step output (stz_set_output)
    int out_flags = 0;
    int pm_writes = 0;          // proj_mag.dat lines written by this handle
    double user_T[9] = {0}, user_dT[9] = {0};  // stz_set_transform: T(t) = user_T + (t - user_t0) user_dT
    double user_t0 = 0;
    int adv_resident[4] = {0, 0, 0, 0};  // resident workgroups of the k_advect_march variants
    std::string err;
};

namespace stz_detail {

// Records the message on the handle and returns code.
int fail(stz_t s, int code, const char *fmt, ...);
// T, T^-1 and dT/dt of the handle's transform at time t.
stz::Xform xform(const stz_t s, double t);
// Loads the transpose of the transformed step's projection operator at T(t)
// (the operator of the step that ends at t) into the handle's solver.
int transposed_trans_operator(stz_t s, double dt, double t);
// Projection operator of the untransformed step for step size dt.
int ensure_operator(stz_t s, double dt);
// Solve options of a step call: opts, or the reference's loop when opts is
// null or asks for it (mode STZMG_SOLVE_REF with tol 0).  Returns true when
// the reference's loop was chosen: its tolerance then follows the operator.
bool step_options(stz_t s, const stzmg_solve_opts *opts, stzmg_solve_opts &o);
// One quasi-static step on the handle's stream.  ref_tol: recompute the
// reference tolerance from the current operator (transformed step).  When
// sigstar is not null it receives the advected stress sigma* (6 Nc doubles,
// device) that enters the projection.
int step_once(stz_t s, double dt, stzmg_solve_opts &o, bool ref_tol, stz_step_stats &S, double *sigstar);
// The per-step output of stz_set_output after a step (stz_output.hip).
int step_output(stz_t s);

}  // namespace stz_detail
