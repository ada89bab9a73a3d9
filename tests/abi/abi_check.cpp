// C++ program compiled against include/stzmg.h and include/stz.h (test
// infrastructure): the binding a reference-side C++ caller would compile.
//
//   abi_check layout   prints sizeof and offsetof of every struct of the two
//                      headers as JSON (tests/test_abi.py compares them with
//                      the ctypes mirror in stzpp_amd/_lib.py; no GPU work)
//   abi_check run      one projection solve (32 x 32 x 33 shear operator,
//                      GMRES-IR to 1e-10) and one QS step (16 x 16 x 8) on the
//                      current GPU, plus the error path of a bad descriptor;
//                      exit status 0 when every check holds
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <vector>

#include "stz.h"
#include "stzmg.h"

#define FIELD(S, F) std::printf("    \"%s\": [%zu, %zu],\n", #F, offsetof(S, F), sizeof(((S *)nullptr)->F))
#define BEGIN(S) std::printf("  \"%s\": {\n    \"__size__\": [0, %zu],\n", #S, sizeof(S))
#define END(last) std::printf("  }%s\n", last ? "" : ",")

static int layout() {
    std::printf("{\n");
    BEGIN(stzmg_desc);
    FIELD(stzmg_desc, m); FIELD(stzmg_desc, n); FIELD(stzmg_desc, o); FIELD(stzmg_desc, prd);
    FIELD(stzmg_desc, coarsest_total); FIELD(stzmg_desc, v_down); FIELD(stzmg_desc, v_up);
    FIELD(stzmg_desc, bottom_sweeps); FIELD(stzmg_desc, sweep_growth);
    std::printf("    \"op_mode\": [%zu, %zu]\n", offsetof(stzmg_desc, op_mode), sizeof(int));
    END(false);
    BEGIN(stzmg_solve_opts);
    FIELD(stzmg_solve_opts, mode); FIELD(stzmg_solve_opts, tol); FIELD(stzmg_solve_opts, extra_cycles);
    FIELD(stzmg_solve_opts, ref_norm_schedule); FIELD(stzmg_solve_opts, rtol); FIELD(stzmg_solve_opts, max_cycles);
    FIELD(stzmg_solve_opts, inner_cycles); FIELD(stzmg_solve_opts, use_graph); FIELD(stzmg_solve_opts, restart);
    std::printf("    \"inner_rtol\": [%zu, %zu]\n", offsetof(stzmg_solve_opts, inner_rtol), sizeof(double));
    END(false);
    BEGIN(stzmg_solve_stats);
    FIELD(stzmg_solve_stats, cycles); FIELD(stzmg_solve_stats, outer); FIELD(stzmg_solve_stats, norms);
    FIELD(stzmg_solve_stats, res0); FIELD(stzmg_solve_stats, res);
    std::printf("    \"ms\": [%zu, %zu]\n", offsetof(stzmg_solve_stats, ms), sizeof(double));
    END(false);
    BEGIN(stz_params);
    FIELD(stz_params, Nx); FIELD(stz_params, Ny); FIELD(stz_params, Nz); FIELD(stz_params, prd);
    FIELD(stz_params, ax); FIELD(stz_params, bx); FIELD(stz_params, ay); FIELD(stz_params, by);
    FIELD(stz_params, az); FIELD(stz_params, bz); FIELD(stz_params, mu); FIELD(stz_params, lamb);
    FIELD(stz_params, sy); FIELD(stz_params, tau0); FIELD(stz_params, c0); FIELD(stz_params, eps0);
    FIELD(stz_params, delta); FIELD(stz_params, omega); FIELD(stz_params, bath_temp); FIELD(stz_params, chi_inf);
    FIELD(stz_params, ez); FIELD(stz_params, diff_l); FIELD(stz_params, U); FIELD(stz_params, eta);
    FIELD(stz_params, transform); FIELD(stz_params, Ut); FIELD(stz_params, rho); FIELD(stz_params, kappa);
    std::printf("    \"tmult\": [%zu, %zu]\n", offsetof(stz_params, tmult), sizeof(double));
    END(false);
    BEGIN(stz_step_stats);
    FIELD(stz_step_stats, steps); FIELD(stz_step_stats, vcycles); FIELD(stz_step_stats, yielded_cells);
    FIELD(stz_step_stats, solve_ms); FIELD(stz_step_stats, step_ms);
    std::printf("    \"last_res\": [%zu, %zu]\n", offsetof(stz_step_stats, last_res), sizeof(double));
    END(false);
    BEGIN(stz_adjoint_opts);
    FIELD(stz_adjoint_opts, checkpoint_every);
    std::printf("    \"rtol\": [%zu, %zu]\n", offsetof(stz_adjoint_opts, rtol), sizeof(double));
    END(false);
    BEGIN(stz_grf_params);
    FIELD(stz_grf_params, gauss_fac); FIELD(stz_grf_params, conv_l); FIELD(stz_grf_params, chi_avg);
    FIELD(stz_grf_params, sigma); FIELD(stz_grf_params, TZ); FIELD(stz_grf_params, sim_case);
    std::printf("    \"seed\": [%zu, %zu]\n", offsetof(stz_grf_params, seed), sizeof(unsigned long long));
    END(true);
    std::printf("}\n");
    return 0;
}

static int check(bool ok, const char *what) {
    std::printf("%s %s\n", ok ? "OK  " : "FAIL", what);
    return ok ? 0 : 1;
}

static int run() {
    int bad = 0;
    // error path: an odd periodic size is refused with a message
    stzmg_desc d;
    stzmg_desc_default(&d, 15, 16, 17, 1, 1, 0);
    stzmg_t h = nullptr;
    int rc = stzmg_create(&d, &h);
    bad += check(rc == STZMG_ERR_UNSUPPORTED && h == nullptr && std::strlen(stzmg_last_error(nullptr)) > 0,
                 "stzmg_create refuses 15x16x17 periodic in x");

    // projection solve: shear operator on 32 x 32 x 33 (periodic x, y; walls in z)
    const int n = 32;
    stzmg_desc_default(&d, n, n, n + 1, 1, 1, 0);
    rc = stzmg_create(&d, &h);
    bad += check(rc == 0 && h != nullptr, "stzmg_create 32x32x33");
    if (rc) return 1;
    double B[243], F[9];
    const double hh = 1.0 / n;
    rc = stzmg_shear_stencil(1.0, 1.5, hh, hh, hh, 1.0, 0.0, B, F);
    rc |= stzmg_set_operator_const(h, B, F);
    bad += check(rc == 0 && stzmg_nlevels(h) >= 2, "operator and Galerkin hierarchy");
    const long N = (long)n * n * (n + 1);
    std::vector<double> x(3 * N, 0.0), r(3 * N);
    unsigned long long st = 88172645463325252ull;
    double mean[3] = {0, 0, 0};
    for (long p = 0; p < N; p++)
        for (int c = 0; c < 3; c++) {
            st ^= st << 13; st ^= st >> 7; st ^= st << 17;
            r[3 * p + c] = (double)(st >> 11) / 9007199254740992.0 - 0.5;
            mean[c] += r[3 * p + c] / N;
        }
    for (long p = 0; p < N; p++) {
        const long k = p / ((long)n * n);
        for (int c = 0; c < 3; c++) r[3 * p + c] = (k == 0 || k == n) ? 0.0 : r[3 * p + c] - mean[c];
    }
    rc = stzmg_set_level_fields(h, 0, x.data(), r.data());
    stzmg_solve_opts o;
    stzmg_solve_opts_default(&o);
    o.mode = STZMG_SOLVE_FP32_GMRES_IR;
    o.rtol = 1e-10;
    o.inner_cycles = 1;
    stzmg_solve_stats ss;
    rc |= stzmg_solve(h, &o, &ss);
    bad += check(rc == 0 && ss.res <= 1e-10 * ss.res0 && ss.cycles > 0, "GMRES-IR solve to 1e-10");
    std::vector<double> res(3 * N);
    rc = stzmg_get_level_fields(h, 0, x.data(), nullptr);
    rc |= stzmg_residual(h, 0, res.data());
    double rn = 0, bn = 0;
    for (long q = 0; q < 3 * N; q++) {
        rn += res[q] * res[q];
        bn += r[q] * r[q];
    }
    bad += check(rc == 0 && std::sqrt(rn) <= 1.01e-10 * std::sqrt(bn), "residual of the returned solution");
    stzmg_destroy(h);

    // one quasi-static step of the simulation API
    stz_params p;
    stz_params_default(&p, 16, 16, 8);
    stz_t s = nullptr;
    rc = stz_create(&p, &s);
    bad += check(rc == 0 && s != nullptr, "stz_create 16x16x8");
    if (rc) return 1;
    const long nc = 16L * 16 * 8, nv = 16L * 16 * 9;
    std::vector<double> u(3 * nv, 0.0), sig(6 * nc, 0.0), chi(nc, 0.09), dpl(nc);
    for (long c = 0; c < nc; c++) sig[6 * c + 1] = 0.5;
    rc = stz_set_state(s, 0.3, u.data(), sig.data(), chi.data());
    stz_step_stats sst;
    rc |= stz_step_qs(s, 0.01, 1, nullptr, &sst);
    double t = 0;
    rc |= stz_get_state(s, &t, u.data(), sig.data(), chi.data(), dpl.data());
    bool finite = true;
    for (double v : u) finite = finite && std::isfinite(v);
    for (double v : sig) finite = finite && std::isfinite(v);
    bad += check(rc == 0 && sst.steps == 1 && sst.vcycles > 0 && std::fabs(t - 0.31) < 1e-14 && finite,
                 "one QS step through stz_step_qs");
    bad += check(stz_set_transform(s, B, B) == STZMG_ERR_STATE, "stz_set_transform refused on an untransformed handle");
    stz_destroy(s);
    return bad ? 1 : 0;
}

int main(int argc, char **argv) {
    if (argc > 1 && std::strcmp(argv[1], "layout") == 0) return layout();
    if (argc > 1 && std::strcmp(argv[1], "run") == 0) return run();
    std::fprintf(stderr, "usage: abi_check layout|run\n");
    return 2;
}
