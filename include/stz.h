/* stz.h -- C ABI of the MI355X quasi-static STZ step.
 *
 * Drop-in for the simulation step API of shear_sim_3d<sym_mat3>
 * (src/shear_sim_3d.hh:35-82) and trans_sim_3d (src/trans_sim_3d.hh:13-36):
 * state in, N quasi-static steps (step_forward_qs, src/shear_sim_3d.cpp:
 * 381-389, src/trans_sim_3d.cpp:72-80), state out.  One subdomain on one
 * GPU; periodic in x and y, walls or periodic in z.  In the transformed step
 * u and sig are the transformed velocity U and stress sigma'.
 *
 * State arrays on the host (AoS, x fastest):
 *   u    vertex velocity, 3 doubles per vertex, Nx*Ny*Nzv vertices with
 *        Nzv = Nz+1 when z is walled (the top plate plane included), Nz when
 *        periodic -- the multigrid grid of region.hh:805
 *   sig  cell stress s11 s12 s13 s22 s23 s33, 6 doubles per cell, Nx*Ny*Nz
 *   chi  cell effective temperature, Nx*Ny*Nz
 *   dpl  cell ad_Dpl (2 mu dt Dpl / sbar scaled as in the reference), Nx*Ny*Nz
 * Status codes are those of stzmg.h.
 */
#ifndef STZ_H
#define STZ_H

#include "stzmg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct stz_s *stz_t;

typedef struct {
    int Nx, Ny, Nz;        /* cells per dimension (gN_x, gN_y, gN_z) */
    int prd[3];            /* periodicity; x and y must be periodic */
    double ax, bx, ay, by, az, bz; /* domain */
    double mu, lamb;       /* elastic moduli */
    double sy, tau0, c0, eps0, delta, omega, bath_temp, chi_inf, ez; /* STZ */
    double diff_l;         /* plasticity-mediated diffusion length */
    double U;              /* plate speed; the plates ramp as U*t for t < 1 */
    double eta;            /* adaptive plastic-step tolerance; 0 = the reference's
                              value for the step: 5e-4 (src/shear_sim_3d.cpp:111),
                              2e-3 transformed (src/trans_sim_3d.cpp:741) */
    int transform;         /* STZ_T_NONE: shear_sim_3d<sym_mat3> step with moving
                              plates; otherwise the transformed-coordinate step of
                              trans_sim_3d with this transform family and walls at rest */
    double Ut;             /* rate of the transform family */
    double rho;            /* density; direct step only (shear_sim driver: rho = mu) */
    double kappa;          /* viscous damping of the direct step (driver: 0.35; the
                              quasi-static driver passes 0, src/shear_sim.cpp:122) */
    double tmult;          /* direct time-step multiplier (src/shear_sim.cpp:53) */
} stz_params;

/* Transform families of trans_sim_3d (src/trans_sim_3d.hh:57-125): pure shear
 * T = diag(e^{Ut t}, 1, e^{-Ut t}) is the one the reference compiles in;
 * simple shear T = I + (Ut t / bz) e_x e_z^T, homogeneous T = e^{Ut t} I and
 * the identity are its commented alternatives. */
enum { STZ_T_NONE = 0, STZ_T_PURE_SHEAR = 1, STZ_T_SIMPLE_SHEAR = 2, STZ_T_HOMOGENEOUS = 3, STZ_T_IDENTITY = 4,
       STZ_T_USER = 5 /* set by stz_set_transform */ };

typedef struct {
    int steps;             /* steps taken */
    int vcycles;           /* V-cycles over all projection solves */
    int yielded_cells;     /* cells with sbar > sy in the last step */
    double solve_ms;       /* wall time in the projection solves */
    double step_ms;        /* wall time of the whole call */
    double last_res;       /* final solve quantity of the last step */
} stz_step_stats;

/* Reference defaults: domain [-1,1)^2 x [-0.5,0.5], x,y periodic, z walled,
 * parameters of the shear_sim driver after nondimensionalisation
 * (src/shear_sim.cpp:18-86). */
int stz_params_default(stz_params *p, int Nx, int Ny, int Nz);
int stz_create(const stz_params *p, stz_t *s);
int stz_destroy(stz_t s);
const char *stz_last_error(stz_t s);
/* Host state in/out; set_state applies the reference's boundary conditions
 * (plate velocities at time t, set_boundary_conditions,
 * src/shear_sim_3d.cpp:2362-2658). */
int stz_set_state(stz_t s, double t, const double *u, const double *sig, const double *chi);
int stz_get_state(stz_t s, double *t, double *u, double *sig, double *chi, double *dpl);
/* nsteps quasi-static steps of size dt.  opts == NULL selects the reference
 * projection loop (sum |r|^4 <= (eps*stencil[13].a11*1e3)^2, then two more
 * cycles, src/shear_sim_3d.cpp:551-600), and so does opts with mode
 * STZMG_SOLVE_REF and tol 0, taking only use_graph from opts; otherwise the
 * given solve mode and options. */
int stz_step_qs(stz_t s, double dt, int nsteps, const stzmg_solve_opts *opts, stz_step_stats *st);
/* The projection solver of the simulation (valid after the first step). */
stzmg_t stz_multigrid(stz_t s);
/* Adjoint of the quasi-static step (no reference counterpart).
 *
 * Loss J = mean over cells of (ad_Dpl / dt)^2 after the last of nsteps
 * steps from the current state.  The forward steps solve each projection
 * to a relative residual (so the discrete model is differentiable; the
 * reference's sum |r|^4 loop is not), keep the state every
 * checkpoint_every steps, and the reverse sweep recomputes one segment of
 * steps from its checkpoint before reversing it.  The reverse projection
 * solves with the transposed operator: the shear operator is symmetric and
 * is reused; the transformed step's operator at T(t) is not, and its
 * transpose is loaded for each step's adjoint solve.
 * Branches (ENO stencils, the yield test, the plastic sub-step count) are
 * differentiated for the branch the forward state selects.
 * Supported: the untransformed and the transformed step (every transform
 * family) with walled z; periodic z is rejected (its projection is
 * singular).  On return the state is the one after the nsteps forward
 * steps. */
typedef struct {
    int checkpoint_every;  /* states kept every k steps (k >= 1) */
    double rtol;           /* relative residual of every projection solve, forward and adjoint */
} stz_adjoint_opts;
void stz_adjoint_opts_default(stz_adjoint_opts *o);
/* grad_chi0: dJ/dchi at the initial state (Nx*Ny*Nz, may be NULL);
 * grad_params: dJ/d(sy, eps0, chi_inf) (3 doubles, may be NULL). */
int stz_adjoint(stz_t s, double dt, int nsteps, const stz_adjoint_opts *o, double *loss, double *grad_chi0,
                double *grad_params);

/* T, T^-1 and dT/dt (row-major 3x3) of a transform family at time t
 * (compute_T, compute_Tinv, compute_dTdt, src/trans_sim_3d.hh:57-125). */
int stz_transform(int family, double Ut, double bz, double t, double *T, double *Tinv, double *dTdt);
/* An arbitrary transform for a transformed-coordinate handle (one created
 * with a transform family), in place of the families of
 * src/trans_sim_3d.hh:57-140: from the handle's current time t0 on,
 * T(t) = T + (t - t0) dTdt with dT/dt = dTdt (row-major 3x3), T^-1 from
 * cofactors.  The step uses T(t) for its advection and T(t + dt) for its
 * projection, so a transform affine in time (simple shear, any constant
 * velocity gradient) is exact, and a caller that sets T(t) and dT/dt(t)
 * before every step gets T(t + dt) to first order in dt.  As in the
 * reference's generated stencil (src/stenc_funcs.hh), the projection
 * operator reads the upper triangle of T only.  STZMG_ERR_STATE for an
 * untransformed handle, STZMG_ERR_SINGULAR for det T = 0. */
int stz_set_transform(stz_t s, const double *T, const double *dTdt);

/* ---- random initial chi (Gaussian random field) -------------------------
 *
 * chi = chi_avg + chi_1 (noise * G) on the Nx x Ny x Nz cells (x fastest),
 * G(i,j,k) = exp(-(i^2 + j^2 + k^2) / conv_l) on the cube |i|,|j|,|k| <= cut,
 * cut = (int)((int)gauss_fac * conv_l), chi_1 = sigma / sqrt(sum G^2)
 * (set_random_initial_conditions, convolve, gauss_normal_factor_long,
 * src/shear_sim_3d.cpp:2793-3078; the int members gauss_fac and cut,
 * src/shear_sim_3d.hh:140-142).  The noise grid is nx x ny x nz with
 * n = N on a periodic axis and N + 2 cut otherwise (x fastest); periodic
 * axes wrap once, so cut must not exceed a periodic N.  The convolution
 * runs on the GPU as three 1-D passes.  When the caller passes no noise it
 * is drawn on the GPU from `seed` (Philox4x64-10 and the polar Box-Muller
 * method); the reference's own noise comes from unseeded rand(). */
typedef struct {
    double gauss_fac;        /* cutoff in correlation lengths (truncated to an int) */
    double conv_l;           /* correlation length in cells */
    double chi_avg;          /* mean of chi */
    double sigma;            /* standard deviation of chi */
    double TZ;               /* temperature scale: enters only the cache file name */
    int sim_case;            /* enters only the cache file name */
    unsigned long long seed; /* noise stream when no noise grid is given */
} stz_grf_params;
/* The shear_sim driver's values (src/shear_sim.cpp:8, :83-86). */
void stz_grf_params_default(stz_grf_params *g);
/* Cutoff, noise-grid size and chi_1 of a configuration (any may be NULL). */
int stz_grf_geometry(const int N[3], const int prd[3], const stz_grf_params *g, int *cut, long long *noise_size,
                     double *chi_1);
/* The reference's cache file name, e.g. "cg_32_32_32_mu550_sig16.dat"
 * (src/shear_sim_3d.cpp:2812-2816). */
int stz_grf_cache_name(const int N[3], const int prd[3], const stz_grf_params *g, char *buf, int len);
/* chi from host noise (NULL: drawn from g->seed) into host memory. */
int stz_grf_field(const int N[3], const int prd[3], const stz_grf_params *g, const double *noise, double *chi);
/* Same on device buffers and a caller's HIP stream (synchronised on return). */
int stz_grf_field_device(const int N[3], const int prd[3], const stz_grf_params *g, const double *noise_dev,
                         double *chi_dev, void *stream);
/* Average device time per launch (ms) of the four kernels -- noise draw,
 * x pass, y pass, z pass -- over reps back-to-back launches each, from HIP
 * events on the library's stream (for measurement). */
int stz_grf_time_kernels(const int N[3], const int prd[3], const stz_grf_params *g, int reps, double ms[4]);
/* Message of the last failing stz_grf_* call of this thread. */
const char *stz_grf_last_error(void);
/* set_random_initial_conditions (src/shear_sim_3d.cpp:2793-2934) on the
 * simulation's chi: with cache_dir, an existing cache file there is loaded
 * instead, and a computed field is written to it (raw FP64, x fastest). */
int stz_set_random_initial_conditions(stz_t s, const stz_grf_params *g, const double *noise, const char *cache_dir);

/* ---- output in the reference's file formats ------------------------------
 *
 * Cube modes are those of Field::fval (src/field.hh:158-178) and write_files
 * (src/shear_sim_3d.cpp:1938-1960): 0-2 u v w, 3-8 s11 s12 s13 s22 s23 s33,
 * 9 p, 10 dev, 11 chi, 12 tem (chi * 21000), 13-15 xi_x xi_y xi_z (the
 * reference map, kept by the direct step only; rejected without one).  A
 * cube is float32, x slowest, then y, z fastest (output_cube, :1963-2161):
 * Nx*Ny*(Nz+1) values for the velocity and the reference map (the top plate
 * plane; the periodic image of plane 0 when z is periodic, xi_z shifted by
 * bz - az), Nx*Ny*Nz otherwise. */
long long stz_cube_size(stz_t s, int mode);   /* -1 for an unknown mode */
int stz_output_cube(stz_t s, int mode, float *out, long long n);
/* Average device time per launch (ms) of the cube kernel of `mode` over
 * reps back-to-back launches into the handle's staging buffer, from HIP
 * events on its stream (for measurement). */
int stz_output_time_kernel(stz_t s, int mode, int reps, double *ms);
/* Appends the cube to <dir>/<name>_grid.<frame> (name: u v w s11 ... tem),
 * creating dir (one level) as the reference does. */
int stz_write_cube(stz_t s, const char *dir, int mode, int frame);
/* write_files(frame) (src/shear_sim_3d.cpp:1938-1960): the modes whose bit
 * is set in fflags (the reference's build uses 4096, tem only), the velocity
 * and stress every n_tp/10 frames, chi when (n_tp/10) % frame == 0 for
 * frame > 0.  n_tp must be at least 10 (the reference divides by n_tp/10). */
int stz_write_files(stz_t s, const char *dir, int frame, int n_tp, int fflags);
/* Wall tractions (compute_net_force, src/shear_sim_3d.cpp:1061-1126,
 * src/trans_sim_3d.cpp:1575-1660): out = top (s13, s23, s33) * dx dy and
 * bottom (s13, s23, s33) * (-dx dy) of the stress extrapolated to the walls,
 * top 1.5 s(Nz-1) - 0.5 s(Nz-2), bottom 1.5 s(x) - 0.5 s(x+1) (the
 * reference pairs each bottom cell with its +x neighbour); the transformed
 * step reports the physical stress T s T^t at the current time. */
int stz_net_force(stz_t s, double out[6]);
/* Appends the traction line to <dir>/traction.dat ("%.16f" x 7 with the
 * time first; transformed: "%g" x 6) and, untransformed, the six doubles to
 * <dir>/trac_grid.bin. */
int stz_write_traction(stz_t s, const char *dir);
/* Sum of sym_mat3::modsq of the last untransformed projection's stress
 * increments (proj_mag, src/shear_sim_3d.cpp:699-724). */
double stz_proj_mag(stz_t s);
/* Per-step output during stz_step_qs, as solve_qs and projection_step write
 * it (src/shear_sim_3d.cpp:336-377, :714-724): STZ_OUT_TRACTION appends the
 * traction line after every step, STZ_OUT_PROJ_MAG one "%g %g" line (time,
 * proj_mag) to proj_mag.dat per untransformed projection (the first line of a
 * handle truncates the file).  flags = 0 turns it off. */
enum { STZ_OUT_TRACTION = 1, STZ_OUT_PROJ_MAG = 2 };
int stz_set_output(stz_t s, const char *dir, int flags);

/* ---- direct (explicit dynamics) step -------------------------------------
 *
 * shear_sim_3d::step_forward (src/shear_sim_3d.cpp:206-213) of the
 * untransformed simulation: forward Euler of the hypo-elastoplastic
 * equations with density rho and viscous damping kappa (stz_params).  Per
 * step every cell takes the elastic, Truesdell, ENO2-advective and STZ
 * plastic stress change and the advected, plastic and diffused chi change
 * (step_stress, step_chi, :774-896, :1193-1215); every vertex 0 < k < Nz
 * takes dt/rho div sigma + dt kappa/rho Lap u - u.grad u (step_velocity,
 * :919-994) and the advected reference map (step_ref, :997-1030); then the
 * top plate moves to U t (U for t >= 1) with reference map
 * (a_x + i dx + U t^2/2 or + U (t - 1/2), a_y + j dy, 0).  As in the
 * reference with one rank, the vertex plane k = 0 is never updated (the
 * skip of zz = 0, :101, also with z periodic) and the bottom plate is not
 * set (:2393, :2468).  Beyond a wall the reference map reads zero (an
 * extrapolated Field carries only its ten primary members,
 * src/field.hh:49-112); periodic images of it are shifted by the domain
 * length.  The reference map is allocated by the first call of
 * stz_init_ref, stz_set_ref_map or stz_step_direct (zero when not set). */
/* dt = min(tmult dx^2 / (6 kappa), dx / 4) (solve_direct, :166-172). */
double stz_direct_dt(stz_t s);
/* init_ref (:1036-1057): the reference map at the vertex positions below
 * the top plate, and the plate's map at the current time. */
int stz_init_ref(stz_t s);
/* Reference map in/out: 3 doubles per vertex, Nx*Ny*(Nz+1) vertices
 * (Nx*Ny*Nz when z is periodic), x fastest. */
int stz_set_ref_map(stz_t s, const double *xi);
int stz_get_ref_map(stz_t s, double *xi);
/* nsteps direct steps of size dt (stats: steps and yielded_cells). */
int stz_step_direct(stz_t s, double dt, int nsteps, stz_step_stats *st);
/* Average device time per launch (ms) of the step's four kernels -- cell
 * plasticity, cell stress and chi, vertex velocity and reference map, top
 * plate -- over reps back-to-back launches each on the current state, from
 * HIP events on the handle's stream (for measurement; the state is left as
 * it was). */
int stz_direct_time_kernels(stz_t s, double dt, int reps, double ms[4]);
/* l2_comparison (src/shear_sim_3d.cpp:3122-3158), the comparison of a
 * direct against a quasi-static simulation in shear_compare
 * (src/shear_compare.cpp:196-205): l2 = the squared differences of a and b
 * (same grid) in velocity (trapezoid rule in z: vertex plane 0 and plane Nz
 * weigh 1/2), stress (six components) and chi, each times
 * dx dy dz / volume, the velocity's divided by U^2 and chi's by chi_inf^2
 * (a's parameters).  The reference's accumulators start from an
 * indeterminate value (:3127); these start from zero. */
int stz_l2_comparison(stz_t a, stz_t b, double l2[3]);

#ifdef __cplusplus
}
#endif
#endif
