/* stzmg.h -- C ABI of the MI355X (gfx950) multigrid projection solver.
 *
 * Drop-in for the reference's duck-typed multigrid<vec3,M>/region<vec3,M>
 * template surface (mg_3d_template/multigrid.hh:19-113, region.hh:17-656) on a
 * single GPU.  Plain pointers and sizes only; no C++ or torch types.
 *
 * Conventions
 *   point index   p = i + m*(j + n*k), x fastest (as region::x0 / r0)
 *   vec3 arrays   host arrays are AoS, 3 doubles per point (reference vec3)
 *   stencil       27 blocks of 3x3, block b = (di+1) + 3*(dj+1) + 9*(dk+1)
 *                 (reference index 13+i+3j+9k, src/shear_sim_3d.cpp:222),
 *                 each block row-major a11 a12 a13 a21 ... a33 (src/mat3.hh)
 *   status        0 = OK, negative = error; stzmg_last_error() gives the text.
 *                 Nothing aborts the process (the reference calls MPI_Abort,
 *                 mg_3d_template/common.cc:52-55).
 *   threading     one HIP stream per handle; calls are synchronous unless
 *                 documented otherwise; a handle is not thread safe, distinct
 *                 handles are independent (no function-static state, unlike
 *                 the reference's std_iters, src/shear_sim_3d.cpp:561).
 */
#ifndef STZMG_H
#define STZMG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct stzmg_s *stzmg_t;

enum {
    STZMG_OK = 0,
    STZMG_ERR_ARG = -1,
    STZMG_ERR_UNSUPPORTED = -2,
    STZMG_ERR_HIP = -3,
    STZMG_ERR_NOMEM = -4,
    STZMG_ERR_STATE = -5,
    STZMG_ERR_NOT_CONVERGED = -6,
    STZMG_ERR_NONFINITE = -7,
    STZMG_ERR_SINGULAR = -8
};

/* How the operator of a level varies in space. */
enum {
    STZMG_OP_PLANE = 0, /* one 27-block stencil per z plane (x, y periodic) */
    STZMG_OP_POINT = 1  /* one 27-block stencil per point (general) */
};

enum { STZMG_NORM_QUARTIC = 0, STZMG_NORM_L2SQ = 1 };

enum {
    STZMG_SOLVE_REF = 0,     /* reference projection loop (sum |r|^4 <= tol, + extra cycles) */
    STZMG_SOLVE_RELRES = 1,  /* FP64 V-cycles to ||r||_2 <= rtol ||b - A x0||_2 */
    STZMG_SOLVE_FP32_IR = 2, /* FP32 V-cycles on the defect inside FP64 refinement */
    STZMG_SOLVE_FP32_GMRES = 3, /* flexible GMRES(restart) in FP64, right-preconditioned by
                                   inner_cycles FP32 V-cycles from zero (GMRES-based
                                   iterative refinement); stops on the true FP64 residual */
    STZMG_SOLVE_FP32_GMRES_IR = 4, /* iterative refinement in FP64 whose correction is an
                                     FP32 GMRES(restart) solve (FP32 Krylov basis and operator
                                     applies, FP64 dot products), right-preconditioned by
                                     inner_cycles FP32 V-cycles, each run until its residual
                                     estimate fell by inner_rtol; stops on the true FP64
                                     residual */
    STZMG_SOLVE_GMRES64 = 5       /* flexible GMRES(restart) entirely in FP64, right-
                                     preconditioned by inner_cycles FP64 V-cycles from zero
                                     (the FP64 counterpart of STZMG_SOLVE_FP32_GMRES, for
                                     measuring what the FP32 precision alone gains); stops on
                                     the true FP64 residual */
};

typedef struct {
    int m, n, o;          /* fine-level vertices per dimension */
    int prd[3];           /* periodicity flags */
    long coarsest_total;  /* coarsen while m*n*o > this (reference: 1024, multigrid.cc:157) */
    int v_down, v_up;     /* pre/post sweeps on level 0 (reference: 2, 2, multigrid.hh:160) */
    int bottom_sweeps;    /* sweeps on the coarsest level (reference: 24, common.hh:57) */
    int sweep_growth;     /* level l gets v_down+growth*l sweeps (reference: 1, multigrid.cc:265) */
    int op_mode;          /* STZMG_OP_PLANE or STZMG_OP_POINT */
} stzmg_desc;

typedef struct {
    int mode;             /* STZMG_SOLVE_* */
    double tol;           /* REF: threshold on sum |r|^4 (region.cc:3031-3057) */
    int extra_cycles;     /* REF: cycles after convergence (reference: 2, src/shear_sim_3d.cpp:599) */
    int ref_norm_schedule;/* REF: 0 = norm after every cycle (shear_sim_3d.cpp:567-582),
                             1 = only once num_iters >= std_iters/16 (trans_sim_3d.cpp:1020-1040) */
    double rtol;          /* RELRES / FP32_IR: relative residual target */
    int max_cycles;       /* cap on V-cycles (the reference loop has none) */
    int inner_cycles;     /* FP32_IR: FP32 V-cycles per refinement step */
    int use_graph;        /* replay V-cycles from a captured hipGraph */
    int restart;          /* FP32_GMRES(_IR): Krylov vectors per restart (1..16) */
    double inner_rtol;    /* FP32_GMRES_IR: reduction of each inner solve (0: 1e-4) */
} stzmg_solve_opts;

typedef struct {
    int cycles;           /* V-cycles run (FP64 or FP32) */
    int outer;            /* FP64 refinement steps (FP32_IR) or restarts (FP32_GMRES) */
    int norms;            /* residual-norm evaluations */
    double res0;          /* initial ||r||_2 (RELRES/FP32_IR) or sum |r|^4 (REF) */
    double res;           /* final value of the same quantity */
    double ms;            /* wall time of the solve in milliseconds */
} stzmg_solve_stats;

/* Fill *d with the reference's defaults for an m x n x o grid. */
int stzmg_desc_default(stzmg_desc *d, int m, int n, int o, int px, int py, int pz);

/* multigrid<V,M>::multigrid + setup_hierarchy (multigrid.hh:157-189,
 * multigrid.cc:136-249): allocate all levels on the current HIP device. */
int stzmg_create(const stzmg_desc *d, stzmg_t *h);
int stzmg_destroy(stzmg_t h);
const char *stzmg_last_error(stzmg_t h);
int stzmg_nlevels(stzmg_t h);
int stzmg_level_dims(stzmg_t h, int level, int dims[3]);

/* Fine operator, then the Galerkin hierarchy on the GPU in FP64
 * (multigrid::setup_matrices -> compute_rats, multigrid.hh:195-200,
 * multigrid.cc:330-344; region::compute_rats, region.cc:1912-1966).
 *   _const : interior stencil everywhere; every point on a non-periodic face
 *            carries the one-block fix, box [0,1)^3 (the shear_sim problem
 *            callbacks, src/shear_sim_3d.hh:388-402); fix may be NULL if all
 *            dimensions are periodic.
 *   _planes: one stencil per z plane (STZMG_OP_PLANE handles).
 *   _points: one stencil per point (STZMG_OP_POINT handles; the reference's
 *            internal()/fill_entries path, region.hh:663-790). */
int stzmg_set_operator_const(stzmg_t h, const double *blocks27x9, const double *fix9);
int stzmg_set_operator_planes(stzmg_t h, const double *planes);
int stzmg_set_operator_points(stzmg_t h, const double *points);
/* Copy out the operator of a level: ncls x 243 doubles, ncls = o (PLANE) or
 * m*n*o (POINT) of that level. */
int stzmg_get_level_operator(stzmg_t h, int level, double *out, long *ncls);

/* Fields (region::setup_fields / x0 / r0, region.hh:209-241). */
int stzmg_set_level_fields(stzmg_t h, int level, const double *x, const double *r);
int stzmg_get_level_fields(stzmg_t h, int level, double *x, double *r);
/* Device-resident SoA variants on level 0: x_dev/r_dev hold 3 planes of
 * m*n*o doubles (component-major).  The copy runs on the handle's stream. */
int stzmg_set_fields_device(stzmg_t h, const double *x_dev, const double *r_dev);
int stzmg_get_x_device(stzmg_t h, double *x_dev);

/* Level operations (multigrid::gauss_seidel / restriction / interpolation,
 * multigrid.cc:412-450).  restriction(l) fills level l from level l-1 and
 * clears x on level l; interpolation(l) adds P x_l into x_{l-1}.  The
 * smoother is 8-colour block Gauss-Seidel (a lexicographic sweep is not
 * parallel). */
int stzmg_gauss_seidel(stzmg_t h, int level, int nsweeps);
int stzmg_restriction(stzmg_t h, int level);
int stzmg_interpolation(stzmg_t h, int level);
/* r - A x on a level, copied to host AoS (region::residual, region.cc:1009-1055). */
int stzmg_residual(stzmg_t h, int level, double *out);
/* STZMG_NORM_QUARTIC = sum |r|^4 (l2_error_all, region.cc:3031-3057),
 * STZMG_NORM_L2SQ = sum |r|^2. */
int stzmg_norm(stzmg_t h, int level, int kind, double *out);
/* stzmg_norm without any device allocation (no reference counterpart; for
 * callers that cannot allocate: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same kernels and summation order
 * as stzmg_norm, so the same value; the norm partials go to a fixed device
 * array of the library, one per GPU (1 MiB, held until the process ends)
 * instead of the handle's lazily allocated buffer.  A level whose residual
 * needs more than 65536 partials (per-point residual kernel above 256 x 256
 * x 256 points) is rejected with STZMG_ERR_UNSUPPORTED.  All handles on one
 * GPU share that array: calls that overlap on one GPU (from several threads
 * or streams) race. */
int stzmg_norm_noalloc(stzmg_t h, int level, int kind, double *out);
/* multigrid::v_cycle (multigrid.cc:253-297), FP64. */
int stzmg_vcycle(stzmg_t h, int n);
/* One FP32 V-cycle on level-0 FP32 fields (x32, r32), for measurement. */
int stzmg_vcycle_fp32(stzmg_t h, int n);
/* The FP32 smoother and residual of the iterative-refinement inner solve on
 * a level (no reference counterpart; for parity tests): the level's FP64 x
 * and r rounded to FP32, then nsweeps FP32 sweeps with x widened back into
 * the FP64 x; or r - A x in FP32, widened and copied to host AoS. */
int stzmg_gauss_seidel_fp32(stzmg_t h, int level, int nsweeps);
int stzmg_residual_fp32(stzmg_t h, int level, double *out);
/* The projection solve loop (src/shear_sim_3d.cpp:567-600) and the added
 * FP32 + FP64 iterative-refinement mode. */
int stzmg_solve(stzmg_t h, const stzmg_solve_opts *opts, stzmg_solve_stats *stats);
void stzmg_solve_opts_default(stzmg_solve_opts *o);

/* Execution options (no reference counterpart; results are unchanged):
 *   STZMG_OPT_COARSE_KERNEL  levels of at most `value` points run as one
 *                            single-workgroup kernel per V-cycle; 1 selects
 *                            the built-in threshold, 0 launches every pass
 *                            separately. */
enum { STZMG_OPT_COARSE_KERNEL = 0, STZMG_OPT_FUSED_SMOOTHER = 1, STZMG_OPT_TILED_RESIDUAL = 2,
       STZMG_OPT_BOTTOM_MAP = 3, STZMG_OPT_CHUNK_CAP = 4, STZMG_OPT_RESTRICT_PLANES = 5,
       STZMG_OPT_CENTRO_GALERKIN = 6 };
/*   STZMG_OPT_FUSED_SMOOTHER 1 (default): per-plane levels with periodic x, y
 *                            (x and y sizes multiples of 8, more than 1024
 *                            points) sweep with two fused launches (one per z
 *                            parity); 0: eight colour launches per sweep.
 *   STZMG_OPT_TILED_RESIDUAL 1 (default): on the same levels the residual
 *                            marches tiles through z with planes staged in
 *                            LDS; 0: one thread per point reading its 27
 *                            neighbours through the caches.  Sums may differ
 *                            in the last bits (another summation order).
 *   STZMG_OPT_BOTTOM_MAP     1 (default): at the first V-cycle after the
 *                            operator is set, the
 *                            coarsest level's bottom sweeps (from x = 0, as
 *                            every V-cycle enters that level) are tabulated
 *                            as the matrix Q of order 3 N (N <= 1024 points,
 *                            at most 75 MB), and each V-cycle applies x = Q r
 *                            across the GPU instead of sweeping in one
 *                            workgroup.  Building Q costs a few milliseconds
 *                            per operator; 0 keeps the sweeps (for operators
 *                            that change every solve).  Results differ in the
 *                            last bits (another summation order).
 *   STZMG_OPT_CHUNK_CAP      0 (default): the fused sweeps and the
 *                            plane-marching residual cut z into as many
 *                            chunks as the device keeps resident at once;
 *                            value > 0 caps that count at `value` workgroups,
 *                            so small grids also march several planes per
 *                            workgroup (sweeps unchanged; residual norms may
 *                            differ in the last bits).
 *   STZMG_OPT_RESTRICT_PLANES 0 (default): coarse planes per restriction
 *                            thread chosen from the grid; value > 0 fixes it
 *                            (results unchanged).
 *   STZMG_OPT_CENTRO_GALERKIN 1 (default): a per-plane Galerkin level whose
 *                            mirrored blocks (offset o and -o) agree in every
 *                            entry to 1e-13 of the level's largest entry --
 *                            the product of a centrosymmetric operator, apart
 *                            by the rounding of the product's summation
 *                            order -- takes each pair's mean, so it is
 *                            centrosymmetric bit for bit and its sweeps add
 *                            opposite neighbours before one multiply.
 *                            Entries change in the last bits; levels whose
 *                            blocks differ by more (walls) are unchanged.
 *                            0: the product as summed. */
int stzmg_set_option(stzmg_t h, int option, int value);

/* The HIP stream (hipStream_t) the handle launches on. */
void *stzmg_stream(stzmg_t h);
/* Mean time of `reps` back-to-back repetitions of one operation on a level,
 * timed with HIP events on the handle's stream, for roofline accounting
 * (no reference counterpart): kind 0 = one FP64 Gauss-Seidel sweep,
 * 1 = one FP64 residual r - A x, 2 = one FP32 sweep, 3 = one FP32 residual;
 * kind 4 (FP64) and 5 (FP32) = the mean duration of one launch of the fused
 * sweep (two per sweep), from an event pair around each launch, without the
 * gaps between launches (STZMG_ERR_UNSUPPORTED on a level without the fused
 * sweep).  The operations run on the level's fields, which are overwritten;
 * the level stays usable (set its fields again before a solve). */
int stzmg_time_kernel(stzmg_t h, int kind, int level, int reps, double *ms_per_unit);

/* Operator builders (host side, FP64).
 * stzmg_shear_stencil: stencil = -dt*stencil_base and fix of
 *   set_up_stencil_base / set_up_stencil_and_matrices
 *   (src/shear_sim_3d.cpp:216-331, src/shear_sim_3d.hh:45-55).
 * stzmg_trans_stencil: the transformed-coordinate blocks of
 *   trans_sim_3d::set_up_stencil_and_matrices (src/trans_sim_3d.cpp:1187-1489,
 *   src/stenc_funcs.hh), scaled by -dt. */
int stzmg_shear_stencil(double mu, double lamb, double dx, double dy, double dz, double dt,
                        double kappa, double *blocks27x9, double *fix9);
int stzmg_trans_stencil(const double *T9, double mu, double lamb, double dx, double dy,
                        double dz, double dt, double *blocks27x9);

#ifdef __cplusplus
}
#endif
#endif
