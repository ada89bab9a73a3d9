/* CPU oracle for the multigrid projection solve (test infrastructure only).
 *
 * This is a plain-C FP64 restatement of the reference's geometric multigrid
 * (mg_3d_template/region.cc, multigrid.cc) on one subdomain.  It is used by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker; the product path never links or calls it.
 *
 * Conventions shared with the GPU library (include/stzmg.h):
 *   point index     p = i + m*(j + n*k)           (x fastest)
 *   vec3 fields     AoS, 3 doubles per point        (reference vec3 layout)
 *   stencil         27 blocks, block b = (di+1) + 3*(dj+1) + 9*(dk+1)
 *                   (reference index 13+i+3j+9k, src/shear_sim_3d.cpp:222)
 *   block           3x3 row-major a11 a12 a13 a21 ... a33 (src/mat3.hh)
 *   operator table  MGO_PLANE: one 27-block stencil per z plane (class = k)
 *                   MGO_POINT: one 27-block stencil per point (class = p)
 */
#ifndef MG_ORACLE_H
#define MG_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

#define MGO_MAX_LEVELS 16
#define MGO_PLANE 0
#define MGO_POINT 1
#define MGO_BLK 243 /* 27 blocks x 9 doubles */

typedef struct {
    int m, n, o;
    int prd[3];
    int mode;
    double *ops; /* per-class 27x9 table */
    double *x;   /* AoS vec3 solution */
    double *r;   /* AoS vec3 source */
} mgo_level;

typedef struct {
    int nlevels;
    mgo_level lv[MGO_MAX_LEVELS];
    int v_down, v_up, bottom_sweeps, growth;
    int smoother; /* 0 = lexicographic with per-sweep ghost snapshot, 1 = 8-colour */
    int ranks[3]; /* subdomain grid of the lexicographic sweep (1,1,1 = one rank) */
} mgo_hier;

int mgo_coarse_size(int nf, int prd);
int mgo_build(mgo_hier *h, int m, int n, int o, const int prd[3], int mode,
              long coarsest_total);
void mgo_free(mgo_hier *h);
const double *mgo_op(const mgo_level *L, int i, int j, int k);
void mgo_set_ops(mgo_hier *h, const double *ops);
void mgo_rat(mgo_hier *h);
void mgo_rat_level(const mgo_level *F, mgo_level *C);
void mgo_apply(const mgo_level *L, const double *x, double *out);
void mgo_residual(const mgo_level *L, double *out);
void mgo_gs_lex(mgo_level *L);
void mgo_gs_lex_blocks(mgo_level *L, const int ranks[3]);
void mgo_gs_color(mgo_level *L);
void mgo_gs(mgo_hier *h, int l);
void mgo_restrict(mgo_hier *h, int l);
void mgo_interp(mgo_hier *h, int l);
void mgo_vcycle(mgo_hier *h);
double mgo_norm_quartic(const mgo_level *L);
double mgo_norm2(const mgo_level *L);
double mgo_dot_rr(const double *a, long n);
void mgo_block_inverse(const double *a, double *inv);
double mgo_w1(int f, int I, int nf, int nc, int prd);
int mgo_color_count(const mgo_level *L);

#ifdef __cplusplus
}
#endif
#endif
