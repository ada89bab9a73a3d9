/* CPU oracle for the multigrid projection solve -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates in plain C the reference multigrid on one subdomain:
 *   hierarchy sizes        mg_3d_template/region.cc:854-861, multigrid.cc:157
 *   residual r - A x       mg_3d_template/region.cc:1009-1055
 *   Gauss-Seidel sweep     mg_3d_template/region.cc:1558-1762 (ghost refresh
 *                          once per sweep, region.cc:1601 / :2760-2847)
 *   restriction R = P^T/8  mg_3d_template/region.cc:1767-1812, :2130-2271
 *   interpolation x += P y mg_3d_template/region.cc:1817-1843, :2062-2121
 *   Galerkin (1/8) P^T A P mg_3d_template/region.cc:1912-1966, :2370-2538
 *   V-cycle schedule       mg_3d_template/multigrid.cc:253-297
 *   bottom solve           mg_3d_template/region.cc:813-815 (24 sweeps,
 *                          common.hh:57)
 *   stopping quantity      mg_3d_template/region.cc:3031-3057 with
 *                          mg3d_mod_sq(vec3) = |r|^4 (region.hh:13,
 *                          src/vec3.hh:110-112)
 *
 * The transfer operators are written as a tensor product of 1-D vertex
 * interpolations (mgo_w1); this reproduces the reference's masked 2x2x2
 * scatter routines, including the even non-periodic ("erow") and odd
 * periodic wrap cases handled by rt_scan_row/grid_map/igrid_map
 * (region.cc:1972-2055, :2588-2647).
 *
 * The 8-colour smoother (mgo_gs_color) has no reference counterpart: it is
 * the ordering the GPU kernels use, kept here so their sweeps can be
 * checked one for one.
 *
 * Loops whose iterations write disjoint points (operator application,
 * restriction, interpolation, the Galerkin product, the sweeps of one
 * colour and the subdomains of a lexicographic sweep) run on OpenMP
 * threads; every result is independent of the thread count.  Sums (the
 * norms) stay serial.
 */
#include "mg_oracle.h"

#include <stdlib.h>
#include <string.h>

int mgo_coarse_size(int nf, int prd) { return (nf + (prd ? 1 : 2)) >> 1; }

static long npts(const mgo_level *L) { return (long)L->m * L->n * L->o; }

static int wrap(int i, int n) { return ((i % n) + n) % n; }

const double *mgo_op(const mgo_level *L, int i, int j, int k) {
    if (L->mode == MGO_PLANE) return L->ops + (long)k * MGO_BLK;
    return L->ops + ((long)i + (long)L->m * (j + (long)L->n * k)) * MGO_BLK;
}

static long ncls(const mgo_level *L) { return L->mode == MGO_PLANE ? L->o : npts(L); }

int mgo_build(mgo_hier *h, int m, int n, int o, const int prd[3], int mode,
              long coarsest_total) {
    memset(h, 0, sizeof(*h));
    h->v_down = 2;
    h->v_up = 2;
    h->bottom_sweeps = 24;
    h->growth = 1;
    h->smoother = 0;
    h->ranks[0] = h->ranks[1] = h->ranks[2] = 1;
    int l = 0;
    for (;;) {
        mgo_level *L = &h->lv[l];
        L->m = m; L->n = n; L->o = o;
        L->prd[0] = prd[0]; L->prd[1] = prd[1]; L->prd[2] = prd[2];
        L->mode = mode;
        long N = npts(L);
        L->ops = (double *)calloc((size_t)ncls(L) * MGO_BLK, sizeof(double));
        L->x = (double *)calloc((size_t)N * 3, sizeof(double));
        L->r = (double *)calloc((size_t)N * 3, sizeof(double));
        if (!L->ops || !L->x || !L->r) return -1;
        l++;
        if (N <= coarsest_total) break;
        if (l == MGO_MAX_LEVELS) return -2;
        m = mgo_coarse_size(m, prd[0]);
        n = mgo_coarse_size(n, prd[1]);
        o = mgo_coarse_size(o, prd[2]);
    }
    h->nlevels = l;
    return 0;
}

void mgo_free(mgo_hier *h) {
    for (int l = 0; l < h->nlevels; l++) {
        free(h->lv[l].ops);
        free(h->lv[l].x);
        free(h->lv[l].r);
    }
    h->nlevels = 0;
}

void mgo_set_ops(mgo_hier *h, const double *ops) {
    memcpy(h->lv[0].ops, ops, sizeof(double) * ncls(&h->lv[0]) * MGO_BLK);
}

/* Weight of coarse vertex I in the 1-D interpolation to fine vertex f.
 * Non-periodic, nf odd : coarse I sits on fine 2I.
 * Non-periodic, nf even: as above except the last coarse vertex sits on the
 *                        last fine vertex nf-1 (region.cc "erow" rows).
 * Periodic, nf even    : coarse I on fine 2I, wrap between nc-1 and 0.
 * Periodic, nf odd     : coarse nc-1 on fine nf-1 and coarse 0 on fine 0 are
 *                        adjacent, with no midpoint between them. */
double mgo_w1(int f, int I, int nf, int nc, int prd) {
    if (prd) {
        f = wrap(f, nf);
        I = wrap(I, nc);
    } else if (f < 0 || f >= nf || I < 0 || I >= nc) {
        return 0.0;
    }
    if (!prd && !(nf & 1) && f == nf - 1) return I == nc - 1 ? 1.0 : 0.0;
    if (!(f & 1)) return I == (f >> 1) ? 1.0 : 0.0;
    int lo = f >> 1, hi = lo + 1;
    if (prd) hi = wrap(hi, nc);
    return (I == lo || I == hi) ? 0.5 : 0.0;
}

/* Coarse parents of fine vertex f: up to two (J, w) pairs. */
static int parents1(int f, int nf, int nc, int prd, int *J, double *w) {
    if (prd) f = wrap(f, nf);
    else if (f < 0 || f >= nf) return 0;
    if (!prd && !(nf & 1) && f == nf - 1) { J[0] = nc - 1; w[0] = 1.0; return 1; }
    if (!(f & 1)) { J[0] = f >> 1; w[0] = 1.0; return 1; }
    J[0] = f >> 1; w[0] = 0.5;
    J[1] = (f >> 1) + 1; w[1] = 0.5;
    if (prd) J[1] = wrap(J[1], nc);
    return 2;
}

static inline void mv_acc(const double *a, const double *v, double *acc) {
    acc[0] += a[0] * v[0] + a[1] * v[1] + a[2] * v[2];
    acc[1] += a[3] * v[0] + a[4] * v[1] + a[5] * v[2];
    acc[2] += a[6] * v[0] + a[7] * v[1] + a[8] * v[2];
}

/* Neighbour index with periodic wrap; -1 when outside a non-periodic dim. */
static long nbr(const mgo_level *L, int i, int j, int k, int *wrapped) {
    int w = 0;
    if (i < 0 || i >= L->m) { if (!L->prd[0]) return -1; i = wrap(i, L->m); w = 1; }
    if (j < 0 || j >= L->n) { if (!L->prd[1]) return -1; j = wrap(j, L->n); w = 1; }
    if (k < 0 || k >= L->o) { if (!L->prd[2]) return -1; k = wrap(k, L->o); w = 1; }
    if (wrapped) *wrapped = w;
    return (long)i + (long)L->m * (j + (long)L->n * k);
}

void mgo_apply(const mgo_level *L, const double *x, double *out) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < L->o; k++)
        for (int j = 0; j < L->n; j++)
            for (int i = 0; i < L->m; i++) {
                const double *A = mgo_op(L, i, j, k);
                double acc[3] = {0, 0, 0};
                for (int b = 0; b < 27; b++) {
                    int di = b % 3 - 1, dj = (b / 3) % 3 - 1, dk = b / 9 - 1;
                    long q = nbr(L, i + di, j + dj, k + dk, NULL);
                    if (q < 0) continue;
                    mv_acc(A + 9 * b, x + 3 * q, acc);
                }
                long p = (long)i + (long)L->m * (j + (long)L->n * k);
                out[3 * p] = acc[0];
                out[3 * p + 1] = acc[1];
                out[3 * p + 2] = acc[2];
            }
}

void mgo_residual(const mgo_level *L, double *out) {
    long N = npts(L);
    mgo_apply(L, L->x, out);
    for (long q = 0; q < 3 * N; q++) out[q] = L->r[q] - out[q];
}

/* Block inverse by cofactors over the determinant (src/mat3.hh:177-183). */
void mgo_block_inverse(const double *a, double *inv) {
    double c00 = a[4] * a[8] - a[5] * a[7];
    double c01 = a[2] * a[7] - a[1] * a[8];
    double c02 = a[1] * a[5] - a[2] * a[4];
    double c10 = a[5] * a[6] - a[3] * a[8];
    double c11 = a[0] * a[8] - a[2] * a[6];
    double c12 = a[2] * a[3] - a[0] * a[5];
    double c20 = a[3] * a[7] - a[4] * a[6];
    double c21 = a[1] * a[6] - a[0] * a[7];
    double c22 = a[0] * a[4] - a[1] * a[3];
    double det = a[0] * c00 + a[3] * c01 + a[6] * c02;
    double id = 1.0 / det;
    inv[0] = c00 * id; inv[1] = c01 * id; inv[2] = c02 * id;
    inv[3] = c10 * id; inv[4] = c11 * id; inv[5] = c12 * id;
    inv[6] = c20 * id; inv[7] = c21 * id; inv[8] = c22 * id;
}

/* One point of block Gauss-Seidel.  With a snapshot, every neighbour outside
 * the sweeping block [lo, hi) or reached through a periodic wrap is read from
 * it: the ghost layer the reference refreshes once per sweep
 * (region.cc:1601, :2760-2847).  Without one (the colour sweeps) all
 * neighbours are read live. */
static void gs_point(mgo_level *L, int i, int j, int k, const double *snap, const int *lo, const int *hi) {
    const double *A = mgo_op(L, i, j, k);
    double acc[3] = {0, 0, 0};
    for (int b = 0; b < 27; b++) {
        if (b == 13) continue;
        int di = b % 3 - 1, dj = (b / 3) % 3 - 1, dk = b / 9 - 1;
        int w = 0;
        long q = nbr(L, i + di, j + dj, k + dk, &w);
        if (q < 0) continue;
        const double *src = L->x;
        if (snap && (w || i + di < lo[0] || i + di >= hi[0] || j + dj < lo[1] || j + dj >= hi[1] ||
                     k + dk < lo[2] || k + dk >= hi[2]))
            src = snap;
        mv_acc(A + 9 * b, src + 3 * q, acc);
    }
    double inv[9], d[3];
    mgo_block_inverse(A + 9 * 13, inv);
    long p = (long)i + (long)L->m * (j + (long)L->n * k);
    d[0] = L->r[3 * p] - acc[0];
    d[1] = L->r[3 * p + 1] - acc[1];
    d[2] = L->r[3 * p + 2] - acc[2];
    L->x[3 * p] = inv[0] * d[0] + inv[1] * d[1] + inv[2] * d[2];
    L->x[3 * p + 1] = inv[3] * d[0] + inv[4] * d[1] + inv[5] * d[2];
    L->x[3 * p + 2] = inv[6] * d[0] + inv[7] * d[1] + inv[8] * d[2];
}

static int imin(int a, int b) { return a < b ? a : b; }

/* Lexicographic in-place block Gauss-Seidel of the reference on a grid of
 * ranks[0] x ranks[1] x ranks[2] subdomains: each subdomain
 * [m*ip/mp, m*(ip+1)/mp) x ... (the split of calc_processor_division's
 * equal blocks, src/common.cpp:7-60) is swept lexicographically, and
 * everything outside it, and every periodic image, is read from a snapshot
 * taken at the start of the sweep, as each rank of the reference reads its
 * ghost layer.  The subdomains are independent within a sweep and run on
 * OpenMP threads; the result does not depend on the thread count.  A 1x1x1
 * grid is the single-rank sweep; a grid of one point per subdomain is block
 * Jacobi.  Levels with fewer points than subdomains along a dimension use
 * one subdomain per point there. */
void mgo_gs_lex_blocks(mgo_level *L, const int ranks[3]) {
    long N = npts(L);
    double *snap = (double *)malloc(sizeof(double) * 3 * N);
    memcpy(snap, L->x, sizeof(double) * 3 * N);
    const int R[3] = {imin(ranks[0], L->m), imin(ranks[1], L->n), imin(ranks[2], L->o)};
    const int nb = R[0] * R[1] * R[2];
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nb; b++) {
        const int bi = b % R[0], bj = (b / R[0]) % R[1], bk = b / (R[0] * R[1]);
        const int lo[3] = {L->m * bi / R[0], L->n * bj / R[1], L->o * bk / R[2]};
        const int hi[3] = {L->m * (bi + 1) / R[0], L->n * (bj + 1) / R[1], L->o * (bk + 1) / R[2]};
        for (int k = lo[2]; k < hi[2]; k++)
            for (int j = lo[1]; j < hi[1]; j++)
                for (int i = lo[0]; i < hi[0]; i++) gs_point(L, i, j, k, snap, lo, hi);
    }
    free(snap);
}

/* The single-rank lexicographic sweep: periodic images are read from a
 * snapshot taken at the start of the sweep (region.cc:1601). */
void mgo_gs_lex(mgo_level *L) {
    const int one[3] = {1, 1, 1};
    mgo_gs_lex_blocks(L, one);
}

static int col1(int i, int n, int prd) { return (prd && (n & 1) && i == n - 1) ? 2 : (i & 1); }
static int ncol1(int n, int prd) { return (prd && (n & 1)) ? 3 : 2; }

int mgo_color_count(const mgo_level *L) {
    return ncol1(L->m, L->prd[0]) * ncol1(L->n, L->prd[1]) * ncol1(L->o, L->prd[2]);
}

/* Multi-colour block Gauss-Seidel: colour (cx, cy, cz) with cx = i mod 2
 * (a third colour for the last vertex of an odd periodic dimension), colours
 * visited with cx fastest.  Points of one colour are never neighbours, so the
 * update within a colour is order independent. */
void mgo_gs_color(mgo_level *L) {
    int ncx = ncol1(L->m, L->prd[0]), ncy = ncol1(L->n, L->prd[1]), ncz = ncol1(L->o, L->prd[2]);
    for (int cz = 0; cz < ncz; cz++)
        for (int cy = 0; cy < ncy; cy++)
            for (int cx = 0; cx < ncx; cx++)
#pragma omp parallel for schedule(static)
                for (int k = 0; k < L->o; k++) {
                    if (col1(k, L->o, L->prd[2]) != cz) continue;
                    for (int j = 0; j < L->n; j++) {
                        if (col1(j, L->n, L->prd[1]) != cy) continue;
                        for (int i = 0; i < L->m; i++)
                            if (col1(i, L->m, L->prd[0]) == cx) gs_point(L, i, j, k, NULL, NULL, NULL);
                    }
                }
}

void mgo_gs(mgo_hier *h, int l) {
    if (h->smoother == 1) mgo_gs_color(&h->lv[l]);
    else mgo_gs_lex_blocks(&h->lv[l], h->ranks);
}

/* Candidate fine vertices of coarse vertex I along one dimension. */
static int cand1(int I, int nf, int prd, int *f) {
    int c = 0;
    for (int d = -1; d <= 1; d++) {
        int g = 2 * I + d;
        if (prd) f[c++] = wrap(g, nf);
        else if (g >= 0 && g < nf) f[c++] = g;
    }
    return c;
}

/* r_c = (1/8) P^T (r_f - A_f x_f); x_c = 0 (multigrid.cc:436-450). */
void mgo_restrict(mgo_hier *h, int l) {
    mgo_level *F = &h->lv[l - 1], *C = &h->lv[l];
    long NF = npts(F), NC = npts(C);
    double *res = (double *)malloc(sizeof(double) * 3 * NF);
    mgo_residual(F, res);
#pragma omp parallel for schedule(static)
    for (int K = 0; K < C->o; K++)
        for (int J = 0; J < C->n; J++)
            for (int I = 0; I < C->m; I++) {
                int fi[3], fj[3], fk[3];
                int ni = cand1(I, F->m, F->prd[0], fi);
                int nj = cand1(J, F->n, F->prd[1], fj);
                int nk = cand1(K, F->o, F->prd[2], fk);
                double acc[3] = {0, 0, 0};
                for (int c = 0; c < nk; c++) {
                    double wk = mgo_w1(fk[c], K, F->o, C->o, F->prd[2]);
                    if (wk == 0) continue;
                    for (int b = 0; b < nj; b++) {
                        double wj = mgo_w1(fj[b], J, F->n, C->n, F->prd[1]);
                        if (wj == 0) continue;
                        for (int a = 0; a < ni; a++) {
                            double wi = mgo_w1(fi[a], I, F->m, C->m, F->prd[0]);
                            if (wi == 0) continue;
                            double w = 0.125 * wi * wj * wk;
                            long q = (long)fi[a] + (long)F->m * (fj[b] + (long)F->n * fk[c]);
                            acc[0] += w * res[3 * q];
                            acc[1] += w * res[3 * q + 1];
                            acc[2] += w * res[3 * q + 2];
                        }
                    }
                }
                long p = (long)I + (long)C->m * (J + (long)C->n * K);
                C->r[3 * p] = acc[0];
                C->r[3 * p + 1] = acc[1];
                C->r[3 * p + 2] = acc[2];
            }
    memset(C->x, 0, sizeof(double) * 3 * NC);
    free(res);
}

/* x_f += P x_c (multigrid.cc:420-430). */
void mgo_interp(mgo_hier *h, int l) {
    mgo_level *F = &h->lv[l - 1], *C = &h->lv[l];
#pragma omp parallel for schedule(static)
    for (int k = 0; k < F->o; k++) {
        int Jk[2]; double wk[2];
        int nk = parents1(k, F->o, C->o, F->prd[2], Jk, wk);
        for (int j = 0; j < F->n; j++) {
            int Jj[2]; double wj[2];
            int nj = parents1(j, F->n, C->n, F->prd[1], Jj, wj);
            for (int i = 0; i < F->m; i++) {
                int Ji[2]; double wi[2];
                int ni = parents1(i, F->m, C->m, F->prd[0], Ji, wi);
                double acc[3] = {0, 0, 0};
                for (int c = 0; c < nk; c++)
                    for (int b = 0; b < nj; b++)
                        for (int a = 0; a < ni; a++) {
                            double w = wi[a] * wj[b] * wk[c];
                            long q = (long)Ji[a] + (long)C->m * (Jj[b] + (long)C->n * Jk[c]);
                            acc[0] += w * C->x[3 * q];
                            acc[1] += w * C->x[3 * q + 1];
                            acc[2] += w * C->x[3 * q + 2];
                        }
                long p = (long)i + (long)F->m * (j + (long)F->n * k);
                F->x[3 * p] += acc[0];
                F->x[3 * p + 1] += acc[1];
                F->x[3 * p + 2] += acc[2];
            }
        }
    }
}

/* Offset of coarse vertex J relative to I in {-1,0,1}, or 99 if further. */
static int coff(int J, int I, int nc, int prd) {
    int d = J - I;
    if (prd) {
        if (d > nc / 2) d -= nc;
        if (d < -(nc / 2)) d += nc;
    }
    return (d >= -1 && d <= 1) ? d : 99;
}

/* Galerkin coarse stencil of coarse vertex (I,J,K):
 *   A_c[I][D] = (1/8) sum_f w(f,I) sum_d A_f[f][d] w(f+d, I+D)
 * (region.cc:2486-2538 builds the same sum row by row). */
static void rat_point(const mgo_level *F, const mgo_level *C, int I, int J, int K, double *out) {
    memset(out, 0, sizeof(double) * MGO_BLK);
    int fi[3], fj[3], fk[3];
    int ni = cand1(I, F->m, F->prd[0], fi);
    int nj = cand1(J, F->n, F->prd[1], fj);
    int nk = cand1(K, F->o, F->prd[2], fk);
    for (int c = 0; c < nk; c++) {
        double wk = mgo_w1(fk[c], K, F->o, C->o, F->prd[2]);
        if (wk == 0) continue;
        for (int b = 0; b < nj; b++) {
            double wj = mgo_w1(fj[b], J, F->n, C->n, F->prd[1]);
            if (wj == 0) continue;
            for (int a = 0; a < ni; a++) {
                double wi = mgo_w1(fi[a], I, F->m, C->m, F->prd[0]);
                if (wi == 0) continue;
                double wr = 0.125 * wi * wj * wk;
                const double *A = mgo_op(F, fi[a], fj[b], fk[c]);
                for (int blk = 0; blk < 27; blk++) {
                    int gi = fi[a] + blk % 3 - 1, gj = fj[b] + (blk / 3) % 3 - 1, gk = fk[c] + blk / 9 - 1;
                    int Pi[2], Pj[2], Pk[2];
                    double pwi[2], pwj[2], pwk[2];
                    int npi = parents1(gi, F->m, C->m, F->prd[0], Pi, pwi);
                    int npj = parents1(gj, F->n, C->n, F->prd[1], Pj, pwj);
                    int npk = parents1(gk, F->o, C->o, F->prd[2], Pk, pwk);
                    const double *Ab = A + 9 * blk;
                    for (int z = 0; z < npk; z++) {
                        int dk = coff(Pk[z], K, C->o, C->prd[2]);
                        for (int y = 0; y < npj; y++) {
                            int dj = coff(Pj[y], J, C->n, C->prd[1]);
                            for (int x = 0; x < npi; x++) {
                                int di = coff(Pi[x], I, C->m, C->prd[0]);
                                if (di == 99 || dj == 99 || dk == 99) continue;
                                double w = wr * pwi[x] * pwj[y] * pwk[z];
                                double *o = out + 9 * ((di + 1) + 3 * (dj + 1) + 9 * (dk + 1));
                                for (int e = 0; e < 9; e++) o[e] += w * Ab[e];
                            }
                        }
                    }
                }
            }
        }
    }
}

void mgo_rat_level(const mgo_level *F, mgo_level *C) {
    if (C->mode == MGO_PLANE) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int K = 0; K < C->o; K++) rat_point(F, C, 0, 0, K, C->ops + (long)K * MGO_BLK);
    } else {
#pragma omp parallel for schedule(dynamic, 1)
        for (int K = 0; K < C->o; K++)
            for (int J = 0; J < C->n; J++)
                for (int I = 0; I < C->m; I++)
                    rat_point(F, C, I, J, K,
                              C->ops + ((long)I + (long)C->m * (J + (long)C->n * K)) * MGO_BLK);
    }
}

void mgo_rat(mgo_hier *h) {
    for (int l = 1; l < h->nlevels; l++) mgo_rat_level(&h->lv[l - 1], &h->lv[l]);
}

/* One V-cycle with the reference's schedule: level l gets v_down+l pre and
 * v_up+l post sweeps, the bottom level bottom_sweeps (multigrid.cc:253-297). */
void mgo_vcycle(mgo_hier *h) {
    int L = h->nlevels;
    for (int i = 0; i < L - 1; i++) {
        int s = h->v_down + h->growth * i;
        for (int q = 0; q < s; q++) mgo_gs(h, i);
        mgo_restrict(h, i + 1);
    }
    for (int q = 0; q < h->bottom_sweeps; q++) mgo_gs(h, L - 1);
    for (int i = L - 1; i > 0; i--) {
        mgo_interp(h, i);
        int s = h->v_up + h->growth * (i - 1);
        for (int q = 0; q < s; q++) mgo_gs(h, i - 1);
    }
}

double mgo_norm_quartic(const mgo_level *L) {
    long N = npts(L);
    double *res = (double *)malloc(sizeof(double) * 3 * N);
    mgo_residual(L, res);
    double s = 0;
    for (long p = 0; p < N; p++) {
        double q = res[3 * p] * res[3 * p] + res[3 * p + 1] * res[3 * p + 1] + res[3 * p + 2] * res[3 * p + 2];
        s += q * q;
    }
    free(res);
    return s;
}

double mgo_norm2(const mgo_level *L) {
    long N = npts(L);
    double *res = (double *)malloc(sizeof(double) * 3 * N);
    mgo_residual(L, res);
    double s = mgo_dot_rr(res, 3 * N);
    free(res);
    return s;
}

double mgo_dot_rr(const double *a, long n) {
    double s = 0;
    for (long q = 0; q < n; q++) s += a[q] * a[q];
    return s;
}
