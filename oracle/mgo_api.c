/* Flat handle API over mg_oracle.c for the Python test harness (ctypes).
 * Test infrastructure only. */
#include <stdlib.h>
#include <string.h>

#include <omp.h>

#include "mg_oracle.h"

void *mgo_new(int m, int n, int o, const int *prd, int mode, long coarsest) {
    mgo_hier *h = (mgo_hier *)calloc(1, sizeof(mgo_hier));
    if (mgo_build(h, m, n, o, prd, mode, coarsest) != 0) { mgo_free(h); free(h); return NULL; }
    return h;
}
void mgo_delete(void *h) { mgo_free((mgo_hier *)h); free(h); }
int mgo_nlev(void *h) { return ((mgo_hier *)h)->nlevels; }
void mgo_dims(void *hp, int l, int *d) {
    mgo_level *L = &((mgo_hier *)hp)->lv[l];
    d[0] = L->m; d[1] = L->n; d[2] = L->o;
}
static long ncl(const mgo_level *L) { return L->mode == MGO_PLANE ? L->o : (long)L->m * L->n * L->o; }
void mgo_set_level_ops(void *hp, int l, const double *ops) {
    mgo_level *L = &((mgo_hier *)hp)->lv[l];
    memcpy(L->ops, ops, sizeof(double) * ncl(L) * MGO_BLK);
}
void mgo_get_level_ops(void *hp, int l, double *ops) {
    mgo_level *L = &((mgo_hier *)hp)->lv[l];
    memcpy(ops, L->ops, sizeof(double) * ncl(L) * MGO_BLK);
}
void mgo_set_fields(void *hp, int l, const double *x, const double *r) {
    mgo_level *L = &((mgo_hier *)hp)->lv[l];
    long N = (long)L->m * L->n * L->o;
    if (x) memcpy(L->x, x, sizeof(double) * 3 * N);
    if (r) memcpy(L->r, r, sizeof(double) * 3 * N);
}
void mgo_get_fields(void *hp, int l, double *x, double *r) {
    mgo_level *L = &((mgo_hier *)hp)->lv[l];
    long N = (long)L->m * L->n * L->o;
    if (x) memcpy(x, L->x, sizeof(double) * 3 * N);
    if (r) memcpy(r, L->r, sizeof(double) * 3 * N);
}
void mgo_params(void *hp, int v_down, int v_up, int bottom, int growth, int smoother) {
    mgo_hier *h = (mgo_hier *)hp;
    h->v_down = v_down; h->v_up = v_up; h->bottom_sweeps = bottom; h->growth = growth;
    h->smoother = smoother;
}
void mgo_set_ranks(void *hp, int px, int py, int pz) {
    mgo_hier *h = (mgo_hier *)hp;
    h->ranks[0] = px < 1 ? 1 : px;
    h->ranks[1] = py < 1 ? 1 : py;
    h->ranks[2] = pz < 1 ? 1 : pz;
}
void mgo_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
}
int mgo_get_threads(void) { return omp_get_max_threads(); }
void mgo_do_rat(void *h) { mgo_rat((mgo_hier *)h); }
void mgo_do_residual(void *hp, int l, double *out) { mgo_residual(&((mgo_hier *)hp)->lv[l], out); }
void mgo_do_apply(void *hp, int l, const double *x, double *out) { mgo_apply(&((mgo_hier *)hp)->lv[l], x, out); }
void mgo_do_gs(void *hp, int l, int n) { for (int q = 0; q < n; q++) mgo_gs((mgo_hier *)hp, l); }
void mgo_do_restrict(void *hp, int l) { mgo_restrict((mgo_hier *)hp, l); }
void mgo_do_interp(void *hp, int l) { mgo_interp((mgo_hier *)hp, l); }
void mgo_do_vcycle(void *hp, int n) { for (int q = 0; q < n; q++) mgo_vcycle((mgo_hier *)hp); }
double mgo_do_norm_quartic(void *hp, int l) { return mgo_norm_quartic(&((mgo_hier *)hp)->lv[l]); }
double mgo_do_norm2(void *hp, int l) { return mgo_norm2(&((mgo_hier *)hp)->lv[l]); }
