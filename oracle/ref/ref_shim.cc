// C-ABI harness around the reference STZ++ sources (test infrastructure only).
//
// Built by oracle/ref/Makefile from the reference sources where they lie
// (src/{shear_sim_3d,trans_sim_3d,common}.cpp and
// mg_3d_template/{buffer,common,geometry}.cc) into oracle/_ref/libstzref.so.
// The classes are driven through their public interfaces on a single MPI rank:
//   multigrid<vec3,mat3> with a problem class that serves per-point 27-block
//   stencils from arrays (duck-typed interface of region.hh:663-790),
//   shear_sim_3d<sym_mat3> QS steps (src/shear_sim_3d.cpp:381-389),
//   the transformed-coordinate stencil functions of src/stenc_funcs.hh.
// Nothing here is linked into the product library.
#include <mpi.h>

#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "shear_sim_3d.hh"
#include "trans_sim_3d.hh"

namespace {

void ensure_mpi() {
    int ini = 0;
    MPI_Initialized(&ini);
    if (!ini) {
        int argc = 0;
        char **argv = nullptr;
        MPI_Init(&argc, &argv);
    }
}

// Problem class for region<vec3,mat3>: every point stores its own box.
struct arr_problem {
    bool x_prd, y_prd, z_prd;
    int m, n, o;
    const int *S;          // 6 ints per point (lo/hi offsets, hi exclusive)
    const double *blocks;  // 243 doubles per point, block (di+1)+3(dj+1)+9(dk+1)
    int pidx(int i, int j, int k) const { return i + m * (j + n * k); }
    int range_xd(int i, int j, int k) { return S[6 * pidx(i, j, k)]; }
    int range_xu(int i, int j, int k) { return S[6 * pidx(i, j, k) + 1]; }
    int range_yd(int i, int j, int k) { return S[6 * pidx(i, j, k) + 2]; }
    int range_yu(int i, int j, int k) { return S[6 * pidx(i, j, k) + 3]; }
    int range_zd(int i, int j, int k) { return S[6 * pidx(i, j, k) + 4]; }
    int range_zu(int i, int j, int k) { return S[6 * pidx(i, j, k) + 5]; }
    bool internal(int, int, int) { return true; }
    int mem_size(int i, int j, int k) {
        const int *s = S + 6 * pidx(i, j, k);
        return (s[1] - s[0]) * (s[3] - s[2]) * (s[5] - s[4]);
    }
    mat3 *external_ptr(int, int, int) { return nullptr; }
    void fill_entries(int i, int j, int k, mat3 *&en) {
        const int *s = S + 6 * pidx(i, j, k);
        const double *B = blocks + 243L * pidx(i, j, k);
        for (int dk = s[4]; dk < s[5]; dk++)
            for (int dj = s[2]; dj < s[3]; dj++)
                for (int di = s[0]; di < s[1]; di++) {
                    const double *b = B + 9 * ((di + 1) + 3 * (dj + 1) + 9 * (dk + 1));
                    *(en++) = mat3(b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7], b[8]);
                }
    }
    vec3 x_field(int, int, int, int, int, int) { return vec3(0); }
    vec3 r_field(int, int, int, int, int, int) { return vec3(0); }
};

struct mg_handle {
    geometry *gm;
    comm_buffer *com;
    arr_problem pr;
    multigrid<vec3, mat3> *mg;
};

}  // namespace

extern "C" {

void *ref_mg_create(int m, int n, int o, const int *prd, const int *S, const double *blocks) {
    ensure_mpi();
    mg_handle *h = new mg_handle;
    h->gm = new geometry(m, n, o, prd[0] != 0, prd[1] != 0, prd[2] != 0);
    h->com = new comm_buffer;
    h->pr.x_prd = prd[0] != 0;
    h->pr.y_prd = prd[1] != 0;
    h->pr.z_prd = prd[2] != 0;
    h->pr.m = m;
    h->pr.n = n;
    h->pr.o = o;
    h->pr.S = S;
    h->pr.blocks = blocks;
    h->mg = new multigrid<vec3, mat3>(h->pr, *h->gm, *h->com);
    h->mg->setup_matrices(h->pr);
    return h;
}

void ref_mg_destroy(void *hp) {
    mg_handle *h = (mg_handle *)hp;
    delete h->mg;
    delete h->com;
    delete h->gm;
    delete h;
}

int ref_mg_nlevels(void *hp) { return ((mg_handle *)hp)->mg->nlevels; }

void ref_mg_dims(void *hp, int l, int *dims) {
    region<vec3, mat3> *R = ((mg_handle *)hp)->mg->reg[l];
    dims[0] = R->sm;
    dims[1] = R->sn;
    dims[2] = R->so;
}

void ref_mg_set_sweeps(void *hp, int v_down, int v_up) {
    mg_handle *h = (mg_handle *)hp;
    h->mg->v_down = v_down;
    h->mg->v_up = v_up;
}

// Expand the per-point box entries of level l into 27-block slots.
int ref_mg_get_operator(void *hp, int l, double *out) {
    region<vec3, mat3> *R = ((mg_handle *)hp)->mg->reg[l];
    int N = R->smno;
    memset(out, 0, sizeof(double) * 243L * N);
    for (int p = 0; p < N; p++) {
        const int *s = R->S + 6 * p;
        const mat3 *a = R->A[p];
        for (int dk = s[4]; dk < s[5]; dk++)
            for (int dj = s[2]; dj < s[3]; dj++)
                for (int di = s[0]; di < s[1]; di++, a++) {
                    if (di < -1 || di > 1 || dj < -1 || dj > 1 || dk < -1 || dk > 1) return -1;
                    double *b = out + 243L * p + 9 * ((di + 1) + 3 * (dj + 1) + 9 * (dk + 1));
                    b[0] += a->a11; b[1] += a->a12; b[2] += a->a13;
                    b[3] += a->a21; b[4] += a->a22; b[5] += a->a23;
                    b[6] += a->a31; b[7] += a->a32; b[8] += a->a33;
                }
    }
    return 0;
}

void ref_mg_get_boxes(void *hp, int l, int *S) {
    region<vec3, mat3> *R = ((mg_handle *)hp)->mg->reg[l];
    memcpy(S, R->S, sizeof(int) * 6 * R->smno);
}

void ref_mg_set_fields(void *hp, int l, const double *x, const double *r) {
    region<vec3, mat3> *R = ((mg_handle *)hp)->mg->reg[l];
    for (int k = 0; k < R->so; k++)
        for (int j = 0; j < R->sn; j++)
            for (int i = 0; i < R->sm; i++) {
                int p = i + R->sm * (j + R->sn * k);
                if (x) R->x0[i + R->mg * (j + R->ng * k)] = vec3(x[3 * p], x[3 * p + 1], x[3 * p + 2]);
                if (r) R->r0[p] = vec3(r[3 * p], r[3 * p + 1], r[3 * p + 2]);
            }
}

void ref_mg_get_fields(void *hp, int l, double *x, double *r) {
    region<vec3, mat3> *R = ((mg_handle *)hp)->mg->reg[l];
    for (int k = 0; k < R->so; k++)
        for (int j = 0; j < R->sn; j++)
            for (int i = 0; i < R->sm; i++) {
                int p = i + R->sm * (j + R->sn * k);
                if (x) {
                    vec3 v = R->x0[i + R->mg * (j + R->ng * k)];
                    x[3 * p] = v.x; x[3 * p + 1] = v.y; x[3 * p + 2] = v.z;
                }
                if (r) {
                    vec3 v = R->r0[p];
                    r[3 * p] = v.x; r[3 * p + 1] = v.y; r[3 * p + 2] = v.z;
                }
            }
}

void ref_mg_residual(void *hp, int l, double *out) {
    region<vec3, mat3> *R = ((mg_handle *)hp)->mg->reg[l];
    R->communicate();
    for (int k = 0; k < R->so; k++)
        for (int j = 0; j < R->sn; j++)
            for (int i = 0; i < R->sm; i++) {
                int p = i + R->sm * (j + R->sn * k);
                vec3 v = R->residual(i, j, k);
                out[3 * p] = v.x; out[3 * p + 1] = v.y; out[3 * p + 2] = v.z;
            }
}

void ref_mg_gs(void *hp, int l, int n) { for (int q = 0; q < n; q++) ((mg_handle *)hp)->mg->gauss_seidel(l); }
void ref_mg_restrict(void *hp, int l) { ((mg_handle *)hp)->mg->restriction(l); }
void ref_mg_interp(void *hp, int l) { ((mg_handle *)hp)->mg->interpolation(l); }
void ref_mg_vcycle(void *hp, int n) { for (int q = 0; q < n; q++) ((mg_handle *)hp)->mg->v_cycle(); }
double ref_mg_l2_error_all(void *hp, int l) { return ((mg_handle *)hp)->mg->reg[l]->l2_error_all(); }

// Transformed-coordinate stencil: the 19 stenc_funcs blocks (corners zero),
// the same name-to-offset convention as src/stencil_writer.py:1-21.
void ref_trans_stencil(const double *T, double mu, double lamb, double dx, double dy, double dz,
                       double *out) {
    double D = T[0] * (T[4] * T[8] - T[5] * T[7]) - T[1] * (T[3] * T[8] - T[5] * T[6]) +
               T[2] * (T[3] * T[7] - T[4] * T[6]);
    memset(out, 0, sizeof(double) * 243);
#define TA T[0], T[1], T[2], T[3], T[4], T[5], T[6], T[7], T[8], mu, lamb, D, dx, dy, dz
#define BLK(b, suf)                                       \
    out[9 * (b) + 0] = U##suf##_1(TA);                    \
    out[9 * (b) + 1] = V##suf##_1(TA);                    \
    out[9 * (b) + 2] = W##suf##_1(TA);                    \
    out[9 * (b) + 3] = U##suf##_2(TA);                    \
    out[9 * (b) + 4] = V##suf##_2(TA);                    \
    out[9 * (b) + 5] = W##suf##_2(TA);                    \
    out[9 * (b) + 6] = U##suf##_3(TA);                    \
    out[9 * (b) + 7] = V##suf##_3(TA);                    \
    out[9 * (b) + 8] = W##suf##_3(TA);
    BLK(9, xmym) BLK(3, xmzm) BLK(12, xm) BLK(21, xmzp) BLK(15, xmyp)
    BLK(1, ymzm) BLK(10, ym) BLK(19, ymzp) BLK(4, zm) BLK(13, c) BLK(22, zp)
    BLK(7, ypzm) BLK(16, yp) BLK(25, ypzp) BLK(11, xpym) BLK(5, xpzm)
    BLK(14, xp) BLK(23, xpzp) BLK(17, xpyp)
#undef BLK
#undef TA
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Shear simulation harness: drives shear_sim_3d<sym_mat3>'s QS step.
namespace {
class ref_shear : public shear_sim_3d<sym_mat3> {
   public:
    using shear_sim_3d<sym_mat3>::shear_sim_3d;
    int nx() const { return N_x; }
    int ny() const { return N_y; }
    int nz() const { return N_z; }
    int gx() const { return nxg; }
    int gxy() const { return nxyg; }
    void boundary() { set_up_ghost_regions(); set_boundary_conditions(); }
    void net_force() { compute_net_force(); }
    void ref_init() { init_ref(); }
    double xi(int i, int j, int k, int a) const {
        const Field &f = grido[i + nxg * j + nxyg * k];
        return a == 0 ? f.xi_x : (a == 1 ? f.xi_y : f.xi_z);
    }
    const sym_mat3 *stencil_ptr() const { return stencil; }
    const sym_mat3 *fix_ptr() const { return fix; }
};

struct shear_handle {
    MPI_Comm cart;
    int dims[3];
    int periods[3];
    ref_shear *sim;
};
}  // namespace

extern "C" {

// p: mu, lamb, sy, tau0, c0, eps0, delta, omega, bath_temp, chi_inf, ez, TZ, diff_l, U
void *ref_shear_create(int Nx, int Ny, int Nz, const double *box, const double *p,
                       const int *periods, const char *outdir) {
    ensure_mpi();
    shear_handle *h = new shear_handle;
    h->dims[0] = h->dims[1] = h->dims[2] = 1;
    for (int a = 0; a < 3; a++) h->periods[a] = periods[a];
    MPI_Cart_create(MPI_COMM_WORLD, 3, h->dims, h->periods, 1, &h->cart);
    h->sim = new ref_shear(Nx, Ny, Nz, box[0], box[2], box[4], box[1], box[3], box[5],
                           0.0, 1.0, 0.5, 1, std::string(outdir), p[0], p[1], p[0], 0.0,
                           p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9], p[10], p[11],
                           p[12], p[13], 5, 3, 0.09, 0.01, 1.0, &h->cart, h->dims, 15, 0.0,
                           0.0, h->periods);
    return h;
}

void ref_shear_setup(void *hp, double dt) {
    shear_handle *h = (shear_handle *)hp;
    h->sim->set_up_stencil_base(dt);
    h->sim->set_up_stencil_and_matrices(dt);
}

void ref_shear_get_stencil(void *hp, double *blocks27x6, double *fix6) {
    shear_handle *h = (shear_handle *)hp;
    const sym_mat3 *s = h->sim->stencil_ptr();
    for (int b = 0; b < 27; b++) {
        double *o = blocks27x6 + 6 * b;
        o[0] = s[b].a11; o[1] = s[b].a12; o[2] = s[b].a13;
        o[3] = s[b].a22; o[4] = s[b].a23; o[5] = s[b].a33;
    }
    const sym_mat3 *f = h->sim->fix_ptr();
    fix6[0] = f->a11; fix6[1] = f->a12; fix6[2] = f->a13;
    fix6[3] = f->a22; fix6[4] = f->a23; fix6[5] = f->a33;
}

// u: vertex vec3 (Nx, Ny, Nz+1 when z is walled), sig: 6 per cell, chi: 1 per cell.
void ref_shear_set_state(void *hp, double t, const double *u, const double *sig, const double *chi) {
    shear_handle *h = (shear_handle *)hp;
    ref_shear *s = h->sim;
    int Nx = s->nx(), Ny = s->ny(), Nz = s->nz(), gx = s->gx(), gxy = s->gxy();
    int nzv = Nz + (h->periods[2] ? 0 : 1);
    s->sim_time = t;
    for (int k = 0; k < nzv; k++)
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Nx; i++) {
                Field &f = s->grido[i + gx * j + gxy * k];
                long q = i + (long)Nx * (j + (long)Ny * k);
                f.u = u[3 * q]; f.v = u[3 * q + 1]; f.w = u[3 * q + 2];
                if (k < Nz) {
                    long c = i + (long)Nx * (j + (long)Ny * k);
                    f.s11 = sig[6 * c]; f.s12 = sig[6 * c + 1]; f.s13 = sig[6 * c + 2];
                    f.s22 = sig[6 * c + 3]; f.s23 = sig[6 * c + 4]; f.s33 = sig[6 * c + 5];
                    f.chi = chi[c];
                    f.cu = f.cv = f.cw = f.cs11 = f.cs12 = f.cs13 = f.cs22 = f.cs23 = f.cs33 = 0;
                    f.cchi = f.cxi_x = f.cxi_y = f.cxi_z = 0;
                    f.ad_Dpl = 0;
                    f.sbar = 0;
                }
            }
    s->boundary();
}

void ref_shear_step(void *hp, double dt, int nsteps) {
    shear_handle *h = (shear_handle *)hp;
    for (int q = 0; q < nsteps; q++) h->sim->step_forward_qs(dt);
}

void ref_shear_get_state(void *hp, double *t, double *u, double *sig, double *chi, double *dpl) {
    shear_handle *h = (shear_handle *)hp;
    ref_shear *s = h->sim;
    int Nx = s->nx(), Ny = s->ny(), Nz = s->nz(), gx = s->gx(), gxy = s->gxy();
    int nzv = Nz + (h->periods[2] ? 0 : 1);
    *t = s->sim_time;
    for (int k = 0; k < nzv; k++)
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Nx; i++) {
                const Field &f = s->grido[i + gx * j + gxy * k];
                long q = i + (long)Nx * (j + (long)Ny * k);
                u[3 * q] = f.u; u[3 * q + 1] = f.v; u[3 * q + 2] = f.w;
                if (k < Nz) {
                    sig[6 * q] = f.s11; sig[6 * q + 1] = f.s12; sig[6 * q + 2] = f.s13;
                    sig[6 * q + 3] = f.s22; sig[6 * q + 4] = f.s23; sig[6 * q + 5] = f.s33;
                    chi[q] = f.chi;
                    dpl[q] = f.ad_Dpl;
                }
            }
}

int ref_shear_vcycles(void *hp) { return ((shear_handle *)hp)->sim->n_vcycles; }

void ref_shear_destroy(void *hp) {
    shear_handle *h = (shear_handle *)hp;
    delete h->sim;
    MPI_Comm_free(&h->cart);
    delete h;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Transformed-coordinate harness: drives trans_sim_3d's QS step (its
// compiled-in transform is pure shear, src/trans_sim_3d.hh:76-94).
namespace {
class ref_trans : public trans_sim_3d {
   public:
    using trans_sim_3d::trans_sim_3d;
    int nx() const { return N_x; }
    int ny() const { return N_y; }
    int nz() const { return N_z; }
    int gx() const { return nxg; }
    int gxy() const { return nxyg; }
    void boundary() { set_up_ghost_regions(); set_boundary_conditions(); }
    void net_force() {
        update_matrices();
        compute_net_force();
    }
    const mat3 *stencil_ptr() const { return stencil; }
    const mat3 *fix_ptr() const { return fix; }
    void matrices(double *T9, double *Ti9, double *dT9) {
        update_matrices();
        const mat3 *m[3] = {&T, &Tinv, &dTdt};
        double *o[3] = {T9, Ti9, dT9};
        for (int q = 0; q < 3; q++) {
            o[q][0] = m[q]->a11; o[q][1] = m[q]->a12; o[q][2] = m[q]->a13;
            o[q][3] = m[q]->a21; o[q][4] = m[q]->a22; o[q][5] = m[q]->a23;
            o[q][6] = m[q]->a31; o[q][7] = m[q]->a32; o[q][8] = m[q]->a33;
        }
    }
};

struct trans_handle {
    MPI_Comm cart;
    int dims[3];
    int periods[3];
    ref_trans *sim;
};

template <typename SIM>
void put_state(SIM *s, const int *periods, double t, const double *u, const double *sig, const double *chi) {
    int Nx = s->nx(), Ny = s->ny(), Nz = s->nz(), gx = s->gx(), gxy = s->gxy();
    int nzv = Nz + (periods[2] ? 0 : 1);
    s->sim_time = t;
    for (int k = 0; k < nzv; k++)
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Nx; i++) {
                Field &f = s->grido[i + gx * j + gxy * k];
                long q = i + (long)Nx * (j + (long)Ny * k);
                f.u = u[3 * q]; f.v = u[3 * q + 1]; f.w = u[3 * q + 2];
                if (k < Nz) {
                    f.s11 = sig[6 * q]; f.s12 = sig[6 * q + 1]; f.s13 = sig[6 * q + 2];
                    f.s22 = sig[6 * q + 3]; f.s23 = sig[6 * q + 4]; f.s33 = sig[6 * q + 5];
                    f.chi = chi[q];
                    f.cu = f.cv = f.cw = f.cs11 = f.cs12 = f.cs13 = f.cs22 = f.cs23 = f.cs33 = 0;
                    f.cchi = f.cxi_x = f.cxi_y = f.cxi_z = 0;
                    f.ad_Dpl = 0;
                    f.sbar = 0;
                }
            }
    s->boundary();
}

template <typename SIM>
void take_state(SIM *s, const int *periods, double *t, double *u, double *sig, double *chi, double *dpl) {
    int Nx = s->nx(), Ny = s->ny(), Nz = s->nz(), gx = s->gx(), gxy = s->gxy();
    int nzv = Nz + (periods[2] ? 0 : 1);
    *t = s->sim_time;
    for (int k = 0; k < nzv; k++)
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Nx; i++) {
                const Field &f = s->grido[i + gx * j + gxy * k];
                long q = i + (long)Nx * (j + (long)Ny * k);
                u[3 * q] = f.u; u[3 * q + 1] = f.v; u[3 * q + 2] = f.w;
                if (k < Nz) {
                    sig[6 * q] = f.s11; sig[6 * q + 1] = f.s12; sig[6 * q + 2] = f.s13;
                    sig[6 * q + 3] = f.s22; sig[6 * q + 4] = f.s23; sig[6 * q + 5] = f.s33;
                    chi[q] = f.chi;
                    dpl[q] = f.ad_Dpl;
                }
            }
}
}  // namespace

extern "C" {

// p as ref_shear_create; Ut: rate of the pure-shear transform.
void *ref_trans_create(int Nx, int Ny, int Nz, const double *box, const double *p, double Ut,
                       const int *periods, const char *outdir) {
    ensure_mpi();
    trans_handle *h = new trans_handle;
    h->dims[0] = h->dims[1] = h->dims[2] = 1;
    for (int a = 0; a < 3; a++) h->periods[a] = periods[a];
    MPI_Cart_create(MPI_COMM_WORLD, 3, h->dims, h->periods, 1, &h->cart);
    h->sim = new ref_trans(Nx, Ny, Nz, box[0], box[2], box[4], box[1], box[3], box[5], 0.0, 1.0, 0.5, 1,
                           std::string(outdir), p[0], p[1], p[0], 0.0, p[2], p[3], p[4], p[5], p[6], p[7],
                           p[8], p[9], p[10], p[11], p[12], Ut, p[13], 5, 3, 0.09, 0.01, 1.0, &h->cart, h->dims,
                           15, h->periods);
    return h;
}

void ref_trans_setup(void *hp, double dt) { ((trans_handle *)hp)->sim->set_up_stencil_and_matrices(dt); }

void ref_trans_get_stencil(void *hp, double *blocks27x9, double *fix9) {
    trans_handle *h = (trans_handle *)hp;
    const mat3 *s = h->sim->stencil_ptr();
    for (int b = 0; b < 27; b++) {
        double *o = blocks27x9 + 9 * b;
        o[0] = s[b].a11; o[1] = s[b].a12; o[2] = s[b].a13;
        o[3] = s[b].a21; o[4] = s[b].a22; o[5] = s[b].a23;
        o[6] = s[b].a31; o[7] = s[b].a32; o[8] = s[b].a33;
    }
    const mat3 *f = h->sim->fix_ptr();
    fix9[0] = f->a11; fix9[1] = f->a12; fix9[2] = f->a13;
    fix9[3] = f->a21; fix9[4] = f->a22; fix9[5] = f->a23;
    fix9[6] = f->a31; fix9[7] = f->a32; fix9[8] = f->a33;
}

void ref_trans_matrices(void *hp, double *T, double *Ti, double *dT) {
    ((trans_handle *)hp)->sim->matrices(T, Ti, dT);
}

void ref_trans_set_state(void *hp, double t, const double *u, const double *sig, const double *chi) {
    trans_handle *h = (trans_handle *)hp;
    put_state(h->sim, h->periods, t, u, sig, chi);
}

void ref_trans_step(void *hp, double dt, int nsteps) {
    trans_handle *h = (trans_handle *)hp;
    for (int q = 0; q < nsteps; q++) h->sim->step_forward_qs(dt);
}

void ref_trans_get_state(void *hp, double *t, double *u, double *sig, double *chi, double *dpl) {
    trans_handle *h = (trans_handle *)hp;
    take_state(h->sim, h->periods, t, u, sig, chi, dpl);
}

int ref_trans_vcycles(void *hp) { return ((trans_handle *)hp)->sim->n_vcycles; }

void ref_trans_destroy(void *hp) {
    trans_handle *h = (trans_handle *)hp;
    delete h->sim;
    MPI_Comm_free(&h->cart);
    delete h;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Random initial chi of shear_sim_3d (set_random_initial_conditions,
// src/shear_sim_3d.cpp:2793-2934).  The reference draws its noise with
// glibc rand() and never seeds it, i.e. from the state srand(1) sets; the
// harness resets that state, asks the reference's own fill_noise_grid for
// the noise grid, resets it again and runs set_random_initial_conditions in
// workdir, where the reference writes its cg_*.dat cache.
namespace {
class ref_grf_sim : public shear_sim_3d<sym_mat3> {
   public:
    using shear_sim_3d<sym_mat3>::shear_sim_3d;
    int cut_pts() const { return cut; }
    double chi_scale() const { return chi_1; }
    double normal_factor() { return gauss_normal_factor_long(); }
    void noise(double *b, double *e) { fill_noise_grid(b, e); }
    void random_ic() { set_random_initial_conditions(); }
    double chi_at(int i, int j, int k) const { return grido[i + nxg * j + nxyg * k].chi; }
};
}  // namespace

extern "C" {

// p: gauss_fac, conv_l, chi_avg, sigma, TZ.  noise_out holds noise_n values
// (the padded grid), chi_out Nx Ny Nz (x fastest).  out3: chi_1 after
// normalisation, gauss_normal_factor_long, cut.  Returns 0, or -1 when
// noise_n is not the padded size.
int ref_grf(int Nx, int Ny, int Nz, const int *periods, const double *p, int sim_case, const char *workdir,
            double *noise_out, long noise_n, double *chi_out, double *out3) {
    ensure_mpi();
    MPI_Comm cart;
    int dims[3] = {1, 1, 1}, per[3] = {periods[0], periods[1], periods[2]};
    MPI_Cart_create(MPI_COMM_WORLD, 3, dims, per, 1, &cart);
    ref_grf_sim *s = new ref_grf_sim(Nx, Ny, Nz, -1, -1, -0.5, 1, 1, 0.5, 0.0, 1.0, 0.5, 1, std::string(workdir),
                                     1.0, 1.5, 1.0, 0.0, 1.0, 1.0, 0.4, 10.0, 1.0, 1.0, 1.0, 0.15, 1.0, p[4], 0.0,
                                     0.01, p[0], p[1], p[2], p[3], 1.0, &cart, dims, sim_case, 0.0, 0.0, per);
    const int cut = s->cut_pts();
    const long nt = (long)(per[0] ? Nx : Nx + 2 * cut) * (per[1] ? Ny : Ny + 2 * cut) * (per[2] ? Nz : Nz + 2 * cut);
    int rc = 0;
    if (nt != noise_n) {
        rc = -1;
    } else {
        srand(1);
        s->noise(noise_out, noise_out + nt);
        char old[4096];
        if (!getcwd(old, sizeof(old))) old[0] = 0;
        if (chdir(workdir) == 0) {
            srand(1);
            s->random_ic();
            if (old[0] && chdir(old) != 0) rc = -2;
        } else {
            rc = -3;
        }
        for (int k = 0; k < Nz; k++)
            for (int j = 0; j < Ny; j++)
                for (int i = 0; i < Nx; i++) chi_out[i + (long)Nx * (j + (long)Ny * k)] = s->chi_at(i, j, k);
        out3[0] = s->chi_scale();
        out3[1] = s->normal_factor();
        out3[2] = cut;
    }
    delete s;
    MPI_Comm_free(&cart);
    return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Output of the reference (output_cube, src/shear_sim_3d.cpp:1963-2161;
// compute_net_force, :1061-1126 and src/trans_sim_3d.cpp:1575-1660): the
// files land in the directory the simulation was created with.
extern "C" {

void ref_shear_output_cube(void *hp, const char *prefix, int mode, int sn) {
    ((shear_handle *)hp)->sim->output_cube(prefix, mode, sn);
}

void ref_shear_net_force(void *hp) { ((shear_handle *)hp)->sim->net_force(); }

void ref_trans_output_cube(void *hp, const char *prefix, int mode, int sn) {
    ((trans_handle *)hp)->sim->output_cube(prefix, mode, sn);
}

void ref_trans_net_force(void *hp) { ((trans_handle *)hp)->sim->net_force(); }

}  // extern "C"

// ---------------------------------------------------------------------------
// Direct (explicit dynamics) step of shear_sim_3d (step_forward,
// src/shear_sim_3d.cpp:206-213) with density rho and viscous damping kappa.
extern "C" {

void *ref_shear_create_direct(int Nx, int Ny, int Nz, const double *box, const double *p, double rho, double kappa,
                              const int *periods, const char *outdir) {
    ensure_mpi();
    shear_handle *h = new shear_handle;
    h->dims[0] = h->dims[1] = h->dims[2] = 1;
    for (int a = 0; a < 3; a++) h->periods[a] = periods[a];
    MPI_Cart_create(MPI_COMM_WORLD, 3, h->dims, h->periods, 1, &h->cart);
    h->sim = new ref_shear(Nx, Ny, Nz, box[0], box[2], box[4], box[1], box[3], box[5], 0.0, 1.0, 0.5, 1,
                           std::string(outdir), p[0], p[1], rho, kappa, p[2], p[3], p[4], p[5], p[6], p[7], p[8],
                           p[9], p[10], p[11], p[12], p[13], 5, 3, 0.09, 0.01, 1.0, &h->cart, h->dims, 15, 0.0, 0.0,
                           h->periods);
    return h;
}

void ref_shear_init_ref(void *hp) {
    shear_handle *h = (shear_handle *)hp;
    h->sim->ref_init();
    h->sim->boundary();
}

void ref_shear_step_direct(void *hp, double dt, int nsteps) {
    shear_handle *h = (shear_handle *)hp;
    for (int q = 0; q < nsteps; q++) h->sim->step_forward(dt);
}

// l2_comparison (src/shear_sim_3d.cpp:3122-3158) of simulation a against b.
void ref_shear_l2_comparison(void *ap, void *bp, double *l2) {
    shear_handle *a = (shear_handle *)ap, *b = (shear_handle *)bp;
    l2[0] = l2[1] = l2[2] = 0;
    a->sim->l2_comparison(*b->sim, l2);
}

void ref_shear_get_ref(void *hp, double *xi) {
    shear_handle *h = (shear_handle *)hp;
    ref_shear *s = h->sim;
    int Nx = s->nx(), Ny = s->ny(), Nz = s->nz();
    int nzv = Nz + (h->periods[2] ? 0 : 1);
    for (int k = 0; k < nzv; k++)
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Nx; i++) {
                long q = i + (long)Nx * (j + (long)Ny * k);
                for (int a = 0; a < 3; a++) xi[3 * q + a] = s->xi(i, j, k, a);
            }
}

}  // extern "C"
