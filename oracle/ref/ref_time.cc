// Times the reference's own multigrid V-cycle (test infrastructure: the CPU
// baseline).  Built by oracle/ref/Makefile from the reference sources where
// they lie into oracle/_ref/ref_time; run under mpiexec with P ranks.
//
//   ref_time shear|trans N reps dt mu lamb sy tau0 c0 eps0 delta omega
//            bath_temp chi_inf ez TZ diff_l U
//
// shear: shear_sim_3d<sym_mat3> on N x N x N cells, x and y periodic, z
//        walled (multigrid N x N x (N+1));
// trans: trans_sim_3d on N x N x N cells, periodic (multigrid N^3).
// The P ranks split the grid by the reference's calc_processor_division
// (src/common.cpp:7-60) as its drivers do (src/parallel_scale.cpp:86-95).
// Each rank fills its velocities and stresses with a deterministic hash of
// the global index, then the reference builds its stencil and Galerkin
// hierarchy (set_up_stencil_and_matrices) and its projection fields
// (multigrid::setup_fields, src/shear_sim_3d.cpp:551), and `reps` V-cycles
// (multigrid::v_cycle, the call timed at src/shear_sim_3d.cpp:567-570) are
// timed between barriers.  Rank 0 prints one JSON line.
#include <mpi.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "common.hh"
#include "shear_sim_3d.hh"
#include "trans_sim_3d.hh"

namespace {

double hash_unit(uint64_t g, int comp) {
    uint64_t z = g * 0x9E3779B97F4A7C15ull + (uint64_t)comp * 0xBF58476D1CE4E5B9ull + 0x94D049BB133111EBull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) / 4503599627370496.0 - 1.0;  // [-1, 1)
}

template <class SIM>
struct timed : public SIM {
    using SIM::SIM;
    // velocities 1e-3 and stresses 0.1 times the hash, chi 0.09, on every
    // vertex plane this rank holds (k <= lNz covers the top wall plane)
    void fill(const int *coords, int N, int Nz) {
        const int lx = this->N_x, ly = this->N_y, lz = this->N_z;
        for (int k = 0; k <= lz; k++)
            for (int j = 0; j < ly; j++)
                for (int i = 0; i < lx; i++) {
                    const uint64_t gi = (uint64_t)coords[0] * lx + i, gj = (uint64_t)coords[1] * ly + j,
                                   gk = (uint64_t)coords[2] * lz + k;
                    const uint64_t g = gi + (uint64_t)N * (gj + (uint64_t)N * gk);
                    Field &f = this->grido[i + this->nxg * j + this->nxyg * k];
                    f.u = 1e-3 * hash_unit(g, 0);
                    f.v = 1e-3 * hash_unit(g, 1);
                    f.w = 1e-3 * hash_unit(g, 2);
                    f.s11 = 0.1 * hash_unit(g, 3);
                    f.s12 = 0.1 * hash_unit(g, 4);
                    f.s13 = 0.1 * hash_unit(g, 5);
                    f.s22 = 0.1 * hash_unit(g, 6);
                    f.s23 = 0.1 * hash_unit(g, 7);
                    f.s33 = 0.1 * hash_unit(g, 8);
                    f.chi = 0.09;
                }
        (void)Nz;
        this->set_up_ghost_regions();
        this->set_boundary_conditions();
    }
    void setup(double dt, bool base) {
        if (base) this->set_up_stencil_base(dt);
        this->set_up_stencil_and_matrices(dt);
    }
    void fields() { this->mgrid.setup_fields(*this); }
    void vcycle() { this->mgrid.v_cycle(); }
    double quartic() { return (*this->mgrid.reg)->l2_error_all(); }
    int nlevels() { return this->mgrid.nlevels; }
};

template <class T>
int run(T &sim, const char *kind, int P, const int *dims, const int *coords, int N, int Nz, int reps,
        double dt, bool base, int rank, const int *mg) {
    sim.fill(coords, N, Nz);
    MPI_Barrier(MPI_COMM_WORLD);
    double t0 = MPI_Wtime();
    sim.setup(dt, base);
    MPI_Barrier(MPI_COMM_WORLD);
    const double t_setup = MPI_Wtime() - t0;
    sim.fields();
    const double res0 = sim.quartic();
    sim.vcycle();  // untimed warm-up cycle
    MPI_Barrier(MPI_COMM_WORLD);
    t0 = MPI_Wtime();
    for (int q = 0; q < reps; q++) sim.vcycle();
    MPI_Barrier(MPI_COMM_WORLD);
    const double t_cyc = (MPI_Wtime() - t0) / reps;
    const double res = sim.quartic();
    if (rank == 0)
        std::printf("{\"case\": \"%s\", \"ranks\": %d, \"dims\": [%d, %d, %d], \"mg\": [%d, %d, %d], "
                    "\"levels\": %d, \"setup_s\": %.6g, \"vcycle_s\": %.6g, \"reps\": %d, "
                    "\"quartic0\": %.6g, \"quartic\": %.6g}\n",
                    kind, P, dims[0], dims[1], dims[2], mg[0], mg[1], mg[2], sim.nlevels(), t_setup, t_cyc,
                    reps, res0, res);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    MPI_Init(&argc, &argv);
    int P = 1, wrank = 0;
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    MPI_Comm_rank(MPI_COMM_WORLD, &wrank);
    if (argc != 19) {
        if (wrank == 0)
            std::fprintf(stderr,
                         "usage: ref_time shear|trans N reps dt mu lamb sy tau0 c0 eps0 delta omega "
                         "bath_temp chi_inf ez TZ diff_l U\n");
        MPI_Finalize();
        return 2;
    }
    const bool trans = std::strcmp(argv[1], "trans") == 0;
    const int N = std::atoi(argv[2]), reps = std::atoi(argv[3]);
    double a[15];
    for (int q = 0; q < 15; q++) a[q] = std::atof(argv[4 + q]);
    const double dt = a[0], mu = a[1], lamb = a[2], sy = a[3], tau0 = a[4], c0 = a[5], eps0 = a[6],
                 delta = a[7], omega = a[8], bath = a[9], chi_inf = a[10], ez = a[11], TZ = a[12],
                 diff_l = a[13], U = a[14];
    int periods[3] = {1, 1, trans ? 1 : 0};
    int dims[3] = {1, 1, 1};
    calc_processor_division(P, N, N, N, trans ? 7u : 3u, dims);
    MPI_Comm cart;
    MPI_Cart_create(MPI_COMM_WORLD, 3, dims, periods, 1, &cart);
    int rank = 0, coords[3] = {0, 0, 0};
    MPI_Comm_rank(cart, &rank);
    MPI_Cart_coords(cart, rank, 3, coords);
    const int lx = N / dims[0], ly = N / dims[1], lz = N / dims[2];
    const std::string out = "ref_time.odr";
    const int mg[3] = {N, N, trans ? N : N + 1};
    int rc;
    if (trans) {
        timed<trans_sim_3d> sim(lx, ly, lz, -1, -1, -0.5, 1, 1, 0.5, 0.0, 1.0, 0.5, 1, out, mu, lamb, mu, 0.0, sy,
                                tau0, c0, eps0, delta, omega, bath, chi_inf, ez, TZ, diff_l, 0.0, U, 5, 3, 0.09,
                                0.01, 1.0, &cart, dims, 15, periods);
        rc = run(sim, "trans", P, dims, coords, N, N, reps, dt, false, rank, mg);
    } else {
        timed<shear_sim_3d<sym_mat3>> sim(lx, ly, lz, -1, -1, -0.5, 1, 1, 0.5, 0.0, 1.0, 0.5, 1, out, mu, lamb, mu,
                                          0.0, sy, tau0, c0, eps0, delta, omega, bath, chi_inf, ez, TZ, diff_l, U,
                                          5, 3, 0.09, 0.01, 1.0, &cart, dims, 15, 0.0, 0.0, periods);
        rc = run(sim, "shear", P, dims, coords, N, N, reps, dt, true, rank, mg);
    }
    MPI_Comm_free(&cart);
    MPI_Finalize();
    return rc;
}
