/* CPU oracle (test infrastructure) of the random initial chi field.
 *
 * Restates, in plain C, the reference's Gaussian random field
 * (shear_sim_3d::set_random_initial_conditions, src/shear_sim_3d.cpp:
 * 2793-2934): a noise grid padded by `cut` points on each side of every
 * non-periodic axis, convolved by brute force over the cube [-cut, cut]^3
 * with weights exp(-llinv (i^2 + j^2 + k^2)) (convolve, :2997-3041; wrap,
 * :3046-3054), then chi = chi_avg + chi_1 conv with chi_1 = sigma divided by
 * gauss_normal_factor_long (:3071-3078, applied in the constructor at :78).
 *
 * The reference draws its noise with glibc rand() and the polar Box-Muller
 * method (box_muller, :3057-3067; fill_noise_grid, :2978-2992).  The noise
 * generator here is this library's own, seeded and counter-based:
 * Philox4x64-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC 2011) feeding the same polar method (with log(r)
 * taken as log1p(r - 1) for r >= 1/2, accurate where log(r) is small), one
 * counter per noise pair and attempt, so every pair can be drawn
 * independently on a GPU.
 */
#include <math.h>
#include <stdint.h>

static inline void mulhilo64(uint64_t a, uint64_t b, uint64_t *hi, uint64_t *lo) {
    const unsigned __int128 p = (unsigned __int128)a * b;
    *hi = (uint64_t)(p >> 64);
    *lo = (uint64_t)p;
}

/* Philox4x64 with 10 rounds: out = block(ctr, key). */
void grf_philox4x64(const uint64_t ctr_in[4], const uint64_t key_in[2], uint64_t out[4]) {
    uint64_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint64_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B97F4A7C15ULL;
            k1 += 0xBB67AE8584CAA73BULL;
        }
        uint64_t hi0, lo0, hi1, lo1;
        mulhilo64(0xD2E7470EE14C6C93ULL, c0, &hi0, &lo0);
        mulhilo64(0xCA5A826395121157ULL, c2, &hi1, &lo1);
        const uint64_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

#define GRF_KEY1 0x53545A2B2B475246ULL /* second key word, fixed */
#define GRF_ATTEMPTS 64                  /* polar-method attempts per pair */

/* Pair p of the noise stream: two standard normals. */
void grf_noise_pair(uint64_t seed, uint64_t p, double *r0, double *r1) {
    const uint64_t key[2] = {seed, GRF_KEY1};
    for (uint64_t a = 0; a < GRF_ATTEMPTS; a++) {
        const uint64_t ctr[4] = {p, a, 0, 0};
        uint64_t o[4];
        grf_philox4x64(ctr, key, o);
        const double x = 2.0 * ((double)(o[0] >> 11) * 0x1.0p-53) - 1.0;
        const double y = 2.0 * ((double)(o[1] >> 11) * 0x1.0p-53) - 1.0;
        const double r = x * x + y * y;
        if (r == 0.0 || r > 1.0) continue;
        /* log(r) through log1p(r - 1) for r >= 1/2 (r - 1 exact there) */
        const double lr = r >= 0.5 ? log1p(r - 1.0) : log(r);
        const double d = sqrt(-2.0 * lr / r);
        *r0 = x * d;
        *r1 = y * d;
        return;
    }
    *r0 = 0.0;
    *r1 = 0.0;
}

/* n noise values: element 2p and 2p+1 from pair p (the last pair of an odd
 * count keeps its first value, as fill_noise_grid does). */
void grf_noise(uint64_t seed, long n, double *out) {
    for (long p = 0; 2 * p < n; p++) {
        double a, b;
        grf_noise_pair(seed, (uint64_t)p, &a, &b);
        out[2 * p] = a;
        if (2 * p + 1 < n) out[2 * p + 1] = b;
    }
}

/* gauss_normal_factor_long (src/shear_sim_3d.cpp:3071-3078), same loop order. */
double grf_normal_factor(int cut, double llinv) {
    double g = 0;
    for (int ii = -cut; ii <= cut; ii++)
        for (int jj = -cut; jj <= cut; jj++)
            for (int kk = -cut; kk <= cut; kk++)
                g += exp(-llinv * (ii * ii + jj * jj + kk * kk)) * exp(-llinv * (ii * ii + jj * jj + kk * kk));
    return sqrt(g);
}

static inline int grf_wrap(int i, int d, int n, int period) {
    const int v = i + d;
    if (period && v > n - 1) return v - n;
    if (period && v < 0) return n + v;
    return v;
}

/* chi on the gN grid (x fastest) from the padded noise grid.  chi_1 is the
 * scale after normalisation (sigma / grf_normal_factor). */
void grf_convolve(const int gN[3], const int prd[3], int cut, double llinv, double chi_avg, double chi_1,
                  const double *noise, double *out) {
    const int nxt = prd[0] ? gN[0] : gN[0] + 2 * cut;
    const int nyt = prd[1] ? gN[1] : gN[1] + 2 * cut;
    const long off = (prd[0] ? 0 : cut) + (long)(prd[1] ? 0 : nxt * cut) + (long)(prd[2] ? 0 : (long)nxt * nyt * cut);
    const double *no = noise + off;
#pragma omp parallel for
    for (int zz = 0; zz < gN[2]; zz++)
        for (int yy = 0; yy < gN[1]; yy++)
            for (int xx = 0; xx < gN[0]; xx++) {
                double cv = 0;
                for (int kk = -cut; kk <= cut; kk++)
                    for (int jj = -cut; jj <= cut; jj++)
                        for (int ii = -cut; ii <= cut; ii++) {
                            const int xw = grf_wrap(xx, ii, gN[0], prd[0]);
                            const int yw = grf_wrap(yy, jj, gN[1], prd[1]);
                            const int zw = grf_wrap(zz, kk, gN[2], prd[2]);
                            const long ind = xw + (long)yw * nxt + (long)zw * nxt * nyt;
                            cv += no[ind] * exp(-llinv * (ii * ii + jj * jj + kk * kk));
                        }
                out[xx + (long)gN[0] * (yy + (long)gN[1] * zz)] = chi_avg + chi_1 * cv;
            }
}

/* The brute-force convolution of grf_convolve at npts chosen cells only
 * (flat indices idx, x fastest), for checks at sizes where the full
 * brute force is out of reach. */
void grf_convolve_points(const int gN[3], const int prd[3], int cut, double llinv, double chi_avg, double chi_1,
                         const double *noise, long npts, const long *idx, double *out) {
    const int nxt = prd[0] ? gN[0] : gN[0] + 2 * cut;
    const int nyt = prd[1] ? gN[1] : gN[1] + 2 * cut;
    const long off = (prd[0] ? 0 : cut) + (long)(prd[1] ? 0 : nxt * cut) + (long)(prd[2] ? 0 : (long)nxt * nyt * cut);
    const double *no = noise + off;
#pragma omp parallel for
    for (long q = 0; q < npts; q++) {
        const int xx = (int)(idx[q] % gN[0]);
        const int yy = (int)((idx[q] / gN[0]) % gN[1]);
        const int zz = (int)(idx[q] / ((long)gN[0] * gN[1]));
        double cv = 0;
        for (int kk = -cut; kk <= cut; kk++)
            for (int jj = -cut; jj <= cut; jj++)
                for (int ii = -cut; ii <= cut; ii++) {
                    const int xw = grf_wrap(xx, ii, gN[0], prd[0]);
                    const int yw = grf_wrap(yy, jj, gN[1], prd[1]);
                    const int zw = grf_wrap(zz, kk, gN[2], prd[2]);
                    const long ind = xw + (long)yw * nxt + (long)zw * nxt * nyt;
                    cv += no[ind] * exp(-llinv * (ii * ii + jj * jj + kk * kk));
                }
        out[q] = chi_avg + chi_1 * cv;
    }
}
