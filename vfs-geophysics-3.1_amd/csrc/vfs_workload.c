/*
 * Synthetic workloads for the fractional-step hot path (host code).
 *
 * The reference reads its grid from grid.dat (init.c:316-436) and its initial
 * field from a restart file or an analytic profile (bcs.c:3179-3400).  None of
 * those inputs ship with it, so this file generates them deterministically:
 *   - channel grids: uniform, tanh wall clustering in y, sinusoidal bottom;
 *   - a counter-based SplitMix64 RNG keyed by (seed, point, component), so
 *     every consumer of a field sees the same bits;
 *   - velocity fields: Poiseuille + noise, and a solenoidal random-Fourier
 *     field with a k^-5/3 energy spectrum;
 *   - Poisson right-hand sides with the aj-weighted mean removed.
 * All arrays use the DA layout: index (k*my + j)*mx + i, vectors as {x,y,z}.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>

#define IX(k, j, i) ((((size_t)(k)) * (size_t)my + (size_t)(j)) * (size_t)mx + (size_t)(i))

static uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

/* uniform in [-1, 1): 53 random bits keyed by (seed, counter) */
double vfs_wl_uniform(uint64_t seed, uint64_t counter)
{
    uint64_t r = splitmix64(splitmix64(seed) ^ counter);
    double u = (double)(r >> 11) * (1.0 / 9007199254740992.0);
    return 2.0 * u - 1.0;
}

/* Node coordinates of a channel of IM x JM x KM nodes on the DA layout
 * (mx = IM+1 ...).  y in [0, Ly]: uniform if beta <= 0, else
 * y = Ly/2 (1 + tanh(beta (2 eta - 1)) / tanh(beta)).  The bottom wall is
 * displaced by amp*sin(2 pi x / Lx) and the displacement fades linearly to
 * zero at the top wall: y += amp sin(2 pi x/Lx) (1 - eta). */
int vfs_wl_channel_grid(int IM, int JM, int KM, double Lx, double Ly, double Lz,
                        double beta, double amp, double *coor)
{
    const int mx = IM + 1, my = JM + 1, mz = KM + 1;
    memset(coor, 0, 3 * (size_t)mx * my * mz * sizeof(double));
    for (int k = 0; k < KM; k++)
        for (int j = 0; j < JM; j++)
            for (int i = 0; i < IM; i++) {
                double x = Lx * (double)i / (double)(IM - 1);
                double z = Lz * (double)k / (double)(KM - 1);
                double eta = (double)j / (double)(JM - 1);
                double y = beta > 0 ? 0.5 * Ly * (1.0 + tanh(beta * (2.0 * eta - 1.0)) / tanh(beta)) : Ly * eta;
                if (amp != 0.0) y += amp * sin(2.0 * M_PI * x / Lx) * (1.0 - eta);
                size_t id = IX(k, j, i);
                coor[3 * id] = x; coor[3 * id + 1] = y; coor[3 * id + 2] = z;
            }
    return 0;
}

/* ucont on every face from cell-centred Cartesian velocity: the face value is
 * the mean of the two adjacent cells dotted with the face area vector, the
 * same construction IB_BC uses (momentum.c:2088-2124).  Faces on no-slip
 * walls (non-periodic boundary faces) carry zero flux; the periodic ghost
 * faces are copies of their partners. */
static void faces_from_ucat(int mx, int my, int mz, const int per[3], const double *ucat,
                            const double *icsi, const double *jeta, const double *kzet, double *ucont)
{
    memset(ucont, 0, 3 * (size_t)mx * my * mz * sizeof(double));
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                if (!per[0] && i == mx - 2) continue;
                size_t a = IX(k, j, i), b = IX(k, j, (per[0] && i == mx - 2) ? 1 : i + 1);
                double u = (ucat[3 * a] + ucat[3 * b]) * 0.5, v = (ucat[3 * a + 1] + ucat[3 * b + 1]) * 0.5,
                       w = (ucat[3 * a + 2] + ucat[3 * b + 2]) * 0.5;
                ucont[3 * a] = u * icsi[3 * a] + v * icsi[3 * a + 1] + w * icsi[3 * a + 2];
            }
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                if (!per[1] && j == my - 2) continue;
                size_t a = IX(k, j, i), b = IX(k, (per[1] && j == my - 2) ? 1 : j + 1, i);
                double u = (ucat[3 * a] + ucat[3 * b]) * 0.5, v = (ucat[3 * a + 1] + ucat[3 * b + 1]) * 0.5,
                       w = (ucat[3 * a + 2] + ucat[3 * b + 2]) * 0.5;
                ucont[3 * a + 1] = u * jeta[3 * a] + v * jeta[3 * a + 1] + w * jeta[3 * a + 2];
            }
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                if (!per[2] && k == mz - 2) continue;
                size_t a = IX(k, j, i), b = IX((per[2] && k == mz - 2) ? 1 : k + 1, j, i);
                double u = (ucat[3 * a] + ucat[3 * b]) * 0.5, v = (ucat[3 * a + 1] + ucat[3 * b + 1]) * 0.5,
                       w = (ucat[3 * a + 2] + ucat[3 * b + 2]) * 0.5;
                ucont[3 * a + 2] = u * kzet[3 * a] + v * kzet[3 * a + 1] + w * kzet[3 * a + 2];
            }
    /* periodic partners: face 0 == face m-2, face m-1 == face 1 */
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++) {
            if (per[0]) { ucont[3 * IX(k, j, 0)] = ucont[3 * IX(k, j, mx - 2)]; ucont[3 * IX(k, j, mx - 1)] = ucont[3 * IX(k, j, 1)]; }
        }
    for (int k = 0; k < mz; k++)
        for (int i = 0; i < mx; i++) {
            if (per[1]) { ucont[3 * IX(k, 0, i) + 1] = ucont[3 * IX(k, my - 2, i) + 1]; ucont[3 * IX(k, my - 1, i) + 1] = ucont[3 * IX(k, 1, i) + 1]; }
        }
    for (int j = 0; j < my; j++)
        for (int i = 0; i < mx; i++) {
            if (per[2]) { ucont[3 * IX(0, j, i) + 2] = ucont[3 * IX(mz - 2, j, i) + 2]; ucont[3 * IX(mz - 1, j, i) + 2] = ucont[3 * IX(1, j, i) + 2]; }
        }
}

/* Poiseuille channel profile u = 1 - (y/h - 1)^2 (h = Ly/2) plus U[-1,1)
 * noise of amplitude `noise` on all three components at interior fluid cells
 * (nvert < 0.1); cells with nvert > 0.1 get zero velocity. */
int vfs_wl_poiseuille(int mx, int my, int mz, const int per[3], double Ly, double noise,
                      uint64_t seed, const double *cent, const double *nvert,
                      const double *icsi, const double *jeta, const double *kzet,
                      double *ucat, double *ucont)
{
    memset(ucat, 0, 3 * (size_t)mx * my * mz * sizeof(double));
    const double h = 0.5 * Ly;
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t id = IX(k, j, i);
                if (nvert && nvert[id] > 0.1) continue;
                double y = cent[3 * id + 1] / h - 1.0;
                ucat[3 * id] = 1.0 - y * y + noise * vfs_wl_uniform(seed, 3 * id + 0);
                ucat[3 * id + 1] = noise * vfs_wl_uniform(seed, 3 * id + 1);
                ucat[3 * id + 2] = noise * vfs_wl_uniform(seed, 3 * id + 2);
            }
    faces_from_ucat(mx, my, mz, per, ucat, icsi, jeta, kzet, ucont);
    return 0;
}

/* Log-law streamwise-k profile of the turbine-wake case: w = (1/0.4)
 * ln(y/y0) u* with u* = 1 / ((1/0.4) ln(Ly/y0)), so that w(Ly) = 1 (y is
 * clamped to at least y0), plus U[-1,1) noise of amplitude `noise` on all
 * three components at interior fluid cells; nvert > 0.1 gets zero. */
int vfs_wl_loglaw_k(int mx, int my, int mz, const int per[3], double Ly, double y0, double noise, uint64_t seed,
                    const double *cent, const double *nvert, const double *icsi, const double *jeta,
                    const double *kzet, double *ucat, double *ucont)
{
    memset(ucat, 0, 3 * (size_t)mx * my * mz * sizeof(double));
    const double us = 1.0 / ((1.0 / 0.4) * log(Ly / y0));
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t id = IX(k, j, i);
                if (nvert && nvert[id] > 0.1) continue;
                const double y = cent[3 * id + 1] > y0 ? cent[3 * id + 1] : y0;
                ucat[3 * id] = noise * vfs_wl_uniform(seed, 3 * id + 0);
                ucat[3 * id + 1] = noise * vfs_wl_uniform(seed, 3 * id + 1);
                ucat[3 * id + 2] = (1.0 / 0.4) * log(y / y0) * us + noise * vfs_wl_uniform(seed, 3 * id + 2);
            }
    faces_from_ucat(mx, my, mz, per, ucat, icsi, jeta, kzet, ucont);
    return 0;
}

/* Solenoidal random-Fourier velocity: u = sum_m A_m cos(k_m . x + phi_m) with
 * A_m perpendicular to k_m, |A_m| ~ |k_m|^(-11/6) (energy ~ k^-5/3),
 * integer mode numbers with 1 <= |n| <= nmax (kx = 2 pi nx/Lx, ky = pi ny/Ly,
 * kz = 2 pi nz/Lz), then multiplied by the wall envelope (y/h)(2 - y/h) and
 * scaled to an rms of `urms`.  A mean streamwise profile 1 - (y/h-1)^2 is
 * added.  Cells with nvert > 0.1 get zero velocity. */
int vfs_wl_turbulent(int mx, int my, int mz, const int per[3], double Lx, double Ly, double Lz,
                     int nmodes, int nmax, double urms, uint64_t seed, const double *cent,
                     const double *nvert, const double *icsi, const double *jeta,
                     const double *kzet, double *ucat, double *ucont)
{
    double *K = (double *)malloc(sizeof(double) * 7 * (size_t)nmodes);
    if (!K) return -1;
    uint64_t ctr = 0;
    for (int m = 0; m < nmodes; m++) {
        int nx, ny, nz;
        double n2;
        do {
            nx = (int)lround(nmax * vfs_wl_uniform(seed, ctr++));
            ny = (int)lround(nmax * vfs_wl_uniform(seed, ctr++));
            nz = (int)lround(nmax * vfs_wl_uniform(seed, ctr++));
            n2 = (double)nx * nx + (double)ny * ny + (double)nz * nz;
        } while (n2 < 1.0 || n2 > (double)nmax * nmax);
        double kx = 2 * M_PI * nx / Lx, ky = M_PI * ny / Ly, kz = 2 * M_PI * nz / Lz;
        double kk = sqrt(kx * kx + ky * ky + kz * kz);
        double r[3] = {vfs_wl_uniform(seed, ctr++), vfs_wl_uniform(seed, ctr++), vfs_wl_uniform(seed, ctr++)};
        double rk = (r[0] * kx + r[1] * ky + r[2] * kz) / (kk * kk);
        double a[3] = {r[0] - rk * kx, r[1] - rk * ky, r[2] - rk * kz};
        double an = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]) + 1e-30;
        double amp = pow(sqrt(n2), -11.0 / 6.0) / an;
        K[7 * m + 0] = kx; K[7 * m + 1] = ky; K[7 * m + 2] = kz;
        K[7 * m + 3] = a[0] * amp; K[7 * m + 4] = a[1] * amp; K[7 * m + 5] = a[2] * amp;
        K[7 * m + 6] = M_PI * vfs_wl_uniform(seed, ctr++);
    }
    memset(ucat, 0, 3 * (size_t)mx * my * mz * sizeof(double));
    const double h = 0.5 * Ly;
    double ss = 0;
    size_t cnt = 0;
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t id = IX(k, j, i);
                if (nvert && nvert[id] > 0.1) continue;
                double x = cent[3 * id], y = cent[3 * id + 1], z = cent[3 * id + 2];
                double u[3] = {0, 0, 0};
                for (int m = 0; m < nmodes; m++) {
                    double ph = cos(K[7 * m] * x + K[7 * m + 1] * y + K[7 * m + 2] * z + K[7 * m + 6]);
                    u[0] += K[7 * m + 3] * ph; u[1] += K[7 * m + 4] * ph; u[2] += K[7 * m + 5] * ph;
                }
                double yy = y / h;
                double env = yy * (2.0 - yy);
                if (env < 0) env = 0;
                for (int q = 0; q < 3; q++) { ucat[3 * id + q] = u[q] * env; ss += u[q] * env * u[q] * env; }
                cnt++;
            }
    double rms = sqrt(ss / (3.0 * (double)(cnt ? cnt : 1)));
    double sc = rms > 0 ? urms / rms : 0;
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t id = IX(k, j, i);
                if (nvert && nvert[id] > 0.1) continue;
                double y = cent[3 * id + 1] / h - 1.0;
                ucat[3 * id] = ucat[3 * id] * sc + (1.0 - y * y);
                ucat[3 * id + 1] *= sc;
                ucat[3 * id + 2] *= sc;
            }
    free(K);
    faces_from_ucat(mx, my, mz, per, ucat, icsi, jeta, kzet, ucont);
    return 0;
}

/* Random field U[-1,1) * amp on a scalar DA array at interior cells with
 * nvert < 0.1; if aj is given, the aj-weighted mean is removed the way
 * Remove_Nullspace_Scale does (poisson_hypre.c:163-172). */
int vfs_wl_random_scalar(int mx, int my, int mz, double amp, uint64_t seed,
                         const double *nvert, const double *aj, double *out)
{
    memset(out, 0, (size_t)mx * my * mz * sizeof(double));
    double sum = 0, sum_aj = 0;
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t id = IX(k, j, i);
                if (nvert && !(nvert[id] < 0.1)) continue;
                out[id] = amp * vfs_wl_uniform(seed, id);
                sum += out[id];
                if (aj) sum_aj += aj[id];
            }
    if (aj && sum_aj != 0)
        for (int k = 1; k < mz - 1; k++)
            for (int j = 1; j < my - 1; j++)
                for (int i = 1; i < mx - 1; i++) {
                    size_t id = IX(k, j, i);
                    if (nvert && !(nvert[id] < 0.1)) continue;
                    out[id] += -sum * aj[id] / sum_aj;
                }
    return 0;
}

/* Random vector field U[-1,1) * amp at every point (ghosts included). */
int vfs_wl_random_vector(int mx, int my, int mz, double amp, uint64_t seed, double *out)
{
    size_t n = 3 * (size_t)mx * my * mz;
    for (size_t q = 0; q < n; q++) out[q] = amp * vfs_wl_uniform(seed, q);
    return 0;
}
