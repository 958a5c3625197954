/*
 * CPU oracle: single-threaded FP64 restatement of the VFS-Geophysics
 * fractional-step hot path.  TEST INFRASTRUCTURE ONLY (see vfs_oracle.h).
 *
 * Every routine follows the reference's loop order and operation order
 * (file:line cited per function), so that it reproduces the reference's
 * floating-point results on the same inputs.  Branches for options that no
 * supported configuration exercises (levelset, rans, skew, clark,
 * inflow/outflow bctypes, ii/jj/kk_periodic) are rejected with a negative
 * return code instead of being restated.
 *
 * Build: gcc -O2 -ffp-contract=off -fPIC -shared (oracle/Makefile).  No FMA
 * contraction, so every a*b+c rounds twice exactly as the reference does.
 */
#include "vfs_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define DIMS  const int mx = c->mx, my = c->my, mz = c->mz; (void)mx; (void)my; (void)mz
#define IX(k, j, i) ((((size_t)(k)) * (size_t)my + (size_t)(j)) * (size_t)mx + (size_t)(i))
#define VX(a, id) ((a)[3 * (id) + 0])
#define VY(a, id) ((a)[3 * (id) + 1])
#define VZ(a, id) ((a)[3 * (id) + 2])

/* Periodic ghost source (the `flag/a/b/c` idiom, e.g. momentum.c:642-669).
 * Sources are never themselves flagged, so in-place copies are order-free. */
static int per_src(const vo_cfg *c, int i, int j, int k, int *a, int *b, int *cc)
{
    DIMS;
    int flag = 0;
    *a = i; *b = j; *cc = k;
    if (c->i_periodic && i == 0) *a = mx - 2, flag = 1;
    else if (c->i_periodic && i == mx - 1) *a = 1, flag = 1;
    if (c->j_periodic && j == 0) *b = my - 2, flag = 1;
    else if (c->j_periodic && j == my - 1) *b = 1, flag = 1;
    if (c->k_periodic && k == 0) *cc = mz - 2, flag = 1;
    else if (c->k_periodic && k == mz - 1) *cc = 1, flag = 1;
    return flag;
}

static void per_copy_vec(const vo_cfg *c, double *v)
{
    DIMS;
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                int a, b, cc;
                if (per_src(c, i, j, k, &a, &b, &cc)) {
                    size_t d = IX(k, j, i), s = IX(cc, b, a);
                    VX(v, d) = VX(v, s); VY(v, d) = VY(v, s); VZ(v, d) = VZ(v, s);
                }
            }
}

static void per_copy_sca(const vo_cfg *c, double *v)
{
    DIMS;
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                int a, b, cc;
                if (per_src(c, i, j, k, &a, &b, &cc)) v[IX(k, j, i)] = v[IX(cc, b, a)];
            }
}

/* Supported boundary set: every face is either no-slip (1) or periodic (7,
 * with the matching i/j/k_periodic flag). */
static int check_cfg(const vo_cfg *c)
{
    if (c->mx < 7 || c->my < 7 || c->mz < 7) return -2;
    const int per[3] = {c->i_periodic, c->j_periodic, c->k_periodic};
    for (int f = 0; f < 6; f++) {
        int bt = c->bctype[f];
        if (per[f / 2]) { if (bt != 7) return -3; }
        else if (bt != 1 && !(bt == -1 && f < 4) && !(bt == 5 && f == 4) && !(bt == 4 && f == 5))
            return -3;   /* -1: wall function (i/j walls); 5 / 4: a k inlet / outlet */
    }
    if (c->immersed != 0 && c->immersed != 1) return -4;
    if (c->weno != 0 && c->weno != 1) return -4;
    if (c->inviscid != 0 && c->inviscid != 1) return -4;
    if (c->skew != 0 && c->skew != 1) return -4;
    if (c->clark != 0 && c->clark != 1) return -4;
    return 0;
}

/* threads the oracle's parallel loops use (1 without OpenMP).  Every such
 * loop writes only its own points, so results do not depend on it. */
int vo_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* level.c:1734-1771 weno3.  pow(x, 2.) of the reference is written x * x:
 * the reference is built with -O3 (makefile:19), where GCC folds pow(x, 2.)
 * into that multiplication. */
static double sq(double x) { return x * x; }

double vo_weno3(double f0, double f1, double f2, double f3, double wavespeed)
{
    double fL, fC, fR;
    if (wavespeed > 0) {
        fL = f0; fC = f1; fR = f2;
    } else {
        fL = f3; fC = f2; fR = f1;   /* mirror */
    }
    const double d0 = 2. / 3., d1 = 1. / 3.;
    const double eps = 1.e-6;
    double beta0 = sq(fC - fR);
    double beta1 = sq(fL - fC);
    double alpha0 = d0 / sq(eps + beta0);
    double alpha1 = d1 / sq(eps + beta1);
    double sumalpha = alpha0 + alpha1;
    double w0 = alpha0 / sumalpha;
    double w1 = alpha1 / sumalpha;
    double u0 = (fC * 0.5 + fR * 0.5);
    double u1 = (-fL * 0.5 + fC * 1.5);
    return w0 * u0 + w1 * u1;
}

static void copy3(double *d, size_t di, const double *s, size_t si)
{
    d[3 * di] = s[3 * si]; d[3 * di + 1] = s[3 * si + 1]; d[3 * di + 2] = s[3 * si + 2];
}

/* ------------------------------------------------------------------------ */
/* metrics.c:4-1346 FormMetrics, NEWMETRIC branch                           */
/* ------------------------------------------------------------------------ */
int vo_form_metrics(const vo_cfg *c, const double *coor, vo_geom *g)
{
    DIMS;
    int rc = check_cfg(c);
    if (rc) return rc;
#define CO(k, j, i, comp) coor[3 * IX(k, j, i) + (comp)]
    /* cell centres, metrics.c:80-99 */
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++)
                for (int q = 0; q < 3; q++)
                    g->cent[3 * IX(k, j, i) + q] = 0.125 *
                        (CO(k, j, i, q) + CO(k, j - 1, i, q) + CO(k - 1, j, i, q) + CO(k - 1, j - 1, i, q) +
                         CO(k, j, i - 1, q) + CO(k, j - 1, i - 1, q) + CO(k - 1, j, i - 1, q) + CO(k - 1, j - 1, i - 1, q));

    /* cell Jacobian and csi/eta/zet, metrics.c:186-251 */
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                double d[3][3]; /* d[comp][dir]: dir 0=c, 1=e, 2=z */
                for (int q = 0; q < 3; q++) {
                    d[q][0] = 0.25 * (CO(k, j, i, q) + CO(k, j - 1, i, q) + CO(k - 1, j, i, q) + CO(k - 1, j - 1, i, q) -
                                      CO(k, j, i - 1, q) - CO(k, j - 1, i - 1, q) - CO(k - 1, j, i - 1, q) - CO(k - 1, j - 1, i - 1, q));
                    d[q][1] = 0.25 * (CO(k, j, i, q) + CO(k, j, i - 1, q) + CO(k - 1, j, i, q) + CO(k - 1, j, i - 1, q) -
                                      CO(k, j - 1, i, q) - CO(k, j - 1, i - 1, q) - CO(k - 1, j - 1, i, q) - CO(k - 1, j - 1, i - 1, q));
                    d[q][2] = 0.25 * (CO(k, j, i, q) + CO(k, j - 1, i, q) + CO(k, j, i - 1, q) + CO(k, j - 1, i - 1, q) -
                                      CO(k - 1, j, i, q) - CO(k - 1, j - 1, i, q) - CO(k - 1, j, i - 1, q) - CO(k - 1, j - 1, i - 1, q));
                }
                double dxdc = d[0][0], dydc = d[1][0], dzdc = d[2][0];
                double dxde = d[0][1], dyde = d[1][1], dzde = d[2][1];
                double dxdz = d[0][2], dydz = d[1][2], dzdz = d[2][2];
                size_t id = IX(k, j, i);
                double a = dxdc * (dyde * dzdz - dzde * dydz) - dydc * (dxde * dzdz - dzde * dxdz) +
                           dzdc * (dxde * dydz - dyde * dxdz);
                g->aj[id] = 1. / a;
                VX(g->csi, id) = dyde * dzdz - dzde * dydz;
                VY(g->csi, id) = -dxde * dzdz + dzde * dxdz;
                VZ(g->csi, id) = dxde * dydz - dyde * dxdz;
                VX(g->eta, id) = dydz * dzdc - dzdz * dydc;
                VY(g->eta, id) = -dxdz * dzdc + dzdz * dxdc;
                VZ(g->eta, id) = dxdz * dydc - dydz * dxdc;
                VX(g->zet, id) = dydc * dzde - dzdc * dyde;
                VY(g->zet, id) = -dxdc * dzde + dzdc * dxde;
                VZ(g->zet, id) = dxdc * dyde - dydc * dxde;
            }
#undef CO
    /* boundary mirror, metrics.c:253-334, in the reference's plane order */
    for (int side = 0; side < 6; side++) {
        for (int k = 0; k < mz; k++)
            for (int j = 0; j < my; j++)
                for (int i = 0; i < mx; i++) {
                    int si = i, sj = j, sk = k;
                    if (side == 0) { if (i != 0) continue; si = 1; }
                    if (side == 1) { if (i != mx - 1) continue; si = mx - 2; }
                    if (side == 2) { if (j != 0) continue; sj = 1; }
                    if (side == 3) { if (j != my - 1) continue; sj = my - 2; }
                    if (side == 4) { if (k != 0) continue; sk = 1; }
                    if (side == 5) { if (k != mz - 1) continue; sk = mz - 2; }
                    size_t d = IX(k, j, i), s = IX(sk, sj, si);
                    copy3(g->csi, d, g->csi, s); copy3(g->eta, d, g->eta, s);
                    copy3(g->zet, d, g->zet, s); g->aj[d] = g->aj[s];
                }
    }
    /* periodic ghosts of the cell metrics, metrics.c:354-389 */
    per_copy_vec(c, g->csi); per_copy_vec(c, g->eta); per_copy_vec(c, g->zet); per_copy_sca(c, g->aj);

    /* face metrics: averages of the two cells, harmonic-mean Jacobian
     * (metrics.c:581-584 i-faces, 689-692 j-faces, 797-800 k-faces) */
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 0; i < mx - 1; i++) {
                size_t id = IX(k, j, i), ip = IX(k, j, i + 1);
                for (int q = 0; q < 3; q++) {
                    g->icsi[3 * id + q] = 0.5 * g->csi[3 * id + q] + 0.5 * g->csi[3 * ip + q];
                    g->ieta[3 * id + q] = 0.5 * g->eta[3 * id + q] + 0.5 * g->eta[3 * ip + q];
                    g->izet[3 * id + q] = 0.5 * g->zet[3 * id + q] + 0.5 * g->zet[3 * ip + q];
                }
                g->iaj[id] = 2. / (1. / g->aj[id] + 1. / g->aj[ip]);
            }
    for (int k = 1; k < mz - 1; k++)
        for (int j = 0; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t id = IX(k, j, i), jp = IX(k, j + 1, i);
                for (int q = 0; q < 3; q++) {
                    g->jcsi[3 * id + q] = 0.5 * g->csi[3 * id + q] + 0.5 * g->csi[3 * jp + q];
                    g->jeta[3 * id + q] = 0.5 * g->eta[3 * id + q] + 0.5 * g->eta[3 * jp + q];
                    g->jzet[3 * id + q] = 0.5 * g->zet[3 * id + q] + 0.5 * g->zet[3 * jp + q];
                }
                g->jaj[id] = 2. / (1. / g->aj[id] + 1. / g->aj[jp]);
            }
    for (int k = 0; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t id = IX(k, j, i), kp = IX(k + 1, j, i);
                for (int q = 0; q < 3; q++) {
                    g->kcsi[3 * id + q] = 0.5 * g->csi[3 * id + q] + 0.5 * g->csi[3 * kp + q];
                    g->keta[3 * id + q] = 0.5 * g->eta[3 * id + q] + 0.5 * g->eta[3 * kp + q];
                    g->kzet[3 * id + q] = 0.5 * g->zet[3 * id + q] + 0.5 * g->zet[3 * kp + q];
                }
                g->kaj[id] = 2. / (1 / g->aj[id] + 1 / g->aj[kp]);
            }
    /* final periodic copy of all sixteen arrays, metrics.c:1214-1250 */
    if (c->i_periodic || c->j_periodic || c->k_periodic) {
        double *v[12] = {g->csi, g->eta, g->zet, g->icsi, g->ieta, g->izet,
                         g->jcsi, g->jeta, g->jzet, g->kcsi, g->keta, g->kzet};
        double *s[4] = {g->aj, g->iaj, g->jaj, g->kaj};
        for (int n = 0; n < 12; n++) per_copy_vec(c, v[n]);
        for (int n = 0; n < 4; n++) per_copy_sca(c, s[n]);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* ibm.c:319-341, 508-536: mask from an analytic sphere                    */
/* ------------------------------------------------------------------------ */
int vo_nvert_sphere(const vo_cfg *c, const vo_geom *g, double cx, double cy,
                    double cz, double r, double *nvert)
{
    DIMS;
    int rc = check_cfg(c);
    if (rc) return rc;
    size_t P = (size_t)mx * my * mz;
    for (size_t n = 0; n < P; n++) nvert[n] = 0.;
    /* inside test at cell centres: 4 = inside this body (ibm.c:224) */
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t id = IX(k, j, i);
                double dx = VX(g->cent, id) - cx, dy = VY(g->cent, id) - cy, dz = VZ(g->cent, id) - cz;
                double d = sqrt(dx * dx + dy * dy + dz * dz) - r;
                if (d < 0.) nvert[id] = 4.;
            }
    /* near-boundary dilation over the clamped 3x3x3 box (ibm.c:319-341) */
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                size_t id = IX(k, j, i);
                if (nvert[id] < 0) nvert[id] = 0;
                int ip = (i < mx - 1 ? (i + 1) : (i)), im = (i > 0 ? (i - 1) : (i));
                int jp = (j < my - 1 ? (j + 1) : (j)), jm = (j > 0 ? (j - 1) : (j));
                int kp = (k < mz - 1 ? (k + 1) : (k)), km = (k > 0 ? (k - 1) : (k));
                if ((int)(nvert[id] + 0.5) != 4) {
                    for (int kk = km; kk < kp + 1; kk++)
                        for (int jj = jm; jj < jp + 1; jj++)
                            for (int ii = im; ii < ip + 1; ii++)
                                if ((int)(nvert[IX(kk, jj, ii)] + 0.5) == 4)
                                    nvert[id] = nvert[id] > 2 ? nvert[id] : 2;
                }
            }
    /* back to 3 (solid) / 1 (IB node) on interior cells (ibm.c:508-516) */
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t id = IX(k, j, i);
                if ((int)(nvert[id] + 0.5) == 2) nvert[id] = 1;
                if ((int)(nvert[id] + 0.5) == 4) nvert[id] = 3;
            }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* rhs.c:57-741 Contra2Cart_2 (bctype 1 / periodic subset)                  */
/* ------------------------------------------------------------------------ */
/* ------------------------------------------------------------------------ */
/* wallfunction.c: the Cabot wall model behind bctype -1                      */
/* ------------------------------------------------------------------------ */
#define WF_KAPPA 0.41          /* wallfunction.c:15 */
#define WF_INTERVAL 2          /* interval_yp, wallfunction.c:288 */
#define WF_MAX_YP 1000000      /* max_yp */
#define WF_NYP (WF_MAX_YP / WF_INTERVAL)

/* wallfunction.c:271-274: nu_t / nu; pow(x, 2.0) is x * x at -O3 */
static double wf_nut(double yplus) { return WF_KAPPA * yplus * sq(1. - exp(-yplus / 19.)); }

/* wallfunction.c:291-323 pre_integrate: cumulative integral of
 * 1 / (1 + nu_t/nu) over y+ in steps of 2, each by 24 Simpson 3/8 panels */
static double *wf_table = NULL;
const double *vo_wf_table(void)
{
    if (wf_table) return wf_table;
    double *b = (double *)malloc(sizeof(double) * (WF_NYP + 1));
    if (!b) return NULL;
    b[0] = 0.;
    double E[25];
    for (int i = 1; i <= WF_NYP; i++) {
        const int N = 24;
        double ya_p = (double)(i - 1) * WF_INTERVAL;
        double yb_p = (double)(i) * WF_INTERVAL;
        double ydiff = yb_p - ya_p, dy_p = ydiff / (double)N;
        double val = 0, ybegin_p = ya_p;
        for (int k = 0; k <= N; k++) E[k] = 1. / (1. + wf_nut(ybegin_p + dy_p * k));
        for (int k = 0; k < N; k++) {
            double F[4];
            F[0] = E[k];
            F[1] = 1. / (1. + wf_nut(ybegin_p + dy_p * 1. / 3.));
            F[2] = 1. / (1. + wf_nut(ybegin_p + dy_p * 2. / 3.));
            F[3] = E[k + 1];
            val += dy_p / 3. * (3 * F[0] + 9 * F[1] + 9 * F[2] + 3 * F[3]) / 8.;
            ybegin_p += dy_p;
        }
        b[i] = b[i - 1] + val;
    }
    wf_table = b;
    return b;
}

/* wallfunction.c:325-371 integrate_F: the table interpolated linearly up to
 * y+ = 1e6, four Simpson 3/8 panels beyond.  The table index is clamped to
 * the table (the reference reads outside it for y+ < 0). */
double vo_wf_integrate_F(double nu, double utau, double yb)
{
    const double *buf = vo_wf_table();
    double val = 0;
    double ya_plus = 0 * utau / nu;
    double yb_plus = yb * utau / nu;
    int ib = (int)(yb_plus / (double)WF_INTERVAL);
    const int N = 4;
    if (yb_plus <= (double)WF_MAX_YP) {
        if (ib < 0) ib = 0;
        if (ib > WF_NYP - 1) ib = WF_NYP - 1;
        double int_a = 0;
        double int_b = (buf[ib + 1] - buf[ib]) / (double)(WF_INTERVAL) * (yb_plus - (double)(ib)*WF_INTERVAL) + buf[ib];
        return (int_b - int_a) / utau;
    }
    val = buf[WF_NYP];
    ya_plus = WF_MAX_YP;
    double ydiff = yb_plus - ya_plus, dy = ydiff / (double)N;
    double E[5];
    double ybegin = ya_plus;
    for (int i = 0; i <= N; i++) E[i] = 1. / (1. + wf_nut(ybegin + dy * i));
    for (int i = 0; i < N; i++) {
        double F[4];
        F[0] = E[i];
        F[1] = 1. / (1. + wf_nut(ybegin + dy * 1. / 3.));
        F[2] = 1. / (1. + wf_nut(ybegin + dy * 2. / 3.));
        F[3] = E[i + 1];
        val += dy / 3. * (3 * F[0] + 9 * F[1] + 9 * F[2] + 3 * F[3]) / 8.;
        ybegin += dy;
    }
    val /= utau;
    return val;
}

/* wallfunction.c:211-214, 234-237, 252-256, 389-404 (dpdn = 0) */
static double wf_u_cabot(double nu, double y, double utau) { return utau * utau * vo_wf_integrate_F(nu, utau, y); }
static double wf_f_cabot(double nu, double u, double y, double utau)
{
    return utau * utau * vo_wf_integrate_F(nu, utau, y) - u;
}
static double wf_df_cabot(double nu, double u, double y, double utau)
{
    double eps = 1.e-7;
    return (wf_f_cabot(nu, u, y, utau + eps) - wf_f_cabot(nu, u, y, utau - eps)) / (2 * eps);
}
double vo_wf_find_utau(double nu, double u, double y, double guess)
{
    double x = guess, x0 = guess;
    for (int i = 0; i < 30; i++) {
        x = x0 - wf_f_cabot(nu, u, y, x0) / wf_df_cabot(nu, u, y, x0);
        if (fabs(x0 - x) < 1.e-10) break;
        x0 = x;
    }
    return x;
}

/* wallfunction.c:18-43 wall_function with Ua = 0: the tangential velocity of
 * the first cell from the wall model, the normal part scaled by sb / sc */
void vo_wall_function(double nu, double sc, double sb, const double Uc[3], double Ub[3], double *ustar,
                      double nx, double ny, double nz)
{
    const double Ua[3] = {0, 0, 0};
    double u_c = Uc[0] - Ua[0], v_c = Uc[1] - Ua[1], w_c = Uc[2] - Ua[2];
    double un = u_c * nx + v_c * ny + w_c * nz;
    double ut = u_c - un * nx, vt = v_c - un * ny, wt = w_c - un * nz;
    double ut_mag = sqrt(ut * ut + vt * vt + wt * wt);
    *ustar = vo_wf_find_utau(nu, ut_mag, sc, 0.01);
    double ut_mag_modeled = wf_u_cabot(nu, sb, *ustar);
    if (ut_mag > 1.e-10) {
        ut *= ut_mag_modeled / ut_mag;
        vt *= ut_mag_modeled / ut_mag;
        wt *= ut_mag_modeled / ut_mag;
    } else ut = vt = wt = 0;
    Ub[0] = ut + sb / sc * un * nx;
    Ub[1] = vt + sb / sc * un * ny;
    Ub[2] = wt + sb / sc * un * nz;
    Ub[0] += Ua[0];
    Ub[1] += Ua[1];
    Ub[2] += Ua[2];
}

/* rhs2.c:539-583 Calculate_Covariant_metrics + Calculate_normal: the unit
 * covariant base vectors (columns of the inverse metric matrix) */
static void calc_normal(const double *csi, const double *eta, const double *zet, double ni[3], double nj[3],
                        double nk[3])
{
    const double a11 = csi[0], a12 = csi[1], a13 = csi[2];
    const double a21 = eta[0], a22 = eta[1], a23 = eta[2];
    const double a31 = zet[0], a32 = zet[1], a33 = zet[2];
    double det = a11 * (a33 * a22 - a32 * a23) - a21 * (a33 * a12 - a32 * a13) + a31 * (a23 * a12 - a22 * a13);
    double G00 = (a33 * a22 - a32 * a23) / det, G01 = -(a33 * a12 - a32 * a13) / det, G02 = (a23 * a12 - a22 * a13) / det;
    double G10 = -(a33 * a21 - a31 * a23) / det, G11 = (a33 * a11 - a31 * a13) / det, G12 = -(a23 * a11 - a21 * a13) / det;
    double G20 = (a32 * a21 - a31 * a22) / det, G21 = -(a32 * a11 - a31 * a12) / det, G22 = (a22 * a11 - a21 * a12) / det;
    double nx_i = G00, ny_i = G10, nz_i = G20;
    double nx_j = G01, ny_j = G11, nz_j = G21;
    double nx_k = G02, ny_k = G12, nz_k = G22;
    double sum_i = sqrt(nx_i * nx_i + ny_i * ny_i + nz_i * nz_i);
    double sum_j = sqrt(nx_j * nx_j + ny_j * ny_j + nz_j * nz_j);
    double sum_k = sqrt(nx_k * nx_k + ny_k * ny_k + nz_k * nz_k);
    nx_i /= sum_i, ny_i /= sum_i, nz_i /= sum_i;
    nx_j /= sum_j, ny_j /= sum_j, nz_j /= sum_j;
    nx_k /= sum_k, ny_k /= sum_k, nz_k /= sum_k;
    ni[0] = nx_i, ni[1] = ny_i, ni[2] = nz_i;
    nj[0] = nx_j, nj[1] = ny_j, nj[2] = nz_j;
    nk[0] = nx_k, nk[1] = ny_k, nk[2] = nz_k;
}

/* momentum.c:2014-2039: with no immersed body, the first cells of the
 * wall-function walls become IB nodes (nvert = 1) */
int vo_wall_mask(const vo_cfg *c, double *nvert)
{
    DIMS;
    int rc = check_cfg(c);
    if (rc) return rc;
    if (c->immersed) return 0;
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t id = IX(k, j, i);
                if (c->bctype[0] == -1 && i == 1) nvert[id] = 1;
                if (c->bctype[1] == -1 && i == mx - 2) nvert[id] = 1;
                if (c->bctype[2] == -1 && j == 1) nvert[id] = 1;
                if (c->bctype[3] == -1 && j == my - 2) nvert[id] = 1;
                if (c->bctype[4] == -1 && k == 1) nvert[id] = 1;
                if (c->bctype[5] == -1 && k == mz - 2) nvert[id] = 1;
            }
    return 0;
}

int vo_contra2cart(const vo_cfg *c, const vo_geom *g, const double *nvert,
                   double *ucont, double *ucat)
{
    return vo_contra2cart_ustar(c, g, nvert, ucont, ucat, NULL);
}

int vo_contra2cart_ustar(const vo_cfg *c, const vo_geom *g, const double *nvert,
                         double *ucont, double *ucat, double *ustar)
{
    DIMS;
    int rc = check_cfg(c);
    if (rc) return rc;
    per_copy_vec(c, ucont); /* rhs.c:121-148 */
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t id = IX(k, j, i);
                if (!(nvert[id] < 0.1)) continue;
                double m00 = VX(g->csi, id), m01 = VY(g->csi, id), m02 = VZ(g->csi, id);
                double m10 = VX(g->eta, id), m11 = VY(g->eta, id), m12 = VZ(g->eta, id);
                double m20 = VX(g->zet, id), m21 = VY(g->zet, id), m22 = VZ(g->zet, id);
                double q0 = 0.5 * (VX(ucont, IX(k, j, i - 1)) + VX(ucont, id));
                double q1 = 0.5 * (VY(ucont, IX(k, j - 1, i)) + VY(ucont, id));
                double q2 = 0.5 * (VZ(ucont, IX(k - 1, j, i)) + VZ(ucont, id));
                double det = m00 * (m11 * m22 - m12 * m21) - m01 * (m10 * m22 - m12 * m20) +
                             m02 * (m10 * m21 - m11 * m20);
                double det0 = q0 * (m11 * m22 - m12 * m21) - q1 * (m01 * m22 - m02 * m21) +
                              q2 * (m01 * m12 - m02 * m11);
                double det1 = -q0 * (m10 * m22 - m12 * m20) + q1 * (m00 * m22 - m02 * m20) -
                              q2 * (m00 * m12 - m02 * m10);
                double det2 = q0 * (m10 * m21 - m11 * m20) - q1 * (m00 * m21 - m01 * m20) +
                              q2 * (m00 * m11 - m01 * m10);
                VX(ucat, id) = det0 / det;
                VY(ucat, id) = det1 / det;
                VZ(ucat, id) = det2 / det;
            }
    per_copy_vec(c, ucat); /* rhs.c:246-279 */
    /* boundary loop reads the snapshot lucat and writes ucat (rhs.c:294-674) */
    size_t P = (size_t)mx * my * mz;
    double *L = (double *)malloc(3 * P * sizeof(double));
    if (!L) return -1;
    memcpy(L, ucat, 3 * P * sizeof(double));
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                size_t id = IX(k, j, i);
                if ((int)(nvert[id] + 0.1) == 3) { VX(ucat, id) = VY(ucat, id) = VZ(ucat, id) = 0; continue; }
                /* wall function for boundary (rhs.c:303-333, 369-401) */
                for (int dir = 0; dir < 2; dir++) {
                    const int n = dir == 0 ? i : j, m = dir == 0 ? mx : my;
                    const int t1 = dir == 0 ? j : i, m1 = dir == 0 ? my : mx;
                    if (!(t1 != 0 && t1 != m1 - 1 && k != 0 && k != mz - 1 &&
                          ((c->bctype[2 * dir] == -1 && n == 1) || (c->bctype[2 * dir + 1] == -1 && n == m - 2))))
                        continue;
                    const double *a = (dir == 0 ? g->csi : g->eta) + 3 * id;
                    double area = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
                    double sb, sc, ni[3], nj[3], nk[3], Ub[3], us;
                    sb = 0.5 / g->aj[id] / area;
                    const size_t nb = dir == 0 ? (n == 1 ? IX(k, j, i + 1) : IX(k, j, i - 1))
                                               : (n == 1 ? IX(k, j + 1, i) : IX(k, j - 1, i));
                    sc = 2 * sb + 0.5 / g->aj[nb] / area;
                    const double Uc[3] = {VX(L, nb), VY(L, nb), VZ(L, nb)};
                    calc_normal(g->csi + 3 * id, g->eta + 3 * id, g->zet + 3 * id, ni, nj, nk);
                    double *nv = dir == 0 ? ni : nj;
                    if (n == m - 2) nv[0] *= -1, nv[1] *= -1, nv[2] *= -1;
                    vo_wall_function(1. / c->ren, sc, sb, Uc, Ub, &us, nv[0], nv[1], nv[2]);
                    if (ustar) ustar[id] = us;
                    VX(ucat, id) = Ub[0]; VY(ucat, id) = Ub[1]; VZ(ucat, id) = Ub[2];
                }
                size_t s = (size_t)-1;
                if (i == 0 && c->bctype[0] == 1 && (j != 0 && k != 0)) s = IX(k, j, i + 1);
                if (i == mx - 1 && c->bctype[1] == 1 && (j != 0 && k != 0)) s = IX(k, j, i - 1);
                if (j == 0 && c->bctype[2] == 1 && (i != 0 && k != 0)) s = IX(k, j + 1, i);
                if (j == my - 1 && c->bctype[3] == 1 && (i != 0 && k != 0)) s = IX(k, j - 1, i);
                if (k == 0 && c->bctype[4] == 1 && (i != 0 && j != 0)) s = IX(k + 1, j, i);
                if (k == mz - 1 && c->bctype[5] == 1 && (i != 0 && j != 0)) s = IX(k - 1, j, i);
                if (s != (size_t)-1) {
                    VX(ucat, id) = -1 * VX(L, s); VY(ucat, id) = -1 * VY(L, s); VZ(ucat, id) = -1 * VZ(L, s);
                }
                /* outlet (rhs.c:646-663): zero behind a solid last cell; the
                 * inlet ghost is InflowFlux's (rhs.c:636-644 is disabled) */
                if (c->bctype[5] == 4 && k == mz - 1 && (i != 0 && i != mx - 1 && j != 0 && j != my - 1) &&
                    nvert[IX(k - 1, j, i)] > 0.1)
                    VX(ucat, id) = VY(ucat, id) = VZ(ucat, id) = 0;
                /* 101220 corner zeroing (rhs.c:668-673) */
                if (i != 0 && i != mx - 1 && j != 0 && j != my - 2 && k != 0 && k != mz - 2) {
                    int z = 0;
                    if (c->bctype[0] <= 1 && c->bctype[2] <= 1 && (i == 1 && j == 1)) z = 1;
                    if (c->bctype[1] <= 1 && c->bctype[2] <= 1 && (i == mx - 2 && j == 1)) z = 1;
                    if (c->bctype[0] <= 1 && c->bctype[3] <= 1 && (i == 1 && j == my - 2)) z = 1;
                    if (c->bctype[1] <= 1 && c->bctype[3] <= 1 && (i == mx - 2 && j == my - 2)) z = 1;
                    if (z) VX(ucat, id) = VY(ucat, id) = VZ(ucat, id) = 0;
                }
            }
    free(L);
    per_copy_vec(c, ucat); /* rhs.c:704-740 */
    return 0;
}

/* ------------------------------------------------------------------------ */
/* k-omega.c:48-278 face gradients, 305-423 cell gradient, 597-610 chain     */
/* ------------------------------------------------------------------------ */
#define SOLID 0.1 /* k-omega.c:11 */
#define U3(k, j, i, q) ucat[3 * IX(k, j, i) + (q)]
#define NV(k, j, i) nvert[IX(k, j, i)]

static void du_i(const vo_cfg *c, int i, int j, int k, const double *ucat, const double *nvert,
                 double dc[3], double de[3], double dz[3])
{
    DIMS;
    for (int q = 0; q < 3; q++) dc[q] = U3(k, j, i + 1, q) - U3(k, j, i, q);
    if (NV(k, j + 1, i) > SOLID || NV(k, j + 1, i + 1) > SOLID)
        for (int q = 0; q < 3; q++) de[q] = (U3(k, j, i + 1, q) + U3(k, j, i, q) - U3(k, j - 1, i + 1, q) - U3(k, j - 1, i, q)) * 0.5;
    else if (NV(k, j - 1, i) > SOLID || NV(k, j - 1, i + 1) > SOLID)
        for (int q = 0; q < 3; q++) de[q] = (U3(k, j + 1, i + 1, q) + U3(k, j + 1, i, q) - U3(k, j, i + 1, q) - U3(k, j, i, q)) * 0.5;
    else
        for (int q = 0; q < 3; q++) de[q] = (U3(k, j + 1, i + 1, q) + U3(k, j + 1, i, q) - U3(k, j - 1, i + 1, q) - U3(k, j - 1, i, q)) * 0.25;
    if (NV(k + 1, j, i) > SOLID || NV(k + 1, j, i + 1) > SOLID)
        for (int q = 0; q < 3; q++) dz[q] = (U3(k, j, i + 1, q) + U3(k, j, i, q) - U3(k - 1, j, i + 1, q) - U3(k - 1, j, i, q)) * 0.5;
    else if (NV(k - 1, j, i) > SOLID || NV(k - 1, j, i + 1) > SOLID)
        for (int q = 0; q < 3; q++) dz[q] = (U3(k + 1, j, i + 1, q) + U3(k + 1, j, i, q) - U3(k, j, i + 1, q) - U3(k, j, i, q)) * 0.5;
    else
        for (int q = 0; q < 3; q++) dz[q] = (U3(k + 1, j, i + 1, q) + U3(k + 1, j, i, q) - U3(k - 1, j, i + 1, q) - U3(k - 1, j, i, q)) * 0.25;
}

static void du_j(const vo_cfg *c, int i, int j, int k, const double *ucat, const double *nvert,
                 double dc[3], double de[3], double dz[3])
{
    DIMS;
    if (NV(k, j, i + 1) > SOLID || NV(k, j + 1, i + 1) > SOLID)
        for (int q = 0; q < 3; q++) dc[q] = (U3(k, j + 1, i, q) + U3(k, j, i, q) - U3(k, j + 1, i - 1, q) - U3(k, j, i - 1, q)) * 0.5;
    else if (NV(k, j, i - 1) > SOLID || NV(k, j + 1, i - 1) > SOLID)
        for (int q = 0; q < 3; q++) dc[q] = (U3(k, j + 1, i + 1, q) + U3(k, j, i + 1, q) - U3(k, j + 1, i, q) - U3(k, j, i, q)) * 0.5;
    else
        for (int q = 0; q < 3; q++) dc[q] = (U3(k, j + 1, i + 1, q) + U3(k, j, i + 1, q) - U3(k, j + 1, i - 1, q) - U3(k, j, i - 1, q)) * 0.25;
    for (int q = 0; q < 3; q++) de[q] = U3(k, j + 1, i, q) - U3(k, j, i, q);
    if (NV(k + 1, j, i) > SOLID || NV(k + 1, j + 1, i) > SOLID)
        for (int q = 0; q < 3; q++) dz[q] = (U3(k, j + 1, i, q) + U3(k, j, i, q) - U3(k - 1, j + 1, i, q) - U3(k - 1, j, i, q)) * 0.5;
    else if (NV(k - 1, j, i) > SOLID || NV(k - 1, j + 1, i) > SOLID)
        for (int q = 0; q < 3; q++) dz[q] = (U3(k + 1, j + 1, i, q) + U3(k + 1, j, i, q) - U3(k, j + 1, i, q) - U3(k, j, i, q)) * 0.5;
    else
        for (int q = 0; q < 3; q++) dz[q] = (U3(k + 1, j + 1, i, q) + U3(k + 1, j, i, q) - U3(k - 1, j + 1, i, q) - U3(k - 1, j, i, q)) * 0.25;
}

static void du_k(const vo_cfg *c, int i, int j, int k, const double *ucat, const double *nvert,
                 double dc[3], double de[3], double dz[3])
{
    DIMS;
    if (NV(k, j, i + 1) > SOLID || NV(k + 1, j, i + 1) > SOLID)
        for (int q = 0; q < 3; q++) dc[q] = (U3(k + 1, j, i, q) + U3(k, j, i, q) - U3(k + 1, j, i - 1, q) - U3(k, j, i - 1, q)) * 0.5;
    else if (NV(k, j, i - 1) > SOLID || NV(k + 1, j, i - 1) > SOLID)
        for (int q = 0; q < 3; q++) dc[q] = (U3(k + 1, j, i + 1, q) + U3(k, j, i + 1, q) - U3(k + 1, j, i, q) - U3(k, j, i, q)) * 0.5;
    else
        for (int q = 0; q < 3; q++) dc[q] = (U3(k + 1, j, i + 1, q) + U3(k, j, i + 1, q) - U3(k + 1, j, i - 1, q) - U3(k, j, i - 1, q)) * 0.25;
    if (NV(k, j + 1, i) > SOLID || NV(k + 1, j + 1, i) > SOLID)
        for (int q = 0; q < 3; q++) de[q] = (U3(k + 1, j, i, q) + U3(k, j, i, q) - U3(k + 1, j - 1, i, q) - U3(k, j - 1, i, q)) * 0.5;
    else if (NV(k, j - 1, i) > SOLID || NV(k + 1, j - 1, i) > SOLID)
        for (int q = 0; q < 3; q++) de[q] = (U3(k + 1, j + 1, i, q) + U3(k, j + 1, i, q) - U3(k + 1, j, i, q) - U3(k, j, i, q)) * 0.5;
    else
        for (int q = 0; q < 3; q++) de[q] = (U3(k + 1, j + 1, i, q) + U3(k, j + 1, i, q) - U3(k + 1, j - 1, i, q) - U3(k, j - 1, i, q)) * 0.25;
    for (int q = 0; q < 3; q++) dz[q] = U3(k + 1, j, i, q) - U3(k, j, i, q);
}

/* k-omega.c:305-423 (note the k==0 quirk at 391) */
static void du_center(const vo_cfg *c, int i, int j, int k, const double *ucat, const double *nvert,
                      double dc[3], double de[3], double dz[3])
{
    DIMS;
    if (NV(k, j, i + 1) > SOLID || (!c->i_periodic && i == mx - 2))
        for (int q = 0; q < 3; q++) dc[q] = (U3(k, j, i, q) - U3(k, j, i - 1, q));
    else if (NV(k, j, i - 1) > SOLID || (!c->i_periodic && i == 1))
        for (int q = 0; q < 3; q++) dc[q] = (U3(k, j, i + 1, q) - U3(k, j, i, q));
    else
        for (int q = 0; q < 3; q++) dc[q] = (U3(k, j, i + 1, q) - U3(k, j, i - 1, q)) * 0.5;
    if (NV(k, j + 1, i) > SOLID || (!c->j_periodic && j == my - 2))
        for (int q = 0; q < 3; q++) de[q] = (U3(k, j, i, q) - U3(k, j - 1, i, q));
    else if (NV(k, j - 1, i) > SOLID || (!c->j_periodic && j == 1))
        for (int q = 0; q < 3; q++) de[q] = (U3(k, j + 1, i, q) - U3(k, j, i, q));
    else
        for (int q = 0; q < 3; q++) de[q] = (U3(k, j + 1, i, q) - U3(k, j - 1, i, q)) * 0.5;
    if (NV(k + 1, j, i) > SOLID || (!c->k_periodic && k == mz - 2))
        for (int q = 0; q < 3; q++) dz[q] = (U3(k, j, i, q) - U3(k - 1, j, i, q));
    else if (NV(k - 1, j, i) > SOLID || (!c->k_periodic && k == 0))
        for (int q = 0; q < 3; q++) dz[q] = (U3(k + 1, j, i, q) - U3(k, j, i, q));
    else
        for (int q = 0; q < 3; q++) dz[q] = (U3(k + 1, j, i, q) - U3(k - 1, j, i, q)) * 0.5;
}

/* ------------------------------------------------------------------------ */
/* les.c:1135-1252 Compute_eddy_viscosity_LES                               */
/* ------------------------------------------------------------------------ */
/* Compute_du_dxyz (k-omega.c:597-610) at cell id from du_center's
 * computational-space differences: D[a][b] = d u_a / d x_b */
static void du_dxyz(const vo_geom *g, size_t id, const double dc[3], const double de[3], const double dz[3],
                    double D[3][3])
{
    const double ajc = g->aj[id];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            D[a][b] = (dc[a] * g->csi[3 * id + b] + de[a] * g->eta[3 * id + b] + dz[a] * g->zet[3 * id + b]) * ajc;
}

/* Compute_du_dxyz (k-omega.c:597-610) with given metrics: the Cartesian
 * gradient at a face from its computational-space differences */
static void du_dxyz_face(double csi0, double csi1, double csi2, double eta0, double eta1, double eta2, double zet0,
                         double zet1, double zet2, double ajc, const double dc[3], const double de[3],
                         const double dz[3], double *du_dx, double *dv_dx, double *dw_dx, double *du_dy,
                         double *dv_dy, double *dw_dy, double *du_dz, double *dv_dz, double *dw_dz)
{
    *du_dx = (dc[0] * csi0 + de[0] * eta0 + dz[0] * zet0) * ajc;
    *du_dy = (dc[0] * csi1 + de[0] * eta1 + dz[0] * zet1) * ajc;
    *du_dz = (dc[0] * csi2 + de[0] * eta2 + dz[0] * zet2) * ajc;
    *dv_dx = (dc[1] * csi0 + de[1] * eta0 + dz[1] * zet0) * ajc;
    *dv_dy = (dc[1] * csi1 + de[1] * eta1 + dz[1] * zet1) * ajc;
    *dv_dz = (dc[1] * csi2 + de[1] * eta2 + dz[1] * zet2) * ajc;
    *dw_dx = (dc[2] * csi0 + de[2] * eta0 + dz[2] * zet0) * ajc;
    *dw_dy = (dc[2] * csi1 + de[2] * eta1 + dz[2] * zet1) * ajc;
    *dw_dz = (dc[2] * csi2 + de[2] * eta2 + dz[2] * zet2) * ajc;
}

/* |S| = sqrt(2 S:S) in the reference's term order (les.c:199-203) */
static double strain_abs(const double D[3][3], double S[3][3])
{
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) S[a][b] = 0.5 * (D[a][b] + D[b][a]);
    double s = S[0][0] * S[0][0];
    for (int n = 1; n < 9; n++) s += S[n / 3][n % 3] * S[n / 3][n % 3];
    return sqrt(2.0 * s);
}

/* nu_t = Cs |S| (1/aj)^(2/3) with Cs either the constant cs or csf[id] */
static int eddy_viscosity(const vo_cfg *c, const vo_geom *g, const double *nvert,
                          const double *ucat, double cs, const double *csf, double *nu_t)
{
    DIMS;
    int rc = check_cfg(c);
    if (rc) return rc;
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t id = IX(k, j, i);
                if (nvert[id] > 1.1) { nu_t[id] = 0; continue; }
                double ajc = g->aj[id];
                double csi0 = VX(g->csi, id), csi1 = VY(g->csi, id), csi2 = VZ(g->csi, id);
                double eta0 = VX(g->eta, id), eta1 = VY(g->eta, id), eta2 = VZ(g->eta, id);
                double zet0 = VX(g->zet, id), zet1 = VY(g->zet, id), zet2 = VZ(g->zet, id);
                double dc[3], de[3], dz[3];
                du_center(c, i, j, k, ucat, nvert, dc, de, dz);
                double du_dx = (dc[0] * csi0 + de[0] * eta0 + dz[0] * zet0) * ajc;
                double du_dy = (dc[0] * csi1 + de[0] * eta1 + dz[0] * zet1) * ajc;
                double du_dz = (dc[0] * csi2 + de[0] * eta2 + dz[0] * zet2) * ajc;
                double dv_dx = (dc[1] * csi0 + de[1] * eta0 + dz[1] * zet0) * ajc;
                double dv_dy = (dc[1] * csi1 + de[1] * eta1 + dz[1] * zet1) * ajc;
                double dv_dz = (dc[1] * csi2 + de[1] * eta2 + dz[1] * zet2) * ajc;
                double dw_dx = (dc[2] * csi0 + de[2] * eta0 + dz[2] * zet0) * ajc;
                double dw_dy = (dc[2] * csi1 + de[2] * eta1 + dz[2] * zet1) * ajc;
                double dw_dz = (dc[2] * csi2 + de[2] * eta2 + dz[2] * zet2) * ajc;
                double Sxx = 0.5 * (du_dx + du_dx), Sxy = 0.5 * (du_dy + dv_dx), Sxz = 0.5 * (du_dz + dw_dx);
                double Syx = Sxy, Syy = 0.5 * (dv_dy + dv_dy), Syz = 0.5 * (dv_dz + dw_dy);
                double Szx = Sxz, Szy = Syz, Szz = 0.5 * (dw_dz + dw_dz);
                double Sabs = sqrt(2.0 * (Sxx * Sxx + Sxy * Sxy + Sxz * Sxz + Syx * Syx + Syy * Syy + Syz * Syz +
                                          Szx * Szx + Szy * Szy + Szz * Szz));
                double filter = pow(1. / g->aj[id], 1. / 3.);
                nu_t[id] = (csf ? csf[id] : cs) * pow(filter, 2.0) * Sabs;
            }
    per_copy_sca(c, nu_t); /* les.c:1215-1243 */
    return 0;
}

int vo_eddy_viscosity(const vo_cfg *c, const vo_geom *g, const double *nvert,
                      const double *ucat, double cs, double *nu_t)
{
    return eddy_viscosity(c, g, nvert, ucat, cs, NULL, nu_t);
}

int vo_eddy_viscosity_cs(const vo_cfg *c, const vo_geom *g, const double *nvert,
                         const double *ucat, const double *cs, double *nu_t)
{
    return eddy_viscosity(c, g, nvert, ucat, 0, cs, nu_t);
}

/* ------------------------------------------------------------------------ */
/* les.c:67-1133 Compute_Smagorinsky_Constant_1, les == 2 with the default    */
/* options: 3-D box test filter (testfilter_ik 0), no homogeneous averaging,  */
/* no Clark term                                                              */
/* ------------------------------------------------------------------------ */
#define LES_EPS 1.e-4  /* les.c:19 */
#define WALL_CS 0.001  /* les.c:20 */
#define PMAX(a, b) ((a) < (b) ? (b) : (a))   /* PetscMax */
#define PMIN(a, b) ((a) < (b) ? (a) : (b))   /* PetscMin */

/* rhs2.c:412-475 integrate_testfilter: the weighted mean over the eight
 * 2x2x2 octants of the 3x3x3 box, in the reference's summation order.
 * val[R][Q][P] with R, Q, P the k, j, i offsets + 1. */
double vo_testfilter(const double val[3][3][3], const double w[3][3][3])
{
    double vs = 0, ws = 0;
    for (int o = 0; o < 8; o++) {
        const int oa = o & 1, oc = (o >> 1) & 1, ob = (o >> 2) & 1;   /* v1..v8 */
        double v = 0, ww = 0;
        for (int t = 0; t < 8; t++) {
            const int a = oa + (t & 1), cc = oc + ((t >> 1) & 1), b = ob + ((t >> 2) & 1);
            v = t ? v + val[a][b][cc] * w[a][b][cc] : val[a][b][cc] * w[a][b][cc];
            ww = t ? ww + w[a][b][cc] : w[a][b][cc];
        }
        vs = o ? vs + v : v;
        ws = o ? ws + ww : ww;
    }
    return vs / ws;
}

/* rhs2.c:388-410 integrate_testfilter_ik: Simpson's rule over the i-k plane
 * of the box (weights 1, 4, 16 over 36); the volumes are not used */
double vo_testfilter_ik(const double val[3][3][3])
{
    return ((val[0][1][0] + val[2][1][0] + val[0][1][2] + val[2][1][2]) +
            4. * (val[0][1][1] + val[1][1][0] + val[2][1][1] + val[1][1][2]) + 16. * val[1][1][1]) / 36.;
}

/* integrate_testfilter with the -testfilter_ik switch (rhs2.c:417-419) */
static double tfilter(const double val[3][3][3], const double w[3][3][3], int ik)
{
    return ik ? vo_testfilter_ik(val) : vo_testfilter(val, w);
}

static const vo_les_opt les_default = {2, 0, 0, 0, 0, 0.5, 2, 0, 0};
#define OPT(o) ((o) ? (o) : &les_default)

/* les.c:23-64 get_weight: 1/aj, 0 in solid cells (the boundary rules are
 * commented out there) */
static void get_weight(const vo_cfg *c, const vo_geom *g, const double *nvert, int i, int j, int k,
                       double w[3][3][3])
{
    DIMS;
    for (int p = -1; p <= 1; p++)
        for (int q = -1; q <= 1; q++)
            for (int r = -1; r <= 1; r++) {
                const size_t id = IX(k + r, j + q, i + p);
                w[r + 1][q + 1][p + 1] = nvert[id] > 1.1 ? 0 : 1. / g->aj[id];
            }
}

/* les.c:717-758 / 1066-1107: weights of the Cs-averaging filters (1/aj, 0 in
 * solids and on non-periodic ghost planes; the j rule tests the centre's j,
 * so the j = my-1 ghost plane keeps its weight) and the periodic sources */
static void cs_weight(const vo_cfg *c, const vo_geom *g, const double *nvert, const double *f, int i, int j, int k,
                      double w[3][3][3], double v[3][3][3])
{
    DIMS;
    for (int a = -1; a <= 1; a++)
        for (int b = -1; b <= 1; b++)
            for (int cc = -1; cc <= 1; cc++) {
                const int R = cc + 1, Q = b + 1, P = a + 1;
                int K = k + cc, J = j + b, I = i + a;
                double wt = 1. / g->aj[IX(K, J, I)];
                if (nvert[IX(K, J, I)] > 1.1) wt = 0;
                if (c->i_periodic) { if (I == 0) I = mx - 2; else if (I == mx - 1) I = 1; }
                else if (I == 0 || I == mx - 1) wt = 0;
                if (c->j_periodic) { if (J == 0) J = my - 2; else if (J == my - 1) J = 1; }
                else if (J == 0 || j == my - 1) wt = 0;
                if (c->k_periodic) { if (K == 0) K = mz - 2; else if (K == mz - 1) K = 1; }
                else if (K == 0 || K == mz - 1) wt = 0;
                w[R][Q][P] = wt;
                v[R][Q][P] = f[IX(K, J, I)];
            }
}

/* les.c:188-257, first pass: velocity gradient grad[9] (du/dx, du/dy, du/dz,
 * dv/dx, ...: the reference's Ax, Ay, Az), |S| and the test-filtered velocity
 * at the non-solid interior cells, 0 elsewhere, then periodic copies. */
int vo_dsm_pass1(const vo_cfg *c, const vo_geom *g, const double *nvert, const double *ucat, const vo_les_opt *o,
                 double *grad, double *sabs, double *ucat_f)
{
    o = OPT(o);
    DIMS;
    int rc = check_cfg(c);
    if (rc) return rc;
    const size_t P = (size_t)mx * my * mz;
    memset(grad, 0, 9 * P * sizeof(double));
    memset(sabs, 0, P * sizeof(double));
    memset(ucat_f, 0, 3 * P * sizeof(double));
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                const size_t id = IX(k, j, i);
                if (nvert[id] > 1.1) continue;
                double dc[3], de[3], dz[3], D[3][3], S[3][3], w[3][3][3], u[3][3][3];
                du_center(c, i, j, k, ucat, nvert, dc, de, dz);
                du_dxyz(g, id, dc, de, dz, D);
                sabs[id] = strain_abs(D, S);
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) grad[9 * id + 3 * a + b] = D[a][b];
                get_weight(c, g, nvert, i, j, k, w);
                for (int q = 0; q < 3; q++) {
                    for (int r = 0; r < 27; r++)
                        u[r / 9][(r / 3) % 3][r % 3] = ucat[3 * IX(k + r / 9 - 1, j + (r / 3) % 3 - 1, i + r % 3 - 1) + q];
                    ucat_f[3 * id + q] = tfilter(u, w, o->testfilter_ik);
                }
            }
    for (int k = 0; k < mz; k++)   /* les.c:271-304 */
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                int a, b, cc;
                if (!per_src(c, i, j, k, &a, &b, &cc)) continue;
                const size_t d = IX(k, j, i), s = IX(cc, b, a);
                sabs[d] = sabs[s];
                for (int q = 0; q < 9; q++) grad[9 * d + q] = grad[9 * s + q];
                for (int q = 0; q < 3; q++) ucat_f[3 * d + q] = ucat_f[3 * s + q];
            }
    return 0;
}

/* rhs2.c:538-552 Calculate_Covariant_metrics: the inverse of [csi; eta; zet] */
static void covariant_metrics(const double gg[3][3], double G[3][3])
{
    const double a11 = gg[0][0], a12 = gg[0][1], a13 = gg[0][2];
    const double a21 = gg[1][0], a22 = gg[1][1], a23 = gg[1][2];
    const double a31 = gg[2][0], a32 = gg[2][1], a33 = gg[2][2];
    const double det = a11 * (a33 * a22 - a32 * a23) - a21 * (a33 * a12 - a32 * a13) + a31 * (a23 * a12 - a22 * a13);
    G[0][0] = (a33 * a22 - a32 * a23) / det, G[0][1] = -(a33 * a12 - a32 * a13) / det, G[0][2] = (a23 * a12 - a22 * a13) / det;
    G[1][0] = -(a33 * a21 - a31 * a23) / det, G[1][1] = (a33 * a11 - a31 * a13) / det, G[1][2] = -(a23 * a11 - a21 * a13) / det;
    G[2][0] = (a32 * a21 - a31 * a22) / det, G[2][1] = -(a32 * a11 - a31 * a12) / det, G[2][2] = (a22 * a11 - a21 * a12) / det;
}

static void dxdydz(double ajc, const double *csi, const double *eta, const double *zet, double *dx, double *dy,
                   double *dz);

/* les.c:306-687, second pass: the Germano identity contracted in the
 * contravariant form, LM = L_ba M_aq G_bq and MM = M_nm M_nl G_ml, at the
 * non-solid interior cells (0 elsewhere), then periodic copies.  With
 * -clark, LM also takes - N_ab M_aq G_bq, N = 4 R - <T> the Clark model's
 * Germano term (les.c:334-340, 410-420, 489-548, 648): T from the gradient
 * at each box point and R from the gradient of the filtered velocity, both
 * with the length scales of the centre cell. */
int vo_dsm_pass2(const vo_cfg *c, const vo_geom *g, const double *nvert, const double *ucat, const vo_les_opt *o,
                 const double *grad, const double *sabs, const double *ucat_f, double *LM, double *MM)
{
    o = OPT(o);
    DIMS;
    int rc = check_cfg(c);
    if (rc) return rc;
    const size_t P = (size_t)mx * my * mz;
    memset(LM, 0, P * sizeof(double));
    memset(MM, 0, P * sizeof(double));
    const double coef[3][3][3] = {{{0.125, 0.250, 0.125}, {0.250, 0.500, 0.250}, {0.125, 0.250, 0.125}},
                                  {{0.250, 0.500, 0.250}, {0.500, 1.000, 0.500}, {0.250, 0.500, 0.250}},
                                  {{0.125, 0.250, 0.125}, {0.250, 0.500, 0.250}, {0.125, 0.250, 0.125}}};
    const double *met[3] = {g->csi, g->eta, g->zet};
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                const size_t id = IX(k, j, i);
                if (nvert[id] > 1.1) continue;
                double w[3][3][3], UU[3][3][3][3][3], SS[3][3][3][3][3], TT[3][3][3][3][3];
                double lx, ly, lz, dl2[3];   /* les.c:339-340: the centre cell's length scales */
                dxdydz(g->aj[id], &g->csi[3 * id], &g->eta[3 * id], &g->zet[3 * id], &lx, &ly, &lz);
                dl2[0] = lx * lx, dl2[1] = ly * ly, dl2[2] = lz * lz;
                get_weight(c, g, nvert, i, j, k, w);
                for (int r = 0; r < 27; r++) {
                    const int R = r / 9, Q = (r / 3) % 3, Pp = r % 3;
                    const size_t n = IX(k + R - 1, j + Q - 1, i + Pp - 1);
                    const double u[3] = {ucat[3 * n], ucat[3 * n + 1], ucat[3 * n + 2]};
                    for (int b = 0; b < 3; b++) {
                        const double *m = met[b] + 3 * n;
                        const double U = u[0] * m[0] + u[1] * m[1] + u[2] * m[2];
                        for (int a = 0; a < 3; a++) UU[b][a][R][Q][Pp] = U * u[a];
                    }
                    double D[3][3], S[3][3];
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) D[a][b] = grad[9 * n + 3 * a + b];
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) S[a][b] = 0.5 * (D[a][b] + D[b][a]);
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) SS[a][b][R][Q][Pp] = S[a][b] * sabs[n];
                    if (c->clark)   /* les.c:412-420 */
                        for (int a = 0; a < 3; a++)
                            for (int b = 0; b < 3; b++)
                                TT[a][b][R][Q][Pp] = (D[a][0] * D[b][0] * dl2[0] + D[a][1] * D[b][1] * dl2[1] +
                                                      D[a][2] * D[b][2] * dl2[2]) / 12.;
                }
                double sum_weight = 0;
                for (int p = -1; p <= 1; p++)
                    for (int q = -1; q <= 1; q++)
                        for (int r = -1; r <= 1; r++) sum_weight += w[r + 1][q + 1][p + 1] * coef[r + 1][q + 1][p + 1];
                const double filter = pow(1. / g->aj[id], 1. / 3.);
                const double test_filter = o->testfilter_ik ? pow(5.0, 1. / 3.) * filter : pow(sum_weight, 1. / 3.);
                const double uf[3] = {ucat_f[3 * id], ucat_f[3 * id + 1], ucat_f[3 * id + 2]};
                double Uf[3];
                for (int b = 0; b < 3; b++) {
                    const double *m = met[b] + 3 * id;
                    Uf[b] = uf[0] * m[0] + uf[1] * m[1] + uf[2] * m[2];
                }
                double dc[3], de[3], dz[3], D[3][3], Sh[3][3];
                du_center(c, i, j, k, ucat_f, nvert, dc, de, dz);
                du_dxyz(g, id, dc, de, dz, D);
                const double S_hat = strain_abs(D, Sh);
                double L[3][3], SSh[3][3], Mc[3][3], M[3][3], gg[3][3], gc[3][3], G[3][3];
                for (int b = 0; b < 3; b++)
                    for (int a = 0; a < 3; a++) L[b][a] = tfilter(UU[b][a], w, o->testfilter_ik) - Uf[b] * uf[a];
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) SSh[a][b] = tfilter(SS[a][b], w, o->testfilter_ik);
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) gg[a][b] = met[a][3 * id + b];
                covariant_metrics(gg, gc);
                /* G = (x_xi)^T (x_xi): dot products of the covariant base vectors */
                for (int a = 0; a < 3; a++)
                    G[a][a] = gc[0][a] * gc[0][a] + gc[1][a] * gc[1][a] + gc[2][a] * gc[2][a];
                G[0][1] = G[1][0] = gc[0][1] * gc[0][0] + gc[1][1] * gc[1][0] + gc[2][1] * gc[2][0];
                G[0][2] = G[2][0] = gc[0][2] * gc[0][0] + gc[1][2] * gc[1][0] + gc[2][2] * gc[2][0];
                G[1][2] = G[2][1] = gc[0][1] * gc[0][2] + gc[1][1] * gc[1][2] + gc[2][1] * gc[2][2];
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++)
                        Mc[a][b] = -sq(test_filter) * S_hat * Sh[a][b] + sq(filter) * SSh[a][b];
                for (int a = 0; a < 3; a++)
                    for (int q = 0; q < 3; q++)
                        M[a][q] = Mc[a][0] * gg[q][0] + Mc[a][1] * gg[q][1] + Mc[a][2] * gg[q][2];
                double N[3][3] = {{0}};
                if (c->clark) {   /* les.c:489-548: N = 4 R - <T>, contravariant in its second index */
                    double Nc[3][3];
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) {
                            const double Rab = (D[a][0] * D[b][0] * dl2[0] + D[a][1] * D[b][1] * dl2[1] +
                                                D[a][2] * D[b][2] * dl2[2]) / 12.;
                            Nc[a][b] = 4.0 * Rab - tfilter(TT[a][b], w, o->testfilter_ik);
                        }
                    for (int a = 0; a < 3; a++)
                        for (int q = 0; q < 3; q++)
                            N[a][q] = Nc[a][0] * gg[q][0] + Nc[a][1] * gg[q][1] + Nc[a][2] * gg[q][2];
                }
                double num = 0, denom = 0;
                for (int q = 0; q < 3; q++)
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) {
                            num += L[b][a] * M[a][q] * G[b][q];
                            if (c->clark) num -= N[a][b] * M[a][q] * G[b][q];   /* les.c:648 */
                        }
                for (int m = 0; m < 3; m++)
                    for (int n = 0; n < 3; n++)
                        for (int l = 0; l < 3; l++) denom += M[n][m] * M[n][l] * G[m][l];
                LM[id] = num;
                MM[id] = denom;
            }
    per_copy_sca(c, LM);   /* les.c:678-706 */
    per_copy_sca(c, MM);
    return 0;
}

/* les.c:708-789, third pass: Cs = max(0.5 <LM> / (<MM> + eps), 0) with <.>
 * the 1/aj-weighted box filter (LM, MM themselves with a homogeneous
 * direction or -les 3; -les 3 relaxes towards the previous Cs once ti >= 100
 * or on a restart); the homogeneous-direction averages (les.c:791-942); the
 * clipping (943-994): 0 in solids and on every ghost point, at least wall_cs
 * at IB nodes, at most max_cs; then periodic copies (les.c:1018-1047).
 * Cs: in (the previous Cs, read by -les 3) and out. */
int vo_dsm_pass3(const vo_cfg *c, const vo_geom *g, const double *nvert, const double *LM, const double *MM,
                 const vo_les_opt *o, double *Cs)
{
    DIMS;
    int rc = check_cfg(c);
    if (rc) return rc;
    o = OPT(o);
    const int homo = o->i_homo_filter || o->j_homo_filter || o->k_homo_filter;
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                const size_t id = IX(k, j, i);
                if (nvert[id] > 1.1) { Cs[id] = 0; continue; }
                double LM_avg, MM_avg;
                if (homo || o->les == 3) {
                    LM_avg = LM[id];
                    MM_avg = MM[id];
                } else {
                    double w[3][3][3], L0[3][3][3], M0[3][3][3];
                    cs_weight(c, g, nvert, LM, i, j, k, w, L0);
                    cs_weight(c, g, nvert, MM, i, j, k, w, M0);
                    LM_avg = tfilter(L0, w, o->testfilter_ik);
                    MM_avg = tfilter(M0, w, o->testfilter_ik);
                }
                double C = 0.5 * LM_avg / (MM_avg + LES_EPS);
                if (o->les == 3 && !(o->ti < 100 && o->tistart == 0 && !o->rstart))
                    C = (1.0 - 0.001) * Cs[id] + 0.001 * C;
                Cs[id] = PMAX(C, 0);
            }
    if (o->les != 3 && o->i_homo_filter && o->k_homo_filter) {
        /* les.c:791-829: averages over the i-k planes of the fluid cells */
        double *sl = calloc(2 * (size_t)my, sizeof(double));
        int *cnt = calloc((size_t)my, sizeof(int));
        if (!sl || !cnt) { free(sl); free(cnt); return -9; }
        for (int k = 1; k < mz - 1; k++)
            for (int j = 1; j < my - 1; j++)
                for (int i = 1; i < mx - 1; i++)
                    if (nvert[IX(k, j, i)] < 0.1) {
                        sl[j] += LM[IX(k, j, i)];
                        sl[my + j] += MM[IX(k, j, i)];
                        cnt[j]++;
                    }
        for (int j = 0; j < my; j++)
            if (cnt[j] > 0) {
                sl[j] /= (double)(cnt[j]);
                sl[my + j] /= (double)(cnt[j]);
            }
        for (int j = 1; j < my - 1; j++)
            for (int k = 1; k < mz - 1; k++)
                for (int i = 1; i < mx - 1; i++) Cs[IX(k, j, i)] = 0.5 * sl[j] / (sl[my + j] + LES_EPS);
        free(sl); free(cnt);
    } else if (o->les != 3 && homo) {
        /* les.c:830-942: averages along the homogeneous direction, for the
         * fluid cells; lines indexed (k,j), (k,i) or (j,i) */
        const int dir = o->i_homo_filter ? 0 : (o->j_homo_filter ? 1 : 2);
        const int n1 = dir == 2 ? my : mz, n2 = dir == 0 ? my : mx;   /* line (p1, p2) */
        const size_t nl = (size_t)n1 * n2;
        double *sl = calloc(2 * nl, sizeof(double));
        int *cnt = calloc(nl, sizeof(int));
        if (!sl || !cnt) { free(sl); free(cnt); return -9; }
        const int len = dir == 0 ? mx : (dir == 1 ? my : mz);
        for (int p1 = 1; p1 < n1 - 1; p1++)
            for (int p2 = 1; p2 < n2 - 1; p2++)
                for (int q = 1; q < len - 1; q++) {
                    const int k = dir == 2 ? q : p1, j = dir == 0 ? p2 : (dir == 1 ? q : p1);
                    const int i = dir == 0 ? q : p2;
                    const size_t id = IX(k, j, i), pos = (size_t)p1 * n2 + p2;
                    if (nvert[id] < 0.1) { sl[pos] += LM[id]; sl[nl + pos] += MM[id]; cnt[pos]++; }
                }
        for (size_t pos = 0; pos < nl; pos++)
            if (cnt[pos] > 0) {
                const double N = (double)(cnt[pos]);
                sl[pos] /= N;
                sl[nl + pos] /= N;
            }
        for (int p1 = 1; p1 < n1 - 1; p1++)
            for (int p2 = 1; p2 < n2 - 1; p2++)
                for (int q = 1; q < len - 1; q++) {
                    const int k = dir == 2 ? q : p1, j = dir == 0 ? p2 : (dir == 1 ? q : p1);
                    const int i = dir == 0 ? q : p2;
                    const size_t id = IX(k, j, i), pos = (size_t)p1 * n2 + p2;
                    if (nvert[id] < 0.1) Cs[id] = 0.5 * sl[pos] / (sl[nl + pos] + LES_EPS);
                }
        free(sl); free(cnt);
    }
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                const size_t id = IX(k, j, i);
                if (nvert[id] > 1.1 || k == 0 || k == mz - 1 || j == 0 || j == my - 1 || i == 0 || i == mx - 1) {
                    Cs[id] = 0;
                } else {
                    if (nvert[id] > 0.1 && nvert[id] < 1.1) Cs[id] = PMAX(WALL_CS, Cs[id]);
                    Cs[id] = PMIN(PMAX(Cs[id], 0), o->max_cs);
                }
            }
    per_copy_sca(c, Cs);
    return 0;
}

/* les.c:1049-1107: the fluid cells' Cs replaced by its 1/aj-weighted box
 * filter (of the clipped field); IB nodes, solids and ghosts keep theirs. */
int vo_dsm_smooth(const vo_cfg *c, const vo_geom *g, const double *nvert, const vo_les_opt *o, double *Cs)
{
    DIMS;
    int rc = check_cfg(c);
    if (rc) return rc;
    o = OPT(o);
    const size_t P = (size_t)mx * my * mz;
    double *tmp = malloc(P * sizeof(double));
    if (!tmp) return -9;
    memcpy(tmp, Cs, P * sizeof(double));
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                const size_t id = IX(k, j, i);
                if (!(nvert[id] < 0.1)) continue;
                double w[3][3][3], C[3][3][3];
                cs_weight(c, g, nvert, tmp, i, j, k, w, C);
                Cs[id] = tfilter(C, w, o->testfilter_ik);
            }
    free(tmp);
    return 0;
}

/* les.c:67-1133 with the fresh-start rule (Cs = 0 while ti < 2 on a run
 * that is not a restart, 69-72) and les == 1 (Cs = 0.01, 74-77).  Cs: in
 * (the previous call's Cs, read by -les 3) and out, every point. */
int vo_smagorinsky_constant(const vo_cfg *c, const vo_geom *g, const double *nvert, const double *ucat,
                            const vo_les_opt *o, double *Cs)
{
    DIMS;
    int rc = check_cfg(c);
    if (rc) return rc;
    o = OPT(o);
    const size_t P = (size_t)mx * my * mz;
    if (o->les < 1 || o->les > 3) return -4;
    if (o->ti < 2 && o->tistart == 0 && !o->rstart) {
        for (size_t n = 0; n < P; n++) Cs[n] = 0.0;
        return 0;
    }
    if (o->les == 1) {
        for (size_t n = 0; n < P; n++) Cs[n] = 0.01;
        return 0;
    }
    double *grad = malloc(9 * P * sizeof(double)), *sabs = malloc(P * sizeof(double));
    double *uf = malloc(3 * P * sizeof(double)), *LM = malloc(P * sizeof(double)), *MM = malloc(P * sizeof(double));
    if (!grad || !sabs || !uf || !LM || !MM) rc = -9;
    if (!rc) rc = vo_dsm_pass1(c, g, nvert, ucat, o, grad, sabs, uf);
    if (!rc) rc = vo_dsm_pass2(c, g, nvert, ucat, o, grad, sabs, uf, LM, MM);
    if (!rc) rc = vo_dsm_pass3(c, g, nvert, LM, MM, o, Cs);
    if (!rc && !(o->i_homo_filter || o->j_homo_filter || o->k_homo_filter) && o->les != 3)
        rc = vo_dsm_smooth(c, g, nvert, o, Cs);
    free(grad); free(sabs); free(uf); free(LM); free(MM);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* momentum.c:197-440 Pressure_Gradient                                      */
/* ------------------------------------------------------------------------ */
#define IB1(x) ((int)((x) + 0.5) == 1)
int vo_pressure_gradient(const vo_cfg *c, const vo_geom *g, const double *nvert,
                         double *p, double *dp)
{
    if (c->k_periodic && !c->dpdz_set) return -5; /* the flux forcing needs vo_pressure_gradient_kf */
    return vo_pressure_gradient_kf(c, g, nvert, p, dp, 0.0, -1.0, 1.0);
}

/* rhs2.c:1562-1645 Calc_k_Flux (no levelset): the fluid (nvert(k-1) +
 * nvert(k) < 0.1) k flux summed over k = 1..mz-2, j, i, divided by mz-2; and
 * the mean over the planes of their fluid |zet| (bcs.c:254-278) */
int vo_calc_k_flux(const vo_cfg *c, const vo_geom *g, const double *nvert, const double *ucont, double *mean_k_flux,
                   double *mean_k_area)
{
    DIMS;
    double f = 0, a = 0;
    for (int k = 1; k < mz - 1; k++) {
        double ak = 0;
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++)
                if (nvert[IX(k - 1, j, i)] + nvert[IX(k, j, i)] < 0.1) {
                    f += ucont[3 * IX(k, j, i) + 2];
                    const double *z = g->zet + 3 * IX(k, j, i);
                    ak += sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
                }
        a += ak;
    }
    *mean_k_flux = f / (mz - 2);
    *mean_k_area = a / (double)(mz - 2);
    return 0;
}

int vo_pressure_gradient_kf(const vo_cfg *c, const vo_geom *g, const double *nvert, double *p, double *dp,
                            double mean_k_flux, double inlet_flux, double mean_k_area)
{
    DIMS;
    int rc = check_cfg(c);
    if (rc) return rc;
    per_copy_sca(c, p);
    size_t P = (size_t)mx * my * mz;
    memset(dp, 0, 3 * P * sizeof(double));
#define PP(k, j, i) p[IX(k, j, i)]
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t id = IX(k, j, i);
                const double *ic = g->icsi + 3 * id, *ie = g->ieta + 3 * id, *iz = g->izet + 3 * id;
                const double *jc = g->jcsi + 3 * id, *je = g->jeta + 3 * id, *jz = g->jzet + 3 * id;
                const double *kc = g->kcsi + 3 * id, *ke = g->keta + 3 * id, *kz = g->kzet + 3 * id;
                double g11_i = (ic[0] * ic[0] + ic[1] * ic[1] + ic[2] * ic[2]);
                double g12_i = (ie[0] * ic[0] + ie[1] * ic[1] + ie[2] * ic[2]);
                double g13_i = (iz[0] * ic[0] + iz[1] * ic[1] + iz[2] * ic[2]);
                double g21_j = (jc[0] * je[0] + jc[1] * je[1] + jc[2] * je[2]);
                double g22_j = (je[0] * je[0] + je[1] * je[1] + je[2] * je[2]);
                double g23_j = (jz[0] * je[0] + jz[1] * je[1] + jz[2] * je[2]);
                double g31_k = (kc[0] * kz[0] + kc[1] * kz[1] + kc[2] * kz[2]);
                double g32_k = (ke[0] * kz[0] + ke[1] * kz[1] + ke[2] * kz[2]);
                double g33_k = (kz[0] * kz[0] + kz[1] * kz[1] + kz[2] * kz[2]);
                double r1 = 1.0, r2 = 1.0, r3 = 1.0;
                double dpdc, dpde, dpdz;
                const int ip = c->i_periodic, jp = c->j_periodic, kp = c->k_periodic;

                if ((i == 0 || i == mx - 2) && ip) dpdc = PP(k, j, 1) - PP(k, j, i);
                else dpdc = PP(k, j, i + 1) - PP(k, j, i);
                if (IB1(NV(k, j + 1, i)) || IB1(NV(k, j + 1, i + 1)) || (j == my - 2 && !jp))
                    dpde = (PP(k, j, i) - PP(k, j - 1, i) + PP(k, j, i + 1) - PP(k, j - 1, i + 1)) * 0.5;
                else if (IB1(NV(k, j - 1, i)) || IB1(NV(k, j - 1, i + 1)) || (j == 1 && !jp))
                    dpde = (PP(k, j + 1, i) - PP(k, j, i) + PP(k, j + 1, i + 1) - PP(k, j, i + 1)) * 0.5;
                else dpde = (PP(k, j + 1, i) - PP(k, j - 1, i) + PP(k, j + 1, i + 1) - PP(k, j - 1, i + 1)) * 0.25;
                if (IB1(NV(k + 1, j, i)) || IB1(NV(k + 1, j, i + 1)) || (k == mz - 2 && !kp))
                    dpdz = (PP(k, j, i) - PP(k - 1, j, i) + PP(k, j, i + 1) - PP(k - 1, j, i + 1)) * 0.5;
                else if (IB1(NV(k - 1, j, i)) || IB1(NV(k - 1, j, i + 1)) || (k == 1 && !kp))
                    dpdz = (PP(k + 1, j, i) - PP(k, j, i) + PP(k + 1, j, i + 1) - PP(k, j, i + 1)) * 0.5;
                else dpdz = (PP(k + 1, j, i) - PP(k - 1, j, i) + PP(k + 1, j, i + 1) - PP(k - 1, j, i + 1)) * 0.25;
                VX(dp, id) = (dpdc * g11_i + dpde * g12_i + dpdz * g13_i) * g->iaj[id] / r1;

                if (IB1(NV(k, j, i + 1)) || IB1(NV(k, j + 1, i + 1)) || (i == mx - 2 && !ip))
                    dpdc = (PP(k, j, i) - PP(k, j, i - 1) + PP(k, j + 1, i) - PP(k, j + 1, i - 1)) * 0.5;
                else if (IB1(NV(k, j, i - 1)) || IB1(NV(k, j + 1, i - 1)) || (i == 1 && !ip))
                    dpdc = (PP(k, j, i + 1) - PP(k, j, i) + PP(k, j + 1, i + 1) - PP(k, j + 1, i)) * 0.5;
                else dpdc = (PP(k, j, i + 1) - PP(k, j, i - 1) + PP(k, j + 1, i + 1) - PP(k, j + 1, i - 1)) * 0.25;
                if ((j == 0 || j == my - 2) && jp) dpde = PP(k, 1, i) - PP(k, j, i);
                else dpde = PP(k, j + 1, i) - PP(k, j, i);
                if (IB1(NV(k + 1, j, i)) || IB1(NV(k + 1, j + 1, i)) || (k == mz - 2 && !kp))
                    dpdz = (PP(k, j, i) - PP(k - 1, j, i) + PP(k, j + 1, i) - PP(k - 1, j + 1, i)) * 0.5;
                else if (IB1(NV(k - 1, j, i)) || IB1(NV(k - 1, j + 1, i)) || (k == 1 && !kp))
                    dpdz = (PP(k + 1, j, i) - PP(k, j, i) + PP(k + 1, j + 1, i) - PP(k, j + 1, i)) * 0.5;
                else dpdz = (PP(k + 1, j, i) - PP(k - 1, j, i) + PP(k + 1, j + 1, i) - PP(k - 1, j + 1, i)) * 0.25;
                VY(dp, id) = (dpdc * g21_j + dpde * g22_j + dpdz * g23_j) * g->jaj[id] / r2;

                if (IB1(NV(k, j, i + 1)) || IB1(NV(k + 1, j, i + 1)) || (i == mx - 2 && !ip))
                    dpdc = (PP(k, j, i) - PP(k, j, i - 1) + PP(k + 1, j, i) - PP(k + 1, j, i - 1)) * 0.5;
                else if (IB1(NV(k, j, i - 1)) || IB1(NV(k + 1, j, i - 1)) || (i == 1 && !ip))
                    dpdc = (PP(k, j, i + 1) - PP(k, j, i) + PP(k + 1, j, i + 1) - PP(k + 1, j, i)) * 0.5;
                else dpdc = (PP(k, j, i + 1) - PP(k, j, i - 1) + PP(k + 1, j, i + 1) - PP(k + 1, j, i - 1)) * 0.25;
                if (IB1(NV(k, j + 1, i)) || IB1(NV(k + 1, j + 1, i)) || (j == my - 2 && !jp))
                    dpde = (PP(k, j, i) - PP(k, j - 1, i) + PP(k + 1, j, i) - PP(k + 1, j - 1, i)) * 0.5;
                else if (IB1(NV(k, j - 1, i)) || IB1(NV(k + 1, j - 1, i)) || (j == 1 && !jp))
                    dpde = (PP(k, j + 1, i) - PP(k, j, i) + PP(k + 1, j + 1, i) - PP(k + 1, j, i)) * 0.5;
                else dpde = (PP(k, j + 1, i) - PP(k, j - 1, i) + PP(k + 1, j + 1, i) - PP(k + 1, j - 1, i)) * 0.25;
                if ((k == 0 || k == mz - 2) && kp) dpdz = PP(1, j, i) - PP(k, j, i);
                else dpdz = (PP(k + 1, j, i) - PP(k, j, i));
                if (kp) { /* momentum.c:382-398 */
                    double dzz = 1. / g->kaj[id] / sqrt(kz[0] * kz[0] + kz[1] * kz[1] + kz[2] * kz[2]);
                    if (c->dpdz_set) dpdz += c->mean_pressure_gradient * dzz;
                    else dpdz += dzz * (mean_k_flux - inlet_flux) / c->dt / mean_k_area;
                }
                VZ(dp, id) = (dpdc * g31_k + dpde * g32_k + dpdz * g33_k) * g->kaj[id] / r3;

                if (i == 0 || NV(k, j, i) + NV(k, j, i + 1) > 0.1 || (!ip && i == mx - 2)) VX(dp, id) = 0;
                if (j == 0 || NV(k, j, i) + NV(k, j + 1, i) > 0.1 || (!jp && j == my - 2)) VY(dp, id) = 0;
                if (k == 0 || NV(k, j, i) + NV(k + 1, j, i) > 0.1 || (!kp && k == mz - 2)) VZ(dp, id) = 0;
            }
#undef PP
    return 0;
}

/* ------------------------------------------------------------------------ */
/* momentum.c:1982-2195 IB_BC                                                */
/* ------------------------------------------------------------------------ */
int vo_ib_bc(const vo_cfg *c, const vo_geom *g, const double *nvert, double *ucont, double *ucat)
{
    DIMS;
    int rc = check_cfg(c);
    if (rc) return rc;
    per_copy_vec(c, ucat); /* momentum.c:2050-2075 */
    if (c->immersed) {
        for (int k = 1; k < mz - 1; k++)
            for (int j = 1; j < my - 1; j++)
                for (int i = 1; i < mx - 1; i++) {
                    size_t id = IX(k, j, i);
                    double f = 1.0, ucx, ucy, ucz;
                    const double *ic = g->icsi + 3 * id, *je = g->jeta + 3 * id, *kz = g->kzet + 3 * id;
                    if ((int)nvert[id] == 1) {
                        ucx = (U3(k, j, i, 0) + U3(k, j, i + 1, 0)) * 0.5;
                        ucy = (U3(k, j, i, 1) + U3(k, j, i + 1, 1)) * 0.5;
                        ucz = (U3(k, j, i, 2) + U3(k, j, i + 1, 2)) * 0.5;
                        VX(ucont, id) = (ucx * ic[0] + ucy * ic[1] + ucz * ic[2]) * f;
                        ucx = (U3(k, j + 1, i, 0) + U3(k, j, i, 0)) * 0.5;
                        ucy = (U3(k, j + 1, i, 1) + U3(k, j, i, 1)) * 0.5;
                        ucz = (U3(k, j + 1, i, 2) + U3(k, j, i, 2)) * 0.5;
                        VY(ucont, id) = (ucx * je[0] + ucy * je[1] + ucz * je[2]) * f;
                        ucx = (U3(k + 1, j, i, 0) + U3(k, j, i, 0)) * 0.5;
                        ucy = (U3(k + 1, j, i, 1) + U3(k, j, i, 1)) * 0.5;
                        ucz = (U3(k + 1, j, i, 2) + U3(k, j, i, 2)) * 0.5;
                        VZ(ucont, id) = (ucx * kz[0] + ucy * kz[1] + ucz * kz[2]) * f;
                    }
                    if ((int)(NV(k, j, i + 1)) == 1) {
                        ucx = (U3(k, j, i, 0) + U3(k, j, i + 1, 0)) * 0.5;
                        ucy = (U3(k, j, i, 1) + U3(k, j, i + 1, 1)) * 0.5;
                        ucz = (U3(k, j, i, 2) + U3(k, j, i + 1, 2)) * 0.5;
                        VX(ucont, id) = (ucx * ic[0] + ucy * ic[1] + ucz * ic[2]) * f;
                    }
                    if ((int)(NV(k, j + 1, i)) == 1) {
                        ucx = (U3(k, j + 1, i, 0) + U3(k, j, i, 0)) * 0.5;
                        ucy = (U3(k, j + 1, i, 1) + U3(k, j, i, 1)) * 0.5;
                        ucz = (U3(k, j + 1, i, 2) + U3(k, j, i, 2)) * 0.5;
                        VY(ucont, id) = (ucx * je[0] + ucy * je[1] + ucz * je[2]) * f;
                    }
                    if ((int)(NV(k + 1, j, i)) == 1) {
                        ucx = (U3(k + 1, j, i, 0) + U3(k, j, i, 0)) * 0.5;
                        ucy = (U3(k + 1, j, i, 1) + U3(k, j, i, 1)) * 0.5;
                        ucz = (U3(k + 1, j, i, 2) + U3(k, j, i, 2)) * 0.5;
                        VZ(ucont, id) = (ucx * kz[0] + ucy * kz[1] + ucz * kz[2]) * f;
                    }
                }
    }
    /* wall-function walls: the wall face flux is zero (momentum.c:2135-2153) */
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                if ((c->bctype[0] == -1 && i == 1) || (c->bctype[1] == -1 && i == mx - 2)) {
                    if (i == 1) VX(ucont, IX(k, j, i - 1)) = 0;
                    else VX(ucont, IX(k, j, i)) = 0;
                }
                if ((c->bctype[2] == -1 && j == 1) || (c->bctype[3] == -1 && j == my - 2)) {
                    if (j == 1) VY(ucont, IX(k, j - 1, i)) = 0;
                    else VY(ucont, IX(k, j, i)) = 0;
                }
            }
    /* periodic face copies, momentum.c:2167-2181 */
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                if (c->i_periodic && i == 0) VX(ucont, IX(k, j, 0)) = VX(ucont, IX(k, j, mx - 2));
                if (c->i_periodic && i == mx - 1) VX(ucont, IX(k, j, mx - 1)) = VX(ucont, IX(k, j, 1));
                if (c->j_periodic && j == 0) VY(ucont, IX(k, 0, i)) = VY(ucont, IX(k, my - 2, i));
                if (c->j_periodic && j == my - 1) VY(ucont, IX(k, my - 1, i)) = VY(ucont, IX(k, 1, i));
                if (c->k_periodic && k == 0) VZ(ucont, IX(0, j, i)) = VZ(ucont, IX(mz - 2, j, i));
                if (c->k_periodic && k == mz - 1) VZ(ucont, IX(mz - 1, j, i)) = VZ(ucont, IX(1, j, i));
            }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* momentum.c:456-1979 Formfunction_2                                        */
/* ------------------------------------------------------------------------ */
/* rhs2.c:592-643 Calculate_normal_and_area + Calculate_dxdydz: the length
 * scales of a cell along x, y, z (volume 1/ajc over the face areas, along the
 * unit covariant base vectors) */
static void dxdydz(double ajc, const double *csi, const double *eta, const double *zet, double *dx, double *dy,
                   double *dz)
{
    double ni[3], nj[3], nk[3];
    const double vol = 1. / ajc;
    calc_normal(csi, eta, zet, ni, nj, nk);
    const double Ai = sqrt(csi[0] * csi[0] + csi[1] * csi[1] + csi[2] * csi[2]);
    const double Aj = sqrt(eta[0] * eta[0] + eta[1] * eta[1] + eta[2] * eta[2]);
    const double Ak = sqrt(zet[0] * zet[0] + zet[1] * zet[1] + zet[2] * zet[2]);
    const double Li = vol / Ai, Lj = vol / Aj, Lk = vol / Ak;
    *dx = fabs(Li * ni[0] + Lj * nj[0] + Lk * nk[0]);
    *dy = fabs(Li * ni[1] + Lj * nj[1] + Lk * nk[1]);
    *dz = fabs(Li * ni[2] + Lj * nj[2] + Lk * nk[2]);
}

int vo_formfunction2(const vo_cfg *c, const vo_geom *g, const double *nvert,
                     double *ucont, const double *ucat, const double *nu_t,
                     double *rhs, double scale)
{
    DIMS;
    int rc = check_cfg(c);
    if (rc) return rc;
    const int ip = c->i_periodic, jp = c->j_periodic, kp = c->k_periodic;
    size_t P = (size_t)mx * my * mz;
    double *buf = (double *)calloc(30 * P, sizeof(double));
    if (!buf) return -1;
    double *div1 = buf, *div2 = buf + 3 * P, *div3 = buf + 6 * P;
    double *visc1 = buf + 9 * P, *visc2 = buf + 12 * P, *visc3 = buf + 15 * P, *fp = buf + 18 * P;
    /* -skew: the advective face fluxes adv1..3 (momentum.c:610-618) */
    double *adv1 = buf + 21 * P, *adv2 = buf + 24 * P, *adv3 = buf + 27 * P;

    per_copy_vec(c, ucont); /* momentum.c:642-669 */

    for (int dir = 0; dir < 3; dir++) {
        const double *fc = dir == 0 ? g->icsi : dir == 1 ? g->jcsi : g->kcsi;
        const double *fe = dir == 0 ? g->ieta : dir == 1 ? g->jeta : g->keta;
        const double *fz = dir == 0 ? g->izet : dir == 1 ? g->jzet : g->kzet;
        const double *faj = dir == 0 ? g->iaj : dir == 1 ? g->jaj : g->kaj;
        double *dv = dir == 0 ? div1 : dir == 1 ? div2 : div3;
        double *vs = dir == 0 ? visc1 : dir == 1 ? visc2 : visc3;
        double *av = dir == 0 ? adv1 : dir == 1 ? adv2 : adv3;
        /* face loops: i-faces momentum.c:672-931, j 934-1166, k 1169-1409 */
#pragma omp parallel for schedule(static)
        for (int k = (dir == 2 ? 0 : 1); k < mz - 1; k++)
            for (int j = (dir == 1 ? 0 : 1); j < my - 1; j++)
                for (int i = (dir == 0 ? 0 : 1); i < mx - 1; i++) {
                    size_t id = IX(k, j, i);
                    double ajc = faj[id];
                    double csi0 = fc[3 * id], csi1 = fc[3 * id + 1], csi2 = fc[3 * id + 2];
                    double eta0 = fe[3 * id], eta1 = fe[3 * id + 1], eta2 = fe[3 * id + 2];
                    double zet0 = fz[3 * id], zet1 = fz[3 * id + 1], zet2 = fz[3 * id + 2];
                    double dc[3], de[3], dz[3];
                    if (dir == 0) du_i(c, i, j, k, ucat, nvert, dc, de, dz);
                    else if (dir == 1) du_j(c, i, j, k, ucat, nvert, dc, de, dz);
                    else du_k(c, i, j, k, ucat, nvert, dc, de, dz);
                    double g11, g21, g31, nrm0, nrm1, nrm2;
                    if (dir == 0) {
                        g11 = csi0 * csi0 + csi1 * csi1 + csi2 * csi2;
                        g21 = eta0 * csi0 + eta1 * csi1 + eta2 * csi2;
                        g31 = zet0 * csi0 + zet1 * csi1 + zet2 * csi2;
                        nrm0 = csi0, nrm1 = csi1, nrm2 = csi2;
                    } else if (dir == 1) {
                        g11 = csi0 * eta0 + csi1 * eta1 + csi2 * eta2;
                        g21 = eta0 * eta0 + eta1 * eta1 + eta2 * eta2;
                        g31 = zet0 * eta0 + zet1 * eta1 + zet2 * eta2;
                        nrm0 = eta0, nrm1 = eta1, nrm2 = eta2;
                    } else {
                        g11 = csi0 * zet0 + csi1 * zet1 + csi2 * zet2;
                        g21 = eta0 * zet0 + eta1 * zet1 + eta2 * zet2;
                        g31 = zet0 * zet0 + zet1 * zet1 + zet2 * zet2;
                        nrm0 = zet0, nrm1 = zet1, nrm2 = zet2;
                    }
                    double r11 = dc[0] * csi0 + de[0] * eta0 + dz[0] * zet0;
                    double r21 = dc[1] * csi0 + de[1] * eta0 + dz[1] * zet0;
                    double r31 = dc[2] * csi0 + de[2] * eta0 + dz[2] * zet0;
                    double r12 = dc[0] * csi1 + de[0] * eta1 + dz[0] * zet1;
                    double r22 = dc[1] * csi1 + de[1] * eta1 + dz[1] * zet1;
                    double r32 = dc[2] * csi1 + de[2] * eta1 + dz[2] * zet1;
                    double r13 = dc[0] * csi2 + de[0] * eta2 + dz[0] * zet2;
                    double r23 = dc[1] * csi2 + de[1] * eta2 + dz[1] * zet2;
                    double r33 = dc[2] * csi2 + de[2] * eta2 + dz[2] * zet2;

                    /* convective scheme selection */
                    int n = dir == 0 ? i : dir == 1 ? j : k;
                    int m = dir == 0 ? mx : dir == 1 ? my : mz;
                    int per = dir == 0 ? ip : dir == 1 ? jp : kp;
                    int L = n - 1, R = n + 2;
                    /* denom: the divisor of the 4th-order ucat difference of -skew;
                     * next to a body only the k faces reset it (momentum.c:705-723,
                     * 967-985, 1202-1219) */
                    double denom = 3.;
#define NB(t) (dir == 0 ? NV(k, j, t) : dir == 1 ? NV(k, t, i) : NV(t, j, i))
#define UB(t, q) (dir == 0 ? U3(k, j, t, q) : dir == 1 ? U3(k, t, i, q) : U3(t, j, i, q))
#define UC(t) (dir == 0 ? ucont[3 * IX(k, j, t) + 0] : dir == 1 ? ucont[3 * IX(k, t, i) + 1] : ucont[3 * IX(t, j, i) + 2])
                    if (n == 0 || n == m - 2) {
                        if (per) L = m - 3, R = 2;
                        else L = n, R = n + 1, denom = 1.;
                    } else if (NB(L) + NB(R) > 0.1) {
                        L = n, R = n + 1;
                        if (dir == 2) denom = 1.;
                    }
                    if (c->second_order) L = n, R = n + 1, denom = 1.;
                    double ucon_f = UC(n);
                    double ucon = ucon_f;
                    if (per && (n == 0 || n == m - 2)) ucon = UC(m - 2);
                    /* outlet: the last face convects with the flux before it (momentum.c:1227-1229) */
                    if (dir == 2 && n == m - 2 && c->bctype[5] == 4 && (int)NV(k, j, i) == 0) ucon = UC(n - 1);
                    double up = -0.5 * (ucon + fabs(ucon));
                    double um = -0.5 * (ucon - fabs(ucon));
                    double d3v[3];
                    if (c->immersed && n != m - 2 && NB(n) > 0.1) {
                        for (int q = 0; q < 3; q++)
                            d3v[q] = um * (0.125 * (-UB(n + 2, q) - 2. * UB(n + 1, q) + 3. * UB(n, q)) + UB(n + 1, q)) +
                                     up * (0.125 * (-UB(n, q) - 2. * UB(n, q) + 3. * UB(n + 1, q)) + UB(n, q));
                    } else if (c->immersed && n != 0 && NB(n + 1) > 0.1) {
                        for (int q = 0; q < 3; q++)
                            d3v[q] = um * (0.125 * (-UB(n + 1, q) - 2. * UB(n + 1, q) + 3. * UB(n, q)) + UB(n + 1, q)) +
                                     up * (0.125 * (-UB(n - 1, q) - 2. * UB(n, q) + 3. * UB(n + 1, q)) + UB(n, q));
                    } else if (c->inviscid) {   /* momentum.c:757-773: WENO3 whatever else is set */
                        for (int q = 0; q < 3; q++)
                            d3v[q] = -ucon_f * vo_weno3(UB(L, q), UB(n, q), UB(n + 1, q), UB(R, q), ucon_f);
                    } else if (c->second_order) {
                        for (int q = 0; q < 3; q++) d3v[q] = -ucon_f * 0.5 * (UB(n, q) + UB(n + 1, q));
                    } else if (c->weno) {   /* momentum.c:782-786, 1044-1048, 1283-1287 */
                        for (int q = 0; q < 3; q++)
                            d3v[q] = -ucon_f * vo_weno3(UB(L, q), UB(n, q), UB(n + 1, q), UB(R, q), ucon_f);
                    } else {
                        for (int q = 0; q < 3; q++)
                            d3v[q] = -ucon_f * 0.0625 * (-UB(L, q) + 9. * UB(n, q) + 9. * UB(n + 1, q) - UB(R, q));
                    }
                    double a3v[3] = {0, 0, 0};
                    if (c->skew) {   /* momentum.c:798-811, 1060-1071, 1298-1309 */
                        const double *dn = dir == 0 ? dc : dir == 1 ? de : dz;
                        for (int q = 0; q < 3; q++) {
                            d3v[q] *= 0.5;
                            const double d4 = (UB(R, q) - UB(L, q)) * (1. / denom);
                            a3v[q] = -0.5 * ucon_f * (9. / 8. * dn[q] - 1. / 8. * d4);
                        }
                    }
                    if (NB(n) + NB(n + 1) > 0.1) {
                        if (!c->immersed) d3v[0] = d3v[1] = d3v[2] = a3v[0] = a3v[1] = a3v[2] = 0;
                    }
                    dv[3 * id] = d3v[0]; dv[3 * id + 1] = d3v[1]; dv[3 * id + 2] = d3v[2];
                    if (c->skew) av[3 * id] = a3v[0], av[3 * id + 1] = a3v[1], av[3 * id + 2] = a3v[2];

                    double nu = 1. / c->ren, nut = 0;
                    double v0, v1, v2;
                    if (c->les) {
                        if ((n == 0 && !per) || NB(n) > 0.1) nut = dir == 0 ? nu_t[IX(k, j, i + 1)] : dir == 1 ? nu_t[IX(k, j + 1, i)] : nu_t[IX(k + 1, j, i)];
                        else if ((n == m - 2 && !per) || NB(n + 1) > 0.1) nut = nu_t[id];
                        else nut = 0.5 * (nu_t[id] + (dir == 0 ? nu_t[IX(k, j, i + 1)] : dir == 1 ? nu_t[IX(k, j + 1, i)] : nu_t[IX(k + 1, j, i)]));
                        v0 = (g11 * dc[0] + g21 * de[0] + g31 * dz[0] + r11 * nrm0 + r21 * nrm1 + r31 * nrm2) * ajc * (nut);
                        v1 = (g11 * dc[1] + g21 * de[1] + g31 * dz[1] + r12 * nrm0 + r22 * nrm1 + r32 * nrm2) * ajc * (nut);
                        v2 = (g11 * dc[2] + g21 * de[2] + g31 * dz[2] + r13 * nrm0 + r23 * nrm1 + r33 * nrm2) * ajc * (nut);
                    } else {
                        v0 = v1 = v2 = 0;
                    }
                    v0 += (g11 * dc[0] + g21 * de[0] + g31 * dz[0] + r11 * nrm0 + r21 * nrm1 + r31 * nrm2) * ajc * (nu);
                    v1 += (g11 * dc[1] + g21 * de[1] + g31 * dz[1] + r12 * nrm0 + r22 * nrm1 + r32 * nrm2) * ajc * (nu);
                    v2 += (g11 * dc[2] + g21 * de[2] + g31 * dz[2] + r13 * nrm0 + r23 * nrm1 + r33 * nrm2) * ajc * (nu);
                    if (c->clark) {   /* momentum.c:907-925, 1142-1160, 1385-1403 */
                        double dx, dy, dzl;
                        /* the cell metrics of point (k,j,i) with the face's Jacobian */
                        dxdydz(ajc, &g->csi[3 * id], &g->eta[3 * id], &g->zet[3 * id], &dx, &dy, &dzl);
                        const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dzl * dzl;
                        double du_dx, du_dy, du_dz, dv_dx, dv_dy, dv_dz, dw_dx, dw_dy, dw_dz;
                        du_dxyz_face(csi0, csi1, csi2, eta0, eta1, eta2, zet0, zet1, zet2, ajc, dc, de, dz, &du_dx, &dv_dx,
                                     &dw_dx, &du_dy, &dv_dy, &dw_dy, &du_dz, &dv_dz, &dw_dz);
                        const double t11 = (du_dx * du_dx * dx2 + du_dy * du_dy * dy2 + du_dz * du_dz * dz2);
                        const double t12 = (du_dx * dv_dx * dx2 + du_dy * dv_dy * dy2 + du_dz * dv_dz * dz2);
                        const double t13 = (du_dx * dw_dx * dx2 + du_dy * dw_dy * dy2 + du_dz * dw_dz * dz2);
                        const double t21 = t12;
                        const double t22 = (dv_dx * dv_dx * dx2 + dv_dy * dv_dy * dy2 + dv_dz * dv_dz * dz2);
                        const double t23 = (dv_dx * dw_dx * dx2 + dv_dy * dw_dy * dy2 + dv_dz * dw_dz * dz2);
                        const double t31 = t13, t32 = t23;
                        const double t33 = (dw_dx * dw_dx * dx2 + dw_dy * dw_dy * dy2 + dw_dz * dw_dz * dz2);
                        v0 -= (t11 * nrm0 + t12 * nrm1 + t13 * nrm2) / 12.;
                        v1 -= (t21 * nrm0 + t22 * nrm1 + t23 * nrm2) / 12.;
                        v2 -= (t31 * nrm0 + t32 * nrm1 + t33 * nrm2) / 12.;
                    }
                    vs[3 * id] = v0; vs[3 * id + 1] = v1; vs[3 * id + 2] = v2;
#undef NB
#undef UB
#undef UC
                }
    }
    /* periodic copies of the face arrays, momentum.c:1463-1504 */
    per_copy_vec(c, div1); per_copy_vec(c, div2); per_copy_vec(c, div3);
    per_copy_vec(c, visc1); per_copy_vec(c, visc2); per_copy_vec(c, visc3);
    if (c->skew) { per_copy_vec(c, adv1); per_copy_vec(c, adv2); per_copy_vec(c, adv3); }

    /* cell divergence with 4th-order deferred correction, momentum.c:1506-1636 */
#pragma omp parallel for schedule(static)
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t id = IX(k, j, i);
                size_t im = IX(k, j, i - 1), jm = IX(k, j - 1, i), km = IX(k - 1, j, i);
                double dv[3], d4[3] = {0, 0, 0}, f[3];
                for (int q = 0; q < 3; q++)
                    dv[q] = (div1[3 * id + q] - div1[3 * im + q] + div2[3 * id + q] - div2[3 * jm + q] +
                             div3[3 * id + q] - div3[3 * km + q]);
                int iR = i + 1, iL = i - 2, jR = j + 1, jL = j - 2, kR = k + 1, kL = k - 2;
                double den_i = 3., den_j = 3., den_k = 3.;
                if (i == 1) { if (ip) iL = mx - 3; else iR = i, iL = i - 1, den_i = 1.; }
                else if (i == 2 || i == mx - 3) { if (!ip) iR = i, iL = i - 1, den_i = 1.; }
                else if (i == mx - 2) { if (ip) iR = 1; else iR = i, iL = i - 1, den_i = 1.; }
                if (j == 1) { if (jp) jL = my - 3; else jR = j, jL = j - 1, den_j = 1.; }
                else if (j == 2 || j == my - 3) { if (!jp) jR = j, jL = j - 1, den_j = 1.; }
                else if (j == my - 2) { if (jp) jR = 1; else jR = j, jL = j - 1, den_j = 1.; }
                if (k == 1) { if (kp) kL = mz - 3; else kR = k, kL = k - 1, den_k = 1.; }
                else if (k == 2 || k == mz - 3) { if (!kp) kR = k, kL = k - 1, den_k = 1.; }
                else if (k == mz - 2) { if (kp) kR = 1; else kR = k, kL = k - 1, den_k = 1.; }
                if (NV(k, j, i - 1) + NV(k, j, i) + NV(k, j, i + 1) > 0.1) iR = i, iL = i - 1, den_i = 1.;
                if (NV(k, j - 1, i) + NV(k, j, i) + NV(k, j + 1, i) > 0.1) jR = j, jL = j - 1, den_j = 1.;
                if (NV(k - 1, j, i) + NV(k, j, i) + NV(k + 1, j, i) > 0.1) kR = k, kL = k - 1, den_k = 1.;
                if (c->second_order || c->inviscid) {   /* momentum.c:1582-1588 */
                    for (int q = 0; q < 3; q++) f[q] = dv[q];
                } else {
                    size_t a1 = IX(k, j, iR), b1 = IX(k, j, iL), a2 = IX(k, jR, i), b2 = IX(k, jL, i);
                    size_t a3 = IX(kR, j, i), b3 = IX(kL, j, i);
                    for (int q = 0; q < 3; q++) {
                        d4[q] += (div1[3 * a1 + q] - div1[3 * b1 + q]) * (1. / den_i);
                        d4[q] += (div2[3 * a2 + q] - div2[3 * b2 + q]) * (1. / den_j);
                        d4[q] += (div3[3 * a3 + q] - div3[3 * b3 + q]) * (1. / den_k);
                        f[q] = 9. / 8. * dv[q] + -1. / 8. * d4[q];
                    }
                }
                if (c->skew && !c->inviscid) {   /* momentum.c:1596-1608 */
                    if (c->second_order) iR = i, iL = i - 1, jR = j, jL = j - 1, kR = k, kL = k - 1;
                    const size_t a1 = IX(k, j, iR), b1 = IX(k, j, iL), a2 = IX(k, jR, i), b2 = IX(k, jL, i);
                    const size_t a3 = IX(kR, j, i), b3 = IX(kL, j, i);
                    for (int q = 0; q < 3; q++) {
                        f[q] += 9. / 8. * 0.5 * (adv1[3 * id + q] + adv1[3 * im + q]) - 1. / 8. * 0.5 * (adv1[3 * a1 + q] + adv1[3 * b1 + q]);
                        f[q] += 9. / 8. * 0.5 * (adv2[3 * id + q] + adv2[3 * jm + q]) - 1. / 8. * 0.5 * (adv2[3 * a2 + q] + adv2[3 * b2 + q]);
                        f[q] += 9. / 8. * 0.5 * (adv3[3 * id + q] + adv3[3 * km + q]) - 1. / 8. * 0.5 * (adv3[3 * a3 + q] + adv3[3 * b3 + q]);
                    }
                }
                if (!c->inviscid)   /* momentum.c:1621-1627 */
                    for (int q = 0; q < 3; q++)
                        f[q] += (visc1[3 * id + q] - visc1[3 * im + q] + visc2[3 * id + q] - visc2[3 * jm + q] +
                                 visc3[3 * id + q] - visc3[3 * km + q]) / 1.0;
                fp[3 * id] = f[0]; fp[3 * id + 1] = f[1]; fp[3 * id + 2] = f[2];
            }
    per_copy_vec(c, fp); /* momentum.c:1643-1671 */

    /* contravariant projection back onto faces, momentum.c:1673-1693 */
#pragma omp parallel for schedule(static)
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t id = IX(k, j, i), a = IX(k, j, i + 1), b = IX(k, j + 1, i), cc = IX(k + 1, j, i);
                const double *cs = g->csi, *et = g->eta, *ze = g->zet;
                rhs[3 * id] += scale * (0.5 * (cs[3 * id] * fp[3 * id] + cs[3 * id + 1] * fp[3 * id + 1] + cs[3 * id + 2] * fp[3 * id + 2]) +
                                        0.5 * (cs[3 * a] * fp[3 * a] + cs[3 * a + 1] * fp[3 * a + 1] + cs[3 * a + 2] * fp[3 * a + 2])) * g->iaj[id];
                rhs[3 * id + 1] += scale * (0.5 * (et[3 * id] * fp[3 * id] + et[3 * id + 1] * fp[3 * id + 1] + et[3 * id + 2] * fp[3 * id + 2]) +
                                            0.5 * (et[3 * b] * fp[3 * b] + et[3 * b + 1] * fp[3 * b + 1] + et[3 * b + 2] * fp[3 * b + 2])) * g->jaj[id];
                rhs[3 * id + 2] += scale * (0.5 * (ze[3 * id] * fp[3 * id] + ze[3 * id + 1] * fp[3 * id + 1] + ze[3 * id + 2] * fp[3 * id + 2]) +
                                            0.5 * (ze[3 * cc] * fp[3 * cc] + ze[3 * cc + 1] * fp[3 * cc + 1] + ze[3 * cc + 2] * fp[3 * cc + 2])) * g->kaj[id];
                if (NV(k, j, i) + NV(k, j, i + 1) > 0.1 || (!ip && i == mx - 2)) rhs[3 * id] = 0;
                if (NV(k, j, i) + NV(k, j + 1, i) > 0.1 || (!jp && j == my - 2)) rhs[3 * id + 1] = 0;
                if (NV(k, j, i) + NV(k + 1, j, i) > 0.1 || (!kp && k == mz - 2)) rhs[3 * id + 2] = 0;
            }
    /* boundary planes, momentum.c:1828-1900 */
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++) {
            size_t a = IX(k, j, 0);
            rhs[3 * a] = rhs[3 * a + 1] = rhs[3 * a + 2] = 0;
        }
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++) {
            if (!ip) rhs[3 * IX(k, j, mx - 2)] = 0;
            size_t a = IX(k, j, mx - 1);
            rhs[3 * a] = rhs[3 * a + 1] = rhs[3 * a + 2] = 0;
        }
    for (int k = 0; k < mz; k++)
        for (int i = 0; i < mx; i++) {
            size_t a = IX(k, 0, i);
            rhs[3 * a] = rhs[3 * a + 1] = rhs[3 * a + 2] = 0;
        }
    for (int k = 0; k < mz; k++)
        for (int i = 0; i < mx; i++) {
            if (!jp) rhs[3 * IX(k, my - 2, i) + 1] = 0;
            size_t a = IX(k, my - 1, i);
            rhs[3 * a] = rhs[3 * a + 1] = rhs[3 * a + 2] = 0;
        }
    for (int j = 0; j < my; j++)
        for (int i = 0; i < mx; i++) {
            size_t a = IX(0, j, i);
            rhs[3 * a] = rhs[3 * a + 1] = rhs[3 * a + 2] = 0;
        }
    for (int j = 0; j < my; j++)
        for (int i = 0; i < mx; i++) {
            if (!kp) rhs[3 * IX(mz - 2, j, i) + 2] = 0;
            size_t a = IX(mz - 1, j, i);
            rhs[3 * a] = rhs[3 * a + 1] = rhs[3 * a + 2] = 0;
        }
    free(buf);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* momentum.c:2197-2301 FormFunction_SNES (time_coeff()==1 branch)          */
/* ------------------------------------------------------------------------ */
int vo_formfunction_snes(const vo_cfg *c, const vo_geom *g, const double *nvert,
                         const double *U, const double *U_o, const double *rhs_o,
                         const double *dP, const double *nu_t, double *lucont,
                         double *ucat, double *rhs)
{
    return vo_formfunction_snes_f(c, g, nvert, U, U_o, rhs_o, dP, nu_t, NULL, lucont, ucat, rhs);
}

/* the same with the body force of the rotor/nacelle models added last
 * (momentum.c:2288); f_eul NULL: none */
int vo_formfunction_snes_f(const vo_cfg *c, const vo_geom *g, const double *nvert,
                           const double *U, const double *U_o, const double *rhs_o,
                           const double *dP, const double *nu_t, const double *f_eul, double *lucont,
                           double *ucat, double *rhs)
{
    DIMS;
    int rc = check_cfg(c);
    if (rc) return rc;
    size_t P = (size_t)mx * my * mz;
    double *Ug = (double *)malloc(3 * P * sizeof(double));
    if (!Ug) return -1;
    memcpy(Ug, U, 3 * P * sizeof(double));
    /* no-slip boundary fluxes (momentum.c:2224-2232) */
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                if (i == 0 && c->bctype[0] == 1) VX(Ug, IX(k, j, i)) = 0;
                if (i == mx - 1 && c->bctype[1] == 1) VX(Ug, IX(k, j, i - 1)) = 0;
                if (j == 0 && c->bctype[2] == 1) VY(Ug, IX(k, j, i)) = 0;
                if (j == my - 1 && c->bctype[3] == 1) VY(Ug, IX(k, j - 1, i)) = 0;
                if (k == 0 && c->bctype[4] == 1) VZ(Ug, IX(k, j, i)) = 0;
                if (k == mz - 1 && c->bctype[5] == 1) VZ(Ug, IX(k - 1, j, i)) = 0;
            }
    memcpy(lucont, Ug, 3 * P * sizeof(double));
    rc = vo_contra2cart(c, g, nvert, lucont, ucat);
    if (!rc) rc = vo_ib_bc(c, g, nvert, lucont, ucat);
    if (rc) { free(Ug); return rc; }
    const double dt = c->dt;
    for (size_t n = 0; n < 3 * P; n++) {
        double r = 0;
        r += -1. / dt * Ug[n];
        r += 1. / dt * U_o[n];
        rhs[n] = r;
    }
    rc = vo_formfunction2(c, g, nvert, lucont, ucat, nu_t, rhs, 0.5);
    for (size_t n = 0; n < 3 * P; n++) {
        rhs[n] += 0.5 * rhs_o[n];
        rhs[n] += -1 * dP[n];
        if (f_eul) rhs[n] += 1. * f_eul[n];
    }
    free(Ug);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Actuator models: the element search of Pre_process (rotor_model.c:       */
/* 2540-2740) and the Eulerian force spreading Calc_F_eul (4792-5137) with  */
/* its discrete delta functions (6042-6200)                                 */
/* ------------------------------------------------------------------------ */
#define VO_PI 3.14159265   /* the truncated pi of dfunc_s4h (rotor_model.c:6104) */

static double dfunc_2h(double r) { return fabs(r) < 1.0 ? 1.0 - fabs(r) : 0.0; }   /* 6085-6095 */

static double dfunc_s4h(double r)   /* 6097-6115 */
{
    if (fabs(r) <= 0.5) return 3.0 / 8.0 + VO_PI / 32.0 - pow(r, 2) / 4.0;
    if (fabs(r) >= 0.5 && fabs(r) <= 1.5)
        return 1.0 / 4.0 + (1.0 - fabs(r)) * sqrt(-2.0 + 8.0 * fabs(r) - 4.0 * pow(r, 2)) / 8.0 -
               asin(sqrt(2.0) * (fabs(r) - 1.0)) / 8.0;
    if (fabs(r) >= 1.5 && fabs(r) <= 2.5)
        return 17.0 / 16.0 - VO_PI / 64.0 - 3.0 * fabs(r) / 4.0 + pow(r, 2) / 8.0 +
               (fabs(r) - 2.0) * sqrt(-14.0 + 16.0 * fabs(r) - 4.0 * pow(r, 2)) / 16.0 +
               asin(sqrt(2.0) * (fabs(r) - 2.0)) / 16.0;
    return 0.0;
}

static double dfunc_sc4h(double r)   /* 6117-6133 (M_PI) */
{
    const double pi = 3.14159265358979323846;
    if (fabs(r) <= 1.5)
        return 0.25 * (pi + 2 * sin(0.25 * pi * (2 * fabs(r) + 1)) - 2 * sin(0.25 * pi * (2 * fabs(r) - 1))) / pi;
    if (fabs(r) >= 1.5 && fabs(r) <= 2.5)
        return -0.125 * (-5 * pi + 2 * pi * fabs(r) + 4 * sin(0.25 * pi * (2 * fabs(r) - 1))) / pi;
    return 0.0;
}

static double dfunc_nh(double r, double n) { return fabs(r) < n ? (n - fabs(r)) / pow(n, 2) : 0.0; }   /* 6183-6191 */

static double dfunc_exp(double r, double n)   /* 6193-6197 */
{
    return exp(-(r / n) * (r / n)) / (pow(n, 1) * pow(3.1415926, 0.5));
}

double vo_dfunc(int df, double r, double n)
{
    switch (df) {
    case 0: return dfunc_2h(r);
    case 6: return dfunc_nh(r, n);
    case 7: return dfunc_exp(r, n);
    case 8: return dfunc_s4h(r);
    default: return dfunc_sc4h(r);
    }
}

/* rotor_model.c:2590-2737 for one block on one rank: the closest grid node
 * of each element centre (every 20th node first, then a +-21 window; the
 * first minimum in i, j, k loop order), the grid widths there, and the
 * element's index box, closest node +- itpwidth clamped to the interior
 * (upper bounds exclusive).  cent: x[n], y[n], z[n]; box: i_min[n], i_max[n],
 * j_min[n], j_max[n], k_min[n], k_max[n]; dh: dhx[n], dhy[n], dhz[n]. */
int vo_rotor_elements(const vo_cfg *c, const vo_geom *g, const double *coor, int n, const double *cent, int itpwidth,
                      int *box, double *dh)
{
    DIMS;
    int rc = check_cfg(c);
    if (rc) return rc;
    const int lxs = 1, lxe = mx - 1, lys = 1, lye = my - 1, lzs = 1, lze = mz - 1;
    for (int l = 0; l < n; l++) {
        const double cx = cent[l], cy = cent[n + l], cz = cent[2 * n + l];
        double dmin = 1.e6;
        int icc = lxs, jcc = lys, kcc = lzs;
        for (int i = lxs; i < lxe; i = i + 20)
            for (int j = lys; j < lye; j = j + 20)
                for (int k = lzs; k < lze; k = k + 20) {
                    const size_t id = IX(k, j, i);
                    double r1 = cx - VX(coor, id), r2 = cy - VY(coor, id), r3 = cz - VZ(coor, id);
                    double d1 = sqrt(r1 * r1 + r2 * r2 + r3 * r3);
                    if (d1 < dmin) { dmin = d1; icc = i; jcc = j; kcc = k; }
                }
        double dm = 1.e13;
        int ic = 0, jc = 0, kc = 0;
        for (int i = (icc - 21 > lxs ? icc - 21 : lxs); i < (icc + 21 < lxe ? icc + 21 : lxe); i++)
            for (int j = (jcc - 21 > lys ? jcc - 21 : lys); j < (jcc + 21 < lye ? jcc + 21 : lye); j++)
                for (int k = (kcc - 21 > lzs ? kcc - 21 : lzs); k < (kcc + 21 < lze ? kcc + 21 : lze); k++) {
                    const size_t id = IX(k, j, i);
                    double r1 = cx - VX(coor, id), r2 = cy - VY(coor, id), r3 = cz - VZ(coor, id);
                    double d1 = sqrt(r1 * r1 + r2 * r2 + r3 * r3);
                    if (d1 < dm) { dm = d1; ic = i; jc = j; kc = k; }
                }
        const size_t id = IX(kc, jc, ic);
        const double *a = g->csi + 3 * id, *b = g->eta + 3 * id, *e = g->zet + 3 * id;
        double area = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        dh[l] = 1.0 / g->aj[id] / area;
        area = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
        dh[n + l] = 1.0 / g->aj[id] / area;
        area = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        dh[2 * n + l] = 1.0 / g->aj[id] / area;
        const int cc[3] = {ic, jc, kc}, lo[3] = {lxs, lys, lzs}, hi[3] = {lxe, lye, lze};
        for (int q = 0; q < 3; q++) {
            const int c1 = cc[q] - itpwidth, c2 = cc[q] + itpwidth;
            if (c1 > hi[q] || c2 < lo[q]) {
                box[2 * q * n + l] = lo[q];
                box[(2 * q + 1) * n + l] = lo[q];
            } else {
                box[2 * q * n + l] = c1 > lo[q] ? c1 : lo[q];
                box[(2 * q + 1) * n + l] = c2 < hi[q] ? c2 : hi[q];
            }
        }
    }
    return 0;
}

/* rotor_model.c:4792-5137 Calc_F_eul for one block on one rank:
 * f_eul (zeroed first) += F_lagr dA times the delta function of each face
 * centre's offset from the element centre, in the covariant directions and
 * scaled by the widths, projected on csi / eta / zet; elements in order,
 * each over its index box.  Then f_eul = 0 where the 3x3x3 nvert sum
 * (clamped to the interior) exceeds 2.9, and the ghost-to-first-cell folds,
 * which add the (zero) ghost values.  width_mode 0: widths from the cell
 * metrics; 1: dh_fixed[3] (-forcewidthfixed); 2: dA; 3: sqrt(dA)
 * (-forcedistrwidth_surfacescale, rotor_model 5). */
int vo_calc_f_eul(const vo_cfg *c, const vo_geom *g, const double *nvert, const double *coor, int n,
                  const double *cent, const double *F, const double *dA, const int *box, int df, double halfwidth,
                  int width_mode, const double *dh_fixed, double *f_eul)
{
    DIMS;
    int rc = check_cfg(c);
    if (rc) return rc;
    const size_t P = (size_t)mx * my * mz;
    memset(f_eul, 0, 3 * P * sizeof(double));
    for (int l = 0; l < n; l++) {
        const double cx = cent[l], cy = cent[n + l], cz = cent[2 * n + l];
        const double Fx = F[l], Fy = F[n + l], Fz = F[2 * n + l];
        for (int k = box[4 * n + l]; k < box[5 * n + l]; k++)
            for (int j = box[2 * n + l]; j < box[3 * n + l]; j++)
                for (int i = box[l]; i < box[n + l]; i++) {
#define CO(kk, jj, ii, q) coor[3 * IX(kk, jj, ii) + (q)]
                    double xi[3], xj[3], xk[3];
                    for (int q = 0; q < 3; q++) {
                        xi[q] = (CO(k, j, i, q) + CO(k - 1, j, i, q) + CO(k, j - 1, i, q) + CO(k - 1, j - 1, i, q)) * 0.25;
                        xj[q] = (CO(k, j, i, q) + CO(k - 1, j, i, q) + CO(k, j, i - 1, q) + CO(k - 1, j, i - 1, q)) * 0.25;
                        xk[q] = (CO(k, j, i, q) + CO(k, j - 1, i, q) + CO(k, j, i - 1, q) + CO(k, j - 1, i - 1, q)) * 0.25;
                    }
#undef CO
                    const size_t id = IX(k, j, i);
                    const double *cs = g->csi + 3 * id, *et = g->eta + 3 * id, *ze = g->zet + 3 * id;
                    double ni[3], nj[3], nk[3];
                    calc_normal(cs, et, ze, ni, nj, nk);
                    double dhx, dhy, dhz;
                    if (width_mode == 1) {
                        dhx = dh_fixed[0]; dhy = dh_fixed[1]; dhz = dh_fixed[2];
                    } else if (width_mode == 2 || width_mode == 3) {
                        dhx = width_mode == 3 ? sqrt(dA[l]) : dA[l];
                        dhy = dhx; dhz = dhx;
                    } else {
                        double area = sqrt(cs[0] * cs[0] + cs[1] * cs[1] + cs[2] * cs[2]);
                        dhx = 1.0 / g->aj[id] / area;
                        area = sqrt(et[0] * et[0] + et[1] * et[1] + et[2] * et[2]);
                        dhy = 1.0 / g->aj[id] / area;
                        area = sqrt(ze[0] * ze[0] + ze[1] * ze[1] + ze[2] * ze[2]);
                        dhz = 1.0 / g->aj[id] / area;
                    }
                    const double rx[3] = {xi[0] - cx, xj[0] - cx, xk[0] - cx};
                    const double ry[3] = {xi[1] - cy, xj[1] - cy, xk[1] - cy};
                    const double rz[3] = {xi[2] - cz, xj[2] - cz, xk[2] - cz};
                    const double vol_eul = 1.0 / (dhx * dhy * dhz);
                    double dfn[3];
                    for (int f = 0; f < 3; f++) {
                        const double r1 = fabs(rx[f] * ni[0] + ry[f] * ni[1] + rz[f] * ni[2]) / dhx;
                        const double r2 = fabs(rx[f] * nj[0] + ry[f] * nj[1] + rz[f] * nj[2]) / dhy;
                        const double r3 = fabs(rx[f] * nk[0] + ry[f] * nk[1] + rz[f] * nk[2]) / dhz;
                        dfn[f] = vol_eul * vo_dfunc(df, r1, halfwidth) * vo_dfunc(df, r2, halfwidth) *
                                 vo_dfunc(df, r3, halfwidth);
                    }
                    const double *m[3] = {cs, et, ze};
                    for (int f = 0; f < 3; f++)
                        f_eul[3 * id + f] += Fx * dfn[f] * dA[l] * m[f][0] + Fy * dfn[f] * dA[l] * m[f][1] +
                                             Fz * dfn[f] * dA[l] * m[f][2];
                }
    }
    /* 5006-5020: no force where the 3x3x3 box (within the interior) is
     * mostly solid */
    double *tmp = malloc(3 * P * sizeof(double));
    if (!tmp) return -9;
    memcpy(tmp, f_eul, 3 * P * sizeof(double));
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                double nv = 0.0;
                for (int kk = (k - 1 > 1 ? k - 1 : 1); kk < (k + 2 < mz - 1 ? k + 2 : mz - 1); kk++)
                    for (int jj = (j - 1 > 1 ? j - 1 : 1); jj < (j + 2 < my - 1 ? j + 2 : my - 1); jj++)
                        for (int ii = (i - 1 > 1 ? i - 1 : 1); ii < (i + 2 < mx - 1 ? i + 2 : mx - 1); ii++)
                            nv += nvert[IX(kk, jj, ii)];
                if (nv > 2.9) tmp[3 * IX(k, j, i)] = tmp[3 * IX(k, j, i) + 1] = tmp[3 * IX(k, j, i) + 2] = 0.0;
            }
    memcpy(f_eul, tmp, 3 * P * sizeof(double));
    free(tmp);
    /* 5021-5118: the ghost-to-first-cell folds, in the reference's order */
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                int a = i, b = j, cc = k;
                const int tests[6] = {i == 1, i == mx - 2, j == 1, j == my - 2, k == 1, k == mz - 2};
                for (int t = 0; t < 6; t++) {
                    if (!tests[t]) continue;
                    if (t == 0) a = 0;
                    else if (t == 1) a = mx - 1;
                    else if (t == 2) b = 0;
                    else if (t == 3) b = my - 1;
                    else if (t == 4) cc = 0;
                    else cc = mz - 1;
                    const size_t d0 = IX(k, j, i), s0 = IX(cc, b, a);
                    for (int q = 0; q < 3; q++) {
                        f_eul[3 * d0 + q] += f_eul[3 * s0 + q];
                        f_eul[3 * s0 + q] = 0.0;
                    }
                }
            }
    return 0;
}

/* momentum.c:104-195 Calc_Minimum_dt */
double vo_min_dt(const vo_cfg *c, const vo_geom *g, const double *ucont)
{
    DIMS;
    double ldt = 1.e7;
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t id = IX(k, j, i);
                double a;
                a = fabs(1. / VX(ucont, id) / g->aj[id]); ldt = a < ldt ? a : ldt;
                a = fabs(1. / VY(ucont, id) / g->aj[id]); ldt = a < ldt ? a : ldt;
                a = fabs(1. / VZ(ucont, id) / g->aj[id]); ldt = a < ldt ? a : ldt;
            }
    return ldt;
}

/* ------------------------------------------------------------------------ */
/* Pressure projection                                                       */
/* ------------------------------------------------------------------------ */
#define POISSON_THRESHOLD 0.1 /* poisson.c:11 */

int vo_setup_gid(const vo_cfg *c, const double *nvert, double *gid)
{
    DIMS;
    size_t P = (size_t)mx * my * mz;
    for (size_t n = 0; n < P; n++) gid[n] = -1;
    int cnt = 0;
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                if (i == 0 || i == mx - 1 || j == 0 || j == my - 1 || k == 0 || k == mz - 1) continue;
                if ((int)NV(k, j, i) > POISSON_THRESHOLD) continue;
                gid[IX(k, j, i)] = (double)cnt++;
            }
    per_copy_sca(c, gid);
    return cnt;
}

/* coefficient slots, poisson.c:951-969 */
enum { CP, EP, WP, NP, SP, TP, BP, NE, SE, NW, SW, TN, BN, TS, BS, TE, BE, TW, BW };

/* g_ab fields with the periodic index map, poisson.c:1096-1132 */
static double gfield(const vo_cfg *c, const vo_geom *g, int which, int k, int j, int i)
{
    DIMS;
    int a, b, cc;
    per_src(c, i, j, k, &a, &b, &cc);
    size_t s = IX(cc, b, a);
    const double *x, *y, *w;
    switch (which) {
    case 11: x = g->icsi; y = g->icsi; w = g->iaj; break;
    case 12: x = g->ieta; y = g->icsi; w = g->iaj; break;
    case 13: x = g->izet; y = g->icsi; w = g->iaj; break;
    case 21: x = g->jcsi; y = g->jeta; w = g->jaj; break;
    case 22: x = g->jeta; y = g->jeta; w = g->jaj; break;
    case 23: x = g->jzet; y = g->jeta; w = g->jaj; break;
    case 31: x = g->kcsi; y = g->kzet; w = g->kaj; break;
    case 32: x = g->keta; y = g->kzet; w = g->kaj; break;
    default: x = g->kzet; y = g->kzet; w = g->kaj; break;
    }
    return (x[3 * s] * y[3 * s] + x[3 * s + 1] * y[3 * s + 1] + x[3 * s + 2] * y[3 * s + 2]) * w[s];
}

/* PoissonLHSNew row body for one fluid cell, poisson.c:1160-1555 */
static void stencil_row(const vo_cfg *c, const vo_geom *g, const double *nvert, int i, int j, int k, double vol[19])
{
    DIMS;
    const int ip = c->i_periodic, jp = c->j_periodic, kp = c->k_periodic;
    const double th = POISSON_THRESHOLD, r = 1.0;
    for (int m = 0; m < 19; m++) vol[m] = 0.;
#define G(w, kk, jj, ii) gfield(c, g, w, kk, jj, ii)
    /* i+1 - i */
    if (NV(k, j, i + 1) < th && (i != mx - 2 || ip)) {
        vol[CP] -= G(11, k, j, i) / r;
        vol[EP] += G(11, k, j, i) / r;
        if ((j == my - 2 && !jp) || NV(k, j + 1, i) + NV(k, j + 1, i + 1) > th) {
            if (NV(k, j - 1, i) + NV(k, j - 1, i + 1) < th && j != 1) {
                vol[CP] += G(12, k, j, i) * 0.5 / r; vol[EP] += G(12, k, j, i) * 0.5 / r;
                vol[SP] -= G(12, k, j, i) * 0.5 / r; vol[SE] -= G(12, k, j, i) * 0.5 / r;
            }
        } else if ((j == 1 && !jp) || NV(k, j - 1, i) + NV(k, j - 1, i + 1) > th) {
            if (NV(k, j + 1, i) + NV(k, j + 1, i + 1) < th) {
                vol[NP] += G(12, k, j, i) * 0.5 / r; vol[NE] += G(12, k, j, i) * 0.5 / r;
                vol[CP] -= G(12, k, j, i) * 0.5 / r; vol[EP] -= G(12, k, j, i) * 0.5 / r;
            }
        } else {
            vol[NP] += G(12, k, j, i) * 0.25 / r; vol[NE] += G(12, k, j, i) * 0.25 / r;
            vol[SP] -= G(12, k, j, i) * 0.25 / r; vol[SE] -= G(12, k, j, i) * 0.25 / r;
        }
        if ((k == mz - 2 && !kp) || NV(k + 1, j, i) + NV(k + 1, j, i + 1) > th) {
            if (NV(k - 1, j, i) + NV(k - 1, j, i + 1) < th && k != 1) {
                vol[CP] += G(13, k, j, i) * 0.5 / r; vol[EP] += G(13, k, j, i) * 0.5 / r;
                vol[BP] -= G(13, k, j, i) * 0.5 / r; vol[BE] -= G(13, k, j, i) * 0.5 / r;
            }
        } else if ((k == 1 && !kp) || NV(k - 1, j, i) + NV(k - 1, j, i + 1) > th) {
            if (NV(k + 1, j, i) + NV(k + 1, j, i + 1) < th) {
                vol[TP] += G(13, k, j, i) * 0.5 / r; vol[TE] += G(13, k, j, i) * 0.5 / r;
                vol[CP] -= G(13, k, j, i) * 0.5 / r; vol[EP] -= G(13, k, j, i) * 0.5 / r;
            }
        } else {
            vol[TP] += G(13, k, j, i) * 0.25 / r; vol[TE] += G(13, k, j, i) * 0.25 / r;
            vol[BP] -= G(13, k, j, i) * 0.25 / r; vol[BE] -= G(13, k, j, i) * 0.25 / r;
        }
    }
    /* i - i-1 */
    if (NV(k, j, i - 1) < th && (i != 1 || ip)) {
        vol[CP] -= G(11, k, j, i - 1) / r;
        vol[WP] += G(11, k, j, i - 1) / r;
        if ((j == my - 2 && !jp) || NV(k, j + 1, i) + NV(k, j + 1, i - 1) > th) {
            if (NV(k, j - 1, i) + NV(k, j - 1, i - 1) < th && j != 1) {
                vol[CP] -= G(12, k, j, i - 1) * 0.5 / r; vol[WP] -= G(12, k, j, i - 1) * 0.5 / r;
                vol[SP] += G(12, k, j, i - 1) * 0.5 / r; vol[SW] += G(12, k, j, i - 1) * 0.5 / r;
            }
        } else if ((j == 1 && !jp) || NV(k, j - 1, i) + NV(k, j - 1, i - 1) > th) {
            if (NV(k, j + 1, i) + NV(k, j + 1, i - 1) < th) {
                vol[NP] -= G(12, k, j, i - 1) * 0.5 / r; vol[NW] -= G(12, k, j, i - 1) * 0.5 / r;
                vol[CP] += G(12, k, j, i - 1) * 0.5 / r; vol[WP] += G(12, k, j, i - 1) * 0.5 / r;
            }
        } else {
            vol[NP] -= G(12, k, j, i - 1) * 0.25 / r; vol[NW] -= G(12, k, j, i - 1) * 0.25 / r;
            vol[SP] += G(12, k, j, i - 1) * 0.25 / r; vol[SW] += G(12, k, j, i - 1) * 0.25 / r;
        }
        if ((k == mz - 2 && !kp) || NV(k + 1, j, i) + NV(k + 1, j, i - 1) > th) {
            if (NV(k - 1, j, i) + NV(k - 1, j, i - 1) < th && k != 1) {
                vol[CP] -= G(13, k, j, i - 1) * 0.5 / r; vol[WP] -= G(13, k, j, i - 1) * 0.5 / r;
                vol[BP] += G(13, k, j, i - 1) * 0.5 / r; vol[BW] += G(13, k, j, i - 1) * 0.5 / r;
            }
        } else if ((k == 1 && !kp) || NV(k - 1, j, i) + NV(k - 1, j, i - 1) > th) {
            if (NV(k + 1, j, i) + NV(k + 1, j, i - 1) < th) {
                vol[TP] -= G(13, k, j, i - 1) * 0.5 / r; vol[TW] -= G(13, k, j, i - 1) * 0.5 / r;
                vol[CP] += G(13, k, j, i - 1) * 0.5 / r; vol[WP] += G(13, k, j, i - 1) * 0.5 / r;
            }
        } else {
            vol[TP] -= G(13, k, j, i - 1) * 0.25 / r; vol[TW] -= G(13, k, j, i - 1) * 0.25 / r;
            vol[BP] += G(13, k, j, i - 1) * 0.25 / r; vol[BW] += G(13, k, j, i - 1) * 0.25 / r;
        }
    }
    /* j+1 - j */
    if (NV(k, j + 1, i) < th && (j != my - 2 || jp)) {
        if ((i == mx - 2 && !ip) || NV(k, j, i + 1) + NV(k, j + 1, i + 1) > th) {
            if (NV(k, j, i - 1) + NV(k, j + 1, i - 1) < th && i != 1) {
                vol[CP] += G(21, k, j, i) * 0.5 / r; vol[NP] += G(21, k, j, i) * 0.5 / r;
                vol[WP] -= G(21, k, j, i) * 0.5 / r; vol[NW] -= G(21, k, j, i) * 0.5 / r;
            }
        } else if ((i == 1 && !ip) || NV(k, j, i - 1) + NV(k, j + 1, i - 1) > th) {
            if (NV(k, j, i + 1) + NV(k, j + 1, i + 1) < th) {
                vol[EP] += G(21, k, j, i) * 0.5 / r; vol[NE] += G(21, k, j, i) * 0.5 / r;
                vol[CP] -= G(21, k, j, i) * 0.5 / r; vol[NP] -= G(21, k, j, i) * 0.5 / r;
            }
        } else {
            vol[EP] += G(21, k, j, i) * 0.25 / r; vol[NE] += G(21, k, j, i) * 0.25 / r;
            vol[WP] -= G(21, k, j, i) * 0.25 / r; vol[NW] -= G(21, k, j, i) * 0.25 / r;
        }
        vol[CP] -= G(22, k, j, i) / r;
        vol[NP] += G(22, k, j, i) / r;
        if ((k == mz - 2 && !kp) || NV(k + 1, j, i) + NV(k + 1, j + 1, i) > th) {
            if (NV(k - 1, j, i) + NV(k - 1, j + 1, i) < th && k != 1) {
                vol[CP] += G(23, k, j, i) * 0.5 / r; vol[NP] += G(23, k, j, i) * 0.5 / r;
                vol[BP] -= G(23, k, j, i) * 0.5 / r; vol[BN] -= G(23, k, j, i) * 0.5 / r;
            }
        } else if ((k == 1 && !kp) || NV(k - 1, j, i) + NV(k - 1, j + 1, i) > th) {
            if (NV(k + 1, j, i) + NV(k + 1, j + 1, i) < th) {
                vol[TP] += G(23, k, j, i) * 0.5 / r; vol[TN] += G(23, k, j, i) * 0.5 / r;
                vol[CP] -= G(23, k, j, i) * 0.5 / r; vol[NP] -= G(23, k, j, i) * 0.5 / r;
            }
        } else {
            vol[TP] += G(23, k, j, i) * 0.25 / r; vol[TN] += G(23, k, j, i) * 0.25 / r;
            vol[BP] -= G(23, k, j, i) * 0.25 / r; vol[BN] -= G(23, k, j, i) * 0.25 / r;
        }
    }
    /* j - j-1 */
    if (NV(k, j - 1, i) < th && (j != 1 || jp)) {
        if ((i == mx - 2 && !ip) || NV(k, j, i + 1) + NV(k, j - 1, i + 1) > th) {
            if (NV(k, j, i - 1) + NV(k, j - 1, i - 1) < th && i != 1) {
                vol[CP] -= G(21, k, j - 1, i) * 0.5 / r; vol[SP] -= G(21, k, j - 1, i) * 0.5 / r;
                vol[WP] += G(21, k, j - 1, i) * 0.5 / r; vol[SW] += G(21, k, j - 1, i) * 0.5 / r;
            }
        } else if ((i == 1 && !ip) || NV(k, j, i - 1) + NV(k, j - 1, i - 1) > th) {
            if (NV(k, j, i + 1) + NV(k, j - 1, i + 1) < th) {
                vol[EP] -= G(21, k, j - 1, i) * 0.5 / r; vol[SE] -= G(21, k, j - 1, i) * 0.5 / r;
                vol[CP] += G(21, k, j - 1, i) * 0.5 / r; vol[SP] += G(21, k, j - 1, i) * 0.5 / r;
            }
        } else {
            vol[EP] -= G(21, k, j - 1, i) * 0.25 / r; vol[SE] -= G(21, k, j - 1, i) * 0.25 / r;
            vol[WP] += G(21, k, j - 1, i) * 0.25 / r; vol[SW] += G(21, k, j - 1, i) * 0.25 / r;
        }
        vol[CP] -= G(22, k, j - 1, i) / r;
        vol[SP] += G(22, k, j - 1, i) / r;
        if ((k == mz - 2 && !kp) || NV(k + 1, j, i) + NV(k + 1, j - 1, i) > th) {
            if (NV(k - 1, j, i) + NV(k - 1, j - 1, i) < th && k != 1) {
                vol[CP] -= G(23, k, j - 1, i) * 0.5 / r; vol[SP] -= G(23, k, j - 1, i) * 0.5 / r;
                vol[BP] += G(23, k, j - 1, i) * 0.5 / r; vol[BS] += G(23, k, j - 1, i) * 0.5 / r;
            }
        } else if ((k == 1 && !kp) || NV(k - 1, j, i) + NV(k - 1, j - 1, i) > th) {
            if (NV(k + 1, j, i) + NV(k + 1, j - 1, i) < th) {
                vol[TP] -= G(23, k, j - 1, i) * 0.5 / r; vol[TS] -= G(23, k, j - 1, i) * 0.5 / r;
                vol[CP] += G(23, k, j - 1, i) * 0.5 / r; vol[SP] += G(23, k, j - 1, i) * 0.5 / r;
            }
        } else {
            vol[TP] -= G(23, k, j - 1, i) * 0.25 / r; vol[TS] -= G(23, k, j - 1, i) * 0.25 / r;
            vol[BP] += G(23, k, j - 1, i) * 0.25 / r; vol[BS] += G(23, k, j - 1, i) * 0.25 / r;
        }
    }
    /* k+1 - k */
    if (NV(k + 1, j, i) < th && (k != mz - 2 || kp)) {
        if ((i == mx - 2 && !ip) || NV(k, j, i + 1) + NV(k + 1, j, i + 1) > th) {
            if (NV(k, j, i - 1) + NV(k + 1, j, i - 1) < th && i != 1) {
                vol[CP] += G(31, k, j, i) * 0.5 / r; vol[TP] += G(31, k, j, i) * 0.5 / r;
                vol[WP] -= G(31, k, j, i) * 0.5 / r; vol[TW] -= G(31, k, j, i) * 0.5 / r;
            }
        } else if ((i == 1 && !ip) || NV(k, j, i - 1) + NV(k + 1, j, i - 1) > th) {
            if (NV(k, j, i + 1) + NV(k + 1, j, i + 1) < th) {
                vol[EP] += G(31, k, j, i) * 0.5 / r; vol[TE] += G(31, k, j, i) * 0.5 / r;
                vol[CP] -= G(31, k, j, i) * 0.5 / r; vol[TP] -= G(31, k, j, i) * 0.5 / r;
            }
        } else {
            vol[EP] += G(31, k, j, i) * 0.25 / r; vol[TE] += G(31, k, j, i) * 0.25 / r;
            vol[WP] -= G(31, k, j, i) * 0.25 / r; vol[TW] -= G(31, k, j, i) * 0.25 / r;
        }
        if ((j == my - 2 && !jp) || NV(k, j + 1, i) + NV(k + 1, j + 1, i) > th) {
            if (NV(k, j - 1, i) + NV(k + 1, j - 1, i) < th && j != 1) {
                vol[CP] += G(32, k, j, i) * 0.5 / r; vol[TP] += G(32, k, j, i) * 0.5 / r;
                vol[SP] -= G(32, k, j, i) * 0.5 / r; vol[TS] -= G(32, k, j, i) * 0.5 / r;
            }
        } else if ((j == 1 && !jp) || NV(k, j - 1, i) + NV(k + 1, j - 1, i) > th) {
            if (NV(k, j + 1, i) + NV(k + 1, j + 1, i) < th) {
                vol[NP] += G(32, k, j, i) * 0.5 / r; vol[TN] += G(32, k, j, i) * 0.5 / r;
                vol[CP] -= G(32, k, j, i) * 0.5 / r; vol[TP] -= G(32, k, j, i) * 0.5 / r;
            }
        } else {
            vol[NP] += G(32, k, j, i) * 0.25 / r; vol[TN] += G(32, k, j, i) * 0.25 / r;
            vol[SP] -= G(32, k, j, i) * 0.25 / r; vol[TS] -= G(32, k, j, i) * 0.25 / r;
        }
        vol[CP] -= G(33, k, j, i) / r;
        vol[TP] += G(33, k, j, i) / r;
    }
    /* k - k-1 */
    if (NV(k - 1, j, i) < th && (k != 1 || kp)) {
        if ((i == mx - 2 && !ip) || NV(k, j, i + 1) + NV(k - 1, j, i + 1) > th) {
            if (NV(k, j, i - 1) + NV(k - 1, j, i - 1) < th && i != 1) {
                vol[CP] -= G(31, k - 1, j, i) * 0.5 / r; vol[BP] -= G(31, k - 1, j, i) * 0.5 / r;
                vol[WP] += G(31, k - 1, j, i) * 0.5 / r; vol[BW] += G(31, k - 1, j, i) * 0.5 / r;
            }
        } else if ((i == 1 && !ip) || NV(k, j, i - 1) + NV(k - 1, j, i - 1) > th) {
            if (NV(k, j, i + 1) + NV(k - 1, j, i + 1) < th) {
                vol[EP] -= G(31, k - 1, j, i) * 0.5 / r; vol[BE] -= G(31, k - 1, j, i) * 0.5 / r;
                vol[CP] += G(31, k - 1, j, i) * 0.5 / r; vol[BP] += G(31, k - 1, j, i) * 0.5 / r;
            }
        } else {
            vol[EP] -= G(31, k - 1, j, i) * 0.25 / r; vol[BE] -= G(31, k - 1, j, i) * 0.25 / r;
            vol[WP] += G(31, k - 1, j, i) * 0.25 / r; vol[BW] += G(31, k - 1, j, i) * 0.25 / r;
        }
        if ((j == my - 2 && !jp) || NV(k, j + 1, i) + NV(k - 1, j + 1, i) > th) {
            if (NV(k, j - 1, i) + NV(k - 1, j - 1, i) < th && j != 1) {
                vol[CP] -= G(32, k - 1, j, i) * 0.5 / r; vol[BP] -= G(32, k - 1, j, i) * 0.5 / r;
                vol[SP] += G(32, k - 1, j, i) * 0.5 / r; vol[BS] += G(32, k - 1, j, i) * 0.5 / r;
            }
        } else if ((j == 1 && !jp) || NV(k, j - 1, i) + NV(k - 1, j - 1, i) > th) {
            if (NV(k, j + 1, i) + NV(k - 1, j + 1, i) < th) {
                vol[NP] -= G(32, k - 1, j, i) * 0.5 / r; vol[BN] -= G(32, k - 1, j, i) * 0.5 / r;
                vol[CP] += G(32, k - 1, j, i) * 0.5 / r; vol[BP] += G(32, k - 1, j, i) * 0.5 / r;
            }
        } else {
            vol[NP] -= G(32, k - 1, j, i) * 0.25 / r; vol[BN] -= G(32, k - 1, j, i) * 0.25 / r;
            vol[SP] += G(32, k - 1, j, i) * 0.25 / r; vol[BS] += G(32, k - 1, j, i) * 0.25 / r;
        }
        vol[CP] -= G(33, k - 1, j, i) / r;
        vol[BP] += G(33, k - 1, j, i) / r;
    }
#undef G
}

/* neighbour offsets of the 19 slots (di, dj, dk) */
static const int SLOT[19][3] = {
    {0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
    {1, 1, 0}, {1, -1, 0}, {-1, 1, 0}, {-1, -1, 0}, {0, 1, 1}, {0, 1, -1}, {0, -1, 1},
    {0, -1, -1}, {1, 0, 1}, {1, 0, -1}, {-1, 0, 1}, {-1, 0, -1}};

int vo_poisson_stencil(const vo_cfg *c, const vo_geom *g, const double *nvert, double *coef)
{
    DIMS;
    int rc = check_cfg(c);
    if (rc) return rc;
    size_t P = (size_t)mx * my * mz;
    memset(coef, 0, 19 * P * sizeof(double));
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t id = IX(k, j, i);
                if (NV(k, j, i) > POISSON_THRESHOLD) continue;
                stencil_row(c, g, nvert, i, j, k, coef + 19 * id);
            }
    return 0;
}

int vo_poisson_lhs(const vo_cfg *c, const vo_geom *g, const double *nvert,
                   const double *gid, int *rowptr, int *col, double *val)
{
    DIMS;
    int rc = check_cfg(c);
    if (rc) return rc;
    int nrow = 0, nnz = 0;
    rowptr[0] = 0;
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                int row = (int)gid[IX(k, j, i)];
                if (i == 0 || i == mx - 1 || j == 0 || j == my - 1 || k == 0 || k == mz - 1) continue;
                double vol[19];
                int idx[19];
                int start = nnz;
                if (NV(k, j, i) > 0.1 && row >= 0) {
                    col[nnz] = row; val[nnz] = 1.0; nnz++;
                } else if (NV(k, j, i) > POISSON_THRESHOLD) {
                    continue;
                } else {
                    stencil_row(c, g, nvert, i, j, k, vol);
                    for (int m = 0; m < 19; m++)
                        idx[m] = (int)gid[IX(k + SLOT[m][2], j + SLOT[m][1], i + SLOT[m][0])];
                    /* drop rule, poisson.c:1604-1608; INSERT_VALUES replaces duplicates */
                    for (int m = 0; m < 19; m++) {
                        if ((fabs(vol[m]) > 1.e-10 && idx[m] >= 0) || m == CP) {
                            int found = -1;
                            for (int q = start; q < nnz; q++) if (col[q] == idx[m]) found = q;
                            if (found >= 0) val[found] = vol[m];
                            else { col[nnz] = idx[m]; val[nnz] = vol[m]; nnz++; }
                        }
                    }
                }
                if (row != nrow) return -6; /* rows are emitted in gid order */
                nrow++;
                rowptr[nrow] = nnz;
            }
    return nnz;
}

int vo_poisson_rhs(const vo_cfg *c, const vo_geom *g, const double *nvert,
                   const double *ucont, double *b)
{
    DIMS;
    int rc = check_cfg(c);
    if (rc) return rc;
    const int ip = c->i_periodic, jp = c->j_periodic, kp = c->k_periodic;
    size_t P = (size_t)mx * my * mz;
    memset(b, 0, P * sizeof(double));
    const double coeff = 1.0; /* time_coeff(), poisson.c:17-25 */
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t id = IX(k, j, i);
                if (NV(k, j, i) >= 0.1) continue;
                double v = 0;
                v -= VX(ucont, id);
                if (i == 1 && ip) v += VX(ucont, IX(k, j, mx - 2));
                else v += VX(ucont, IX(k, j, i - 1));
                v -= VY(ucont, id);
                if (j == 1 && jp) v += VY(ucont, IX(k, my - 2, i));
                else v += VY(ucont, IX(k, j - 1, i));
                v -= VZ(ucont, id);
                if (k == 1 && kp) v += VZ(ucont, IX(mz - 2, j, i));
                else v += VZ(ucont, IX(k - 1, j, i));
                v *= -1.0 / c->dt * c->st * coeff;
                b[id] = v;
            }
    /* Remove_Nullspace_Scale, poisson_hypre.c:110-192 */
    double sum = 0, sum_aj = 0;
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t id = IX(k, j, i);
                if (NV(k, j, i) < 0.1) { sum += b[id]; sum_aj += g->aj[id]; }
            }
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t id = IX(k, j, i);
                if (NV(k, j, i) < 0.1) b[id] += -sum * g->aj[id] / sum_aj;
            }
    return 0;
}

/* poisson.c:3570-3910 VolumeFlux, flg=0 */
int vo_volume_flux(const vo_cfg *c, const vo_geom *g, const double *nvert,
                   double *ucont, double *flux_out, double *area_out)
{
    DIMS;
    int rc = check_cfg(c);
    if (rc) return rc;
    const double eps = 1.e-8, ibmval = 1.1;
    double flux = 0, area = 0;
#define AREA(v, id) sqrt(VX(v, id) * VX(v, id) + VY(v, id) * VY(v, id) + VZ(v, id) * VZ(v, id))
    if (!c->immersed) {
        /* poisson.c:3306-3500: every fluid / non-fluid face, then the
         * correction and, for immersed == 0, zero flux on those faces */
        for (int pass = 0; pass < 3; pass++) {
            double corr = 0;
            if (pass == 1) corr = area > 1.e-15 ? flux / area : 0;
            if (pass != 1) flux = 0, area = 0;
            for (int k = 1; k < mz - 1; k++)
                for (int j = 1; j < my - 1; j++)
                    for (int i = 1; i < mx - 1; i++) {
                        size_t id = IX(k, j, i);
                        const double *ar[3] = {g->csi, g->eta, g->zet};
                        double nb[3] = {NV(k, j, i + 1), NV(k, j + 1, i), NV(k + 1, j, i)};
                        for (int q = 0; q < 3; q++)
                            if (NV(k, j, i) < 0.1 && nb[q] > 0.1) {
                                double *u = &ucont[3 * id + q];
                                if (pass == 1) { *u -= AREA(ar[q], id) * corr; *u = 0; }
                                else { flux += *u; area += AREA(ar[q], id); }
                            }
                        for (int q = 0; q < 3; q++)
                            if (NV(k, j, i) > 0.1 && nb[q] < 0.1) {
                                double *u = &ucont[3 * id + q];
                                if (pass == 1) { *u += AREA(ar[q], id) * corr; *u = 0; }
                                else { flux -= *u; area += AREA(ar[q], id); }
                            }
                    }
        }
        if (flux_out) *flux_out = flux;
        if (area_out) *area_out = area;
        return 0;
    }
    for (int pass = 0; pass < 3; pass++) {
        double corr = 0;
        if (pass == 1) corr = area > 1.e-15 ? flux / area : 0;
        if (pass != 1) flux = 0, area = 0;
        for (int k = 1; k < mz - 1; k++)
            for (int j = 1; j < my - 1; j++)
                for (int i = 1; i < mx - 1; i++) {
                    size_t id = IX(k, j, i);
                    const double *ar[3] = {g->csi, g->eta, g->zet};
                    double nb[3] = {NV(k, j, i + 1), NV(k, j + 1, i), NV(k + 1, j, i)};
                    int inb[3] = {i < mx - 2, j < my - 2, k < mz - 2};
                    for (int q = 0; q < 3; q++) {
                        double *u = &ucont[3 * id + q];
                        double a = AREA(ar[q], id);
                        /* fluid cell, IB neighbour: outflow into the body */
                        if (NV(k, j, i) < 0.1 && nb[q] > 0.1 && nb[q] < ibmval && inb[q]) {
                            if (pass == 0) {
                                if (fabs(*u) > eps) { flux += *u; area += a; }
                                else *u = 0.;
                            } else if (pass == 1) {
                                if (fabs(*u) > eps) *u -= a * corr;
                            } else { flux += *u; area += a; }
                        }
                    }
                    for (int q = 0; q < 3; q++) {
                        double *u = &ucont[3 * id + q];
                        double a = AREA(ar[q], id);
                        if (NV(k, j, i) > 0.1 && NV(k, j, i) < ibmval && nb[q] < 0.1 && inb[q]) {
                            if (pass == 0) {
                                if (fabs(*u) > eps) { flux -= *u; area += a; }
                                else *u = 0.;
                            } else if (pass == 1) {
                                if (fabs(*u) > eps) *u += a * corr;
                            } else { flux -= *u; area += a; }
                        }
                    }
                }
    }
#undef AREA
    if (flux_out) *flux_out = flux;
    if (area_out) *area_out = area;
    return 0;
}

int vo_csr_apply(const vo_cfg *c, const double *gid, const int *rowptr,
                 const int *col, const double *val, const double *x, double *y)
{
    DIMS;
    size_t P = (size_t)mx * my * mz;
    int n = 0;
    for (size_t q = 0; q < P; q++) if (gid[q] >= 0) n = (int)gid[q] + 1 > n ? (int)gid[q] + 1 : n;
    double *xc = (double *)malloc((size_t)(n + 1) * sizeof(double));
    if (!xc) return -1;
    /* compress (interior cells only; periodic ghosts share the gid) */
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                int r = (int)gid[IX(k, j, i)];
                if (r >= 0) xc[r] = x[IX(k, j, i)];
            }
    memset(y, 0, P * sizeof(double));
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                int r = (int)gid[IX(k, j, i)];
                if (r < 0) continue;
                double s = 0;
                for (int q = rowptr[r]; q < rowptr[r + 1]; q++) s += val[q] * xc[col[q]];
                y[IX(k, j, i)] = s;
            }
    free(xc);
    return 0;
}

/* GMRES(m) with Jacobi right preconditioning, modified Gram-Schmidt */
static double dot(const double *a, const double *b, int n)
{
    double s = 0;
    for (int q = 0; q < n; q++) s += a[q] * b[q];
    return s;
}

static void spmv(int n, const int *rp, const int *col, const double *val, const double *x, double *y)
{
    for (int r = 0; r < n; r++) {
        double s = 0;
        for (int q = rp[r]; q < rp[r + 1]; q++) s += val[q] * x[col[q]];
        y[r] = s;
    }
}

int vo_poisson_solve(const vo_cfg *c, const double *nvert, const double *gid,
                     const int *rowptr, const int *col, const double *val,
                     const double *b, double rtol, int maxit, double *phi, double *relres)
{
    DIMS;
    size_t P = (size_t)mx * my * mz;
    int n = 0;
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) if (gid[IX(k, j, i)] >= 0) n++;
    const int M = 80;
    double *bc = calloc(n, sizeof(double)), *x = calloc(n, sizeof(double)), *dinv = calloc(n, sizeof(double));
    double *V = calloc((size_t)(M + 1) * n, sizeof(double)), *w = calloc(n, sizeof(double)), *z = calloc(n, sizeof(double));
    double *H = calloc((size_t)(M + 1) * M, sizeof(double)), *cs = calloc(M, sizeof(double));
    double *sn = calloc(M, sizeof(double)), *s = calloc(M + 1, sizeof(double)), *yv = calloc(M, sizeof(double));
    if (!bc || !x || !dinv || !V || !w || !z || !H || !cs || !sn || !s || !yv) return -1;
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                int r = (int)gid[IX(k, j, i)];
                if (r >= 0) bc[r] = b[IX(k, j, i)];
            }
    for (int r = 0; r < n; r++) {
        double d = 1;
        for (int q = rowptr[r]; q < rowptr[r + 1]; q++) if (col[q] == r) d = val[q];
        dinv[r] = d != 0 ? 1. / d : 1.;
    }
    double bn = sqrt(dot(bc, bc, n));
    if (bn == 0) bn = 1;
    int it = 0;
    double res = 0;
    while (it < maxit) {
        spmv(n, rowptr, col, val, x, w);
        for (int r = 0; r < n; r++) V[r] = bc[r] - w[r];
        double beta = sqrt(dot(V, V, n));
        res = beta / bn;
        if (res <= rtol) break;
        for (int r = 0; r < n; r++) V[r] /= beta;
        for (int q = 0; q <= M; q++) s[q] = 0;
        s[0] = beta;
        int kk;
        for (kk = 0; kk < M && it < maxit; kk++, it++) {
            for (int r = 0; r < n; r++) z[r] = dinv[r] * V[(size_t)kk * n + r];
            spmv(n, rowptr, col, val, z, w);
            for (int q = 0; q <= kk; q++) {
                double h = dot(w, V + (size_t)q * n, n);
                H[q * M + kk] = h;
                for (int r = 0; r < n; r++) w[r] -= h * V[(size_t)q * n + r];
            }
            double hn = sqrt(dot(w, w, n));
            H[(kk + 1) * M + kk] = hn;
            if (hn > 0) for (int r = 0; r < n; r++) V[(size_t)(kk + 1) * n + r] = w[r] / hn;
            for (int q = 0; q < kk; q++) {
                double t = cs[q] * H[q * M + kk] + sn[q] * H[(q + 1) * M + kk];
                H[(q + 1) * M + kk] = -sn[q] * H[q * M + kk] + cs[q] * H[(q + 1) * M + kk];
                H[q * M + kk] = t;
            }
            double a = H[kk * M + kk], bb = H[(kk + 1) * M + kk];
            double den = sqrt(a * a + bb * bb);
            cs[kk] = den > 0 ? a / den : 1; sn[kk] = den > 0 ? bb / den : 0;
            H[kk * M + kk] = den; H[(kk + 1) * M + kk] = 0;
            s[kk + 1] = -sn[kk] * s[kk]; s[kk] = cs[kk] * s[kk];
            if (fabs(s[kk + 1]) / bn <= rtol * 0.5) { kk++; it++; break; }
        }
        for (int q = kk - 1; q >= 0; q--) {
            double t = s[q];
            for (int p = q + 1; p < kk; p++) t -= H[q * M + p] * yv[p];
            yv[q] = t / H[q * M + q];
        }
        for (int r = 0; r < n; r++) {
            double t = 0;
            for (int q = 0; q < kk; q++) t += V[(size_t)q * n + r] * yv[q];
            x[r] += dinv[r] * t;
        }
    }
    spmv(n, rowptr, col, val, x, w);
    double rn = 0;
    for (int r = 0; r < n; r++) rn += (bc[r] - w[r]) * (bc[r] - w[r]);
    res = sqrt(rn) / bn;
    if (relres) *relres = res;
    /* Remove_Nullspace: plain mean over nvert<0.1 cells, poisson_hypre.c:194-255 */
    double sum = 0;
    int cnt = 0;
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++)
                if (NV(k, j, i) < 0.1) { sum += x[(int)gid[IX(k, j, i)]]; cnt++; }
    double v = -sum / (double)cnt;
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++)
                if (NV(k, j, i) < 0.1) x[(int)gid[IX(k, j, i)]] += v;
    /* Convert_Phi2_Phi, poisson.c:92-180 */
    memset(phi, 0, P * sizeof(double));
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                if ((int)NV(k, j, i) > POISSON_THRESHOLD) phi[IX(k, j, i)] = 0;
                else phi[IX(k, j, i)] = x[(int)gid[IX(k, j, i)]];
            }
    per_copy_sca(c, phi);
    free(bc); free(x); free(dinv); free(V); free(w); free(z); free(H); free(cs); free(sn); free(s); free(yv);
    return it;
}

/* poisson.c:3144-3304 UpdatePressure */
int vo_update_pressure(const vo_cfg *c, const double *nvert, double *p, double *phi)
{
    DIMS;
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t id = IX(k, j, i);
                if (nvert[id] < 0.1) p[id] += phi[id];
                if (nvert[id] > 1.1) p[id] = 0;
            }
    per_copy_sca(c, p);
    per_copy_sca(c, phi);
    return 0;
}

/* poisson.c:2707-3073 Projection */
int vo_projection(const vo_cfg *c, const vo_geom *g, const double *nvert,
                  const double *phi, double *ucont)
{
    DIMS;
    int rc = check_cfg(c);
    if (rc) return rc;
    const int ip = c->i_periodic, jp = c->j_periodic, kp = c->k_periodic;
    const double th = POISSON_THRESHOLD, coeff = 1.0, r1 = 1.0, r2 = 1.0, r3 = 1.0;
#define PH(k, j, i) phi[IX(k, j, i)]
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t id = IX(k, j, i);
                double dpdc, dpde, dpdz;
                if (i < mx - 2 || (ip && i == mx - 2)) {
                    dpdc = PH(k, j, i + 1) - PH(k, j, i);
                    if (i == mx - 2 && ip) dpdc = PH(k, j, 1) - PH(k, j, i);
                    dpde = 0.; dpdz = 0.;
                    if ((j == my - 2 && !jp) || NV(k, j + 1, i) + NV(k, j + 1, i + 1) > th) {
                        if (NV(k, j - 1, i) + NV(k, j - 1, i + 1) < th && (j != 1))
                            dpde = (PH(k, j, i) + PH(k, j, i + 1) - PH(k, j - 1, i) - PH(k, j - 1, i + 1)) * 0.5;
                    } else if ((j == 1 && !jp) || NV(k, j - 1, i) + NV(k, j - 1, i + 1) > th) {
                        if (NV(k, j + 1, i) + NV(k, j + 1, i + 1) < th)
                            dpde = (PH(k, j + 1, i) + PH(k, j + 1, i + 1) - PH(k, j, i) - PH(k, j, i + 1)) * 0.5;
                    } else dpde = (PH(k, j + 1, i) + PH(k, j + 1, i + 1) - PH(k, j - 1, i) - PH(k, j - 1, i + 1)) * 0.25;
                    if ((k == mz - 2 && !kp) || NV(k + 1, j, i) + NV(k + 1, j, i + 1) > th) {
                        if (NV(k - 1, j, i) + NV(k - 1, j, i + 1) < th && (k != 1))
                            dpdz = (PH(k, j, i) + PH(k, j, i + 1) - PH(k - 1, j, i) - PH(k - 1, j, i + 1)) * 0.5;
                    } else if ((k == 1 && !kp) || NV(k - 1, j, i) + NV(k - 1, j, i + 1) > th) {
                        if (NV(k + 1, j, i) + NV(k + 1, j, i + 1) < th)
                            dpdz = (PH(k + 1, j, i) + PH(k + 1, j, i + 1) - PH(k, j, i) - PH(k, j, i + 1)) * 0.5;
                    } else dpdz = (PH(k + 1, j, i) + PH(k + 1, j, i + 1) - PH(k - 1, j, i) - PH(k - 1, j, i + 1)) * 0.25;
                    if ((NV(k, j, i) + NV(k, j, i + 1)) < th) {
                        const double *ic = g->icsi + 3 * id, *ie = g->ieta + 3 * id, *iz = g->izet + 3 * id;
                        double a = g->iaj[id];
                        ucont[3 * id] -= (dpdc * (ic[0] * ic[0] + ic[1] * ic[1] + ic[2] * ic[2]) * a +
                                          dpde * (ie[0] * ic[0] + ie[1] * ic[1] + ie[2] * ic[2]) * a +
                                          dpdz * (iz[0] * ic[0] + iz[1] * ic[1] + iz[2] * ic[2]) * a) *
                                         c->dt * c->st / coeff / r1;
                    }
                }
                if (j < my - 2 || (jp && j == my - 2)) {
                    dpdc = 0.; dpdz = 0.;
                    if ((i == mx - 2 && !ip) || NV(k, j, i + 1) + NV(k, j + 1, i + 1) > th) {
                        if (NV(k, j, i - 1) + NV(k, j + 1, i - 1) < th && (i != 1))
                            dpdc = (PH(k, j, i) + PH(k, j + 1, i) - PH(k, j, i - 1) - PH(k, j + 1, i - 1)) * 0.5;
                    } else if ((i == 1 && !ip) || NV(k, j, i - 1) + NV(k, j + 1, i - 1) > th) {
                        if (NV(k, j, i + 1) + NV(k, j + 1, i + 1) < th)
                            dpdc = (PH(k, j, i + 1) + PH(k, j + 1, i + 1) - PH(k, j, i) - PH(k, j + 1, i)) * 0.5;
                    } else dpdc = (PH(k, j, i + 1) + PH(k, j + 1, i + 1) - PH(k, j, i - 1) - PH(k, j + 1, i - 1)) * 0.25;
                    if (j == my - 2 && jp) dpde = PH(k, 1, i) - PH(k, j, i);
                    else dpde = PH(k, j + 1, i) - PH(k, j, i);
                    if ((k == mz - 2 && !kp) || NV(k + 1, j, i) + NV(k + 1, j + 1, i) > th) {
                        if (NV(k - 1, j, i) + NV(k - 1, j + 1, i) < th && (k != 1))
                            dpdz = (PH(k, j, i) + PH(k, j + 1, i) - PH(k - 1, j, i) - PH(k - 1, j + 1, i)) * 0.5;
                    } else if ((k == 1 && !kp) || NV(k - 1, j, i) + NV(k - 1, j + 1, i) > th) {
                        if (NV(k + 1, j, i) + NV(k + 1, j + 1, i) < th)
                            dpdz = (PH(k + 1, j, i) + PH(k + 1, j + 1, i) - PH(k, j, i) - PH(k, j + 1, i)) * 0.5;
                    } else dpdz = (PH(k + 1, j, i) + PH(k + 1, j + 1, i) - PH(k - 1, j, i) - PH(k - 1, j + 1, i)) * 0.25;
                    if ((NV(k, j, i) + NV(k, j + 1, i)) < th || (int)(NV(k, j, i) + NV(k, j + 1, i)) == 5) {
                        const double *jc = g->jcsi + 3 * id, *je = g->jeta + 3 * id, *jz = g->jzet + 3 * id;
                        double a = g->jaj[id];
                        ucont[3 * id + 1] -= (dpdc * (jc[0] * je[0] + jc[1] * je[1] + jc[2] * je[2]) * a +
                                              dpde * (je[0] * je[0] + je[1] * je[1] + je[2] * je[2]) * a +
                                              dpdz * (jz[0] * je[0] + jz[1] * je[1] + jz[2] * je[2]) * a) *
                                             c->dt * c->st / coeff / r2;
                    }
                }
                if (k < mz - 2 || (kp && k == mz - 2)) {
                    dpdc = 0.; dpde = 0.;
                    if ((i == mx - 2 && !ip) || NV(k, j, i + 1) + NV(k + 1, j, i + 1) > th) {
                        if (NV(k, j, i - 1) + NV(k + 1, j, i - 1) < th && (i != 1))
                            dpdc = (PH(k, j, i) + PH(k + 1, j, i) - PH(k, j, i - 1) - PH(k + 1, j, i - 1)) * 0.5;
                    } else if ((i == 1 && !ip) || NV(k, j, i - 1) + NV(k + 1, j, i - 1) > th) {
                        if (NV(k, j, i + 1) + NV(k + 1, j, i + 1) < th)
                            dpdc = (PH(k, j, i + 1) + PH(k + 1, j, i + 1) - PH(k, j, i) - PH(k + 1, j, i)) * 0.5;
                    } else dpdc = (PH(k, j, i + 1) + PH(k + 1, j, i + 1) - PH(k, j, i - 1) - PH(k + 1, j, i - 1)) * 0.25;
                    if ((j == my - 2 && !jp) || NV(k, j + 1, i) + NV(k + 1, j + 1, i) > th) {
                        if (NV(k, j - 1, i) + NV(k + 1, j - 1, i) < th && (j != 1))
                            dpde = (PH(k, j, i) + PH(k + 1, j, i) - PH(k, j - 1, i) - PH(k + 1, j - 1, i)) * 0.5;
                    } else if ((j == 1 && !jp) || NV(k, j - 1, i) + NV(k + 1, j - 1, i) > th) {
                        if (NV(k, j + 1, i) + NV(k + 1, j + 1, i) < th)
                            dpde = (PH(k, j + 1, i) + PH(k + 1, j + 1, i) - PH(k, j, i) - PH(k + 1, j, i)) * 0.5;
                    } else dpde = (PH(k, j + 1, i) + PH(k + 1, j + 1, i) - PH(k, j - 1, i) - PH(k + 1, j - 1, i)) * 0.25;
                    if (k == mz - 2 && kp) dpdz = PH(1, j, i) - PH(k, j, i);
                    else dpdz = PH(k + 1, j, i) - PH(k, j, i);
                    if ((NV(k, j, i) + NV(k + 1, j, i)) < th) {
                        const double *kc = g->kcsi + 3 * id, *ke = g->keta + 3 * id, *kz = g->kzet + 3 * id;
                        double a = g->kaj[id];
                        ucont[3 * id + 2] -= (dpdc * (kc[0] * kz[0] + kc[1] * kz[1] + kc[2] * kz[2]) * a +
                                              dpde * (ke[0] * kz[0] + ke[1] * kz[1] + ke[2] * kz[2]) * a +
                                              dpdz * (kz[0] * kz[0] + kz[1] * kz[1] + kz[2] * kz[2]) * a) *
                                             c->dt * c->st / coeff / r3;
                    }
                }
            }
#undef PH
    /* per-component periodic face copies, poisson.c:3003-3043 */
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                int fi = 0, fj = 0, fk = 0, a = i, b = j, cc = k;
                if (ip && i == 0) a = mx - 2, fi = 1; else if (ip && i == mx - 1) a = 1, fi = 1;
                if (jp && j == 0) b = my - 2, fj = 1; else if (jp && j == my - 1) b = 1, fj = 1;
                if (kp && k == 0) cc = mz - 2, fk = 1; else if (kp && k == mz - 1) cc = 1, fk = 1;
                size_t d = IX(k, j, i), s = IX(cc, b, a);
                if (fi) VX(ucont, d) = VX(ucont, s);
                if (fj) VY(ucont, d) = VY(ucont, s);
                if (fk) VZ(ucont, d) = VZ(ucont, s);
            }
    return 0;
}

/* main.c:1047-1177 Divergence */
double vo_divergence(const vo_cfg *c, const vo_geom *g, const double *nvert,
                     const double *ucont, double *div)
{
    DIMS;
    size_t P = (size_t)mx * my * mz;
    memset(div, 0, P * sizeof(double));
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t id = IX(k, j, i);
                double m = fabs((VX(ucont, id) - VX(ucont, IX(k, j, i - 1)) + VY(ucont, id) - VY(ucont, IX(k, j - 1, i)) +
                                 VZ(ucont, id) - VZ(ucont, IX(k - 1, j, i))) * g->aj[id]);
                if (NV(k, j, i) + NV(k + 1, j, i) + NV(k - 1, j, i) + NV(k, j + 1, i) + NV(k, j - 1, i) +
                    NV(k, j, i + 1) + NV(k, j, i - 1) > 0.1) m = 0.;
                div[id] = m;
            }
    double mx_ = div[0];
    for (size_t n = 0; n < P; n++) if (div[n] > mx_) mx_ = div[n];
    return mx_;
}

/* rhs.c:1637-1659: one Runge-Kutta stage with coefficient alfa:
 * Contra2Cart; Rhs = FF2(U) - dP; U = alfa*dt*st*Rhs + U0; Contra2Cart. */
int vo_rk_stage(const vo_cfg *c, const vo_geom *g, const double *nvert,
                const double *u0, double *ucont, double *ucat, const double *nu_t,
                const double *dP, double alfa, double *rhs)
{
    DIMS;
    size_t P = (size_t)mx * my * mz;
    int rc = vo_contra2cart(c, g, nvert, ucont, ucat);
    if (rc) return rc;
    memset(rhs, 0, 3 * P * sizeof(double));
    rc = vo_formfunction2(c, g, nvert, ucont, ucat, nu_t, rhs, 1.0);
    if (rc) return rc;
    double a = alfa * c->dt * c->st;
    for (size_t n = 0; n < 3 * P; n++) {
        rhs[n] += -1 * dP[n];
        ucont[n] = a * rhs[n] + u0[n];
    }
    return vo_contra2cart(c, g, nvert, ucont, ucat);
}

/* ------------------------------------------------------------------------ */
/* CURVIB reconstruction: ibm_search_advanced's node geometry (ibm.c:366-497, */
/* BoundingSphere 1276-1364, nearestcell1 1366-1511, ICP 1659-1757,          */
/* InterceptionPoint 1759-2250; compgeom.c:56-230) and                        */
/* ibm_interpolation_advanced (ibm.c:3294-3698)                               */
/* ------------------------------------------------------------------------ */

/* ibm.c:28 Dist */
static double ib_dist(const double a[3], const double b[3])
{
    return sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
}

static void ib_vertex(const vo_ib_surface *s, int v, double p[3])
{
    p[0] = s->x_bp[v]; p[1] = s->y_bp[v]; p[2] = s->z_bp[v];
}

/* ibm.c:30-33 Cross */
static void ib_cross(double r[3], const double a[3], const double b[3])
{
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

/* ibm.c:1276-1364 BoundingSphere of element e: centre q, radius */
void vo_ib_bounding_sphere(const vo_ib_surface *s, int e, double q[3], double *rad)
{
    double p1[3], p2[3], p3[3], pa[3], pb[3], pc[3], pf[3], pu[3], pv[3], pt[3], pd[3];
    const double p0[3] = {0, 0, 0};
    ib_vertex(s, s->nv1[e], p1); ib_vertex(s, s->nv2[e], p2); ib_vertex(s, s->nv3[e], p3);
    const double l12 = ib_dist(p1, p2), l23 = ib_dist(p2, p3), l31 = ib_dist(p3, p1);
    const double *a, *b, *cc;   /* the longest edge is a-b */
    if (l12 > l23) {
        if (l12 > l31) a = p1, b = p2, cc = p3;
        else a = p3, b = p1, cc = p2;
    } else {
        if (l31 < l23) a = p2, b = p3, cc = p1;
        else a = p3, b = p1, cc = p2;
    }
    for (int q0 = 0; q0 < 3; q0++) pa[q0] = a[q0], pb[q0] = b[q0], pc[q0] = cc[q0];
    for (int q0 = 0; q0 < 3; q0++) pf[q0] = 0.5 * (pa[q0] + pb[q0]);
    for (int q0 = 0; q0 < 3; q0++) pu[q0] = pa[q0] - pf[q0], pv[q0] = pc[q0] - pf[q0];
    ib_cross(pt, pu, pv);
    ib_cross(pd, pt, pu);
    double gama = -(ib_dist(pu, p0) * ib_dist(pu, p0) - ib_dist(pv, p0) * ib_dist(pv, p0));
    for (int q0 = 0; q0 < 3; q0++) pt[q0] = pv[q0] - pu[q0];
    double lamda = 2 * (pd[0] * pt[0] + pd[1] * pt[1] + pd[2] * pt[2]);
    gama /= lamda;
    lamda = gama < 0 ? 0 : gama;
    for (int q0 = 0; q0 < 3; q0++) q[q0] = pf[q0] + lamda * pd[q0];
    *rad = ib_dist(q, pa);
}

/* compgeom.c:56-110 ISSameSide2D, as it runs: the t1/t2 of the "Change"
 * block decide, with |t1|/|p1p2| < 1e-10 counted as on the line */
static int ib_same_side(const double p[2], const double p1[2], const double p2[2], const double p3[2])
{
    const double t1 = (p[0] - p1[0]) * (p2[1] - p1[1]) - (p[1] - p1[1]) * (p2[0] - p1[0]);
    const double t2 = (p3[0] - p1[0]) * (p2[1] - p1[1]) - (p3[1] - p1[1]) * (p2[0] - p1[0]);
    const double lt = sqrt((p1[0] - p2[0]) * (p1[0] - p2[0]) + (p1[1] - p2[1]) * (p1[1] - p2[1]));
    if (fabs(t1 / lt) < 1.e-10) return 1;
    if (t1 > 0) return t2 > 0 ? 1 : -1;
    return t2 < 0 ? 1 : -1;
}

/* compgeom.c:112-158 ISInsideTriangle2D on the plane ISPointInTriangle
 * chooses from the normal: 1 inside (or on an edge), -1 outside */
static int ib_point_in_triangle(const double p[3], const double p1[3], const double p2[3], const double p3[3],
                                double nx, double ny, double nz)
{
    int u, v;
    if (fabs(nz) >= fabs(nx) && fabs(nz) >= fabs(ny)) u = 0, v = 1;
    else if (fabs(nx) >= fabs(ny) && fabs(nx) >= fabs(nz)) u = 2, v = 1;
    else u = 0, v = 2;
    const double a[2] = {p[u], p[v]}, b1[2] = {p1[u], p1[v]}, b2[2] = {p2[u], p2[v]}, b3[2] = {p3[u], p3[v]};
    int ls = ib_same_side(a, b1, b2, b3);
    if (ls < 0) return ls;
    ls = ib_same_side(a, b2, b3, b1);
    if (ls < 0) return ls;
    return ib_same_side(a, b3, b1, b2);
}

/* compgeom.c:160-186 Dis_P_Line: closest point po on segment p1p2, distance d */
static void ib_dis_p_line(const double p[3], const double p1[3], const double p2[3], double po[3], double *d)
{
    const double dx21 = p2[0] - p1[0], dy21 = p2[1] - p1[1], dz21 = p2[2] - p1[2];
    const double dx31 = p[0] - p1[0], dy31 = p[1] - p1[1], dz31 = p[2] - p1[2];
    const double t = (dx31 * dx21 + dy31 * dy21 + dz31 * dz21) / (dx21 * dx21 + dy21 * dy21 + dz21 * dz21);
    if (t < 0) {
        po[0] = p1[0], po[1] = p1[1], po[2] = p1[2];
        *d = sqrt(dx31 * dx31 + dy31 * dy31 + dz31 * dz31);
        return;
    }
    if (t > 1) po[0] = p2[0], po[1] = p2[1], po[2] = p2[2];
    else po[0] = p1[0] + t * dx21, po[1] = p1[1] + t * dy21, po[2] = p1[2] + t * dz21;
    *d = sqrt((p[0] - po[0]) * (p[0] - po[0]) + (p[1] - po[1]) * (p[1] - po[1]) + (p[2] - po[2]) * (p[2] - po[2]));
}

/* compgeom.c:188-230 triangle_intp2 / triangle_intp: barycentric weights of
 * p in the plane (u, v); the projection is chosen as nearestcell1 and ICP do
 * (x-dominant: (y, z); y-dominant: (x, z); z-dominant: (y, x)) */
static void ib_triangle_intp(const double p[3], const double p1[3], const double p2[3], const double p3[3],
                             double nx, double ny, double nz, double w[3])
{
    int u, v;
    if (fabs(nx) >= fabs(ny) && fabs(nx) >= fabs(nz)) u = 1, v = 2;
    else if (fabs(ny) >= fabs(nx) && fabs(ny) >= fabs(nz)) u = 0, v = 2;
    else if (fabs(nz) >= fabs(ny) && fabs(nz) >= fabs(nx)) u = 1, v = 0;
    else return;   /* NaN normal: the weights are left as they were */
    const double x13 = p1[u] - p3[u], y13 = p1[v] - p3[v];
    const double x23 = p2[u] - p3[u], y23 = p2[v] - p3[v];
    const double xp3 = p[u] - p3[u], yp3 = p[v] - p3[v];
    const double a = x13 * y23 - x23 * y13;
    w[0] = (y23 * xp3 - x23 * yp3) / a;
    w[1] = (-y13 * xp3 + x13 * yp3) / a;
    w[2] = 1. - w[0] - w[1];
}

/* ibm.c:1366-1511 nearestcell1: the closest surface element to p among the
 * elements whose plane has p on its positive side, scanned in element order
 * with the bounding-sphere cut.  Returns -1 if no element qualifies. */
static int ib_nearest(const vo_ib_surface *s, const double *qv, const double *rad, const double p[3],
                      vo_ib_node *nd)
{
    double dmin = 1.e20, pmin[3] = {0, 0, 0};
    int cell_min = -100;
    for (int e = 0; e < s->n_elmt; e++) {
        const double q[3] = {qv[3 * e], qv[3 * e + 1], qv[3 * e + 2]};
        const double d_center = ib_dist(p, q);
        if (!(d_center - rad[e] < dmin)) continue;
        const int n1 = s->nv1[e];
        const double nfx = s->nf_x[e], nfy = s->nf_y[e], nfz = s->nf_z[e];
        double p1[3], p2[3], p3[3], po[3], d;
        ib_vertex(s, n1, p1); ib_vertex(s, s->nv2[e], p2); ib_vertex(s, s->nv3[e], p3);
        double tf = ((p[0] - s->x_bp[n1]) * nfx + (p[1] - s->y_bp[n1]) * nfy + (p[2] - s->z_bp[n1]) * nfz);
        if (fabs(tf) < 1.e-10) tf = 1.e-10;
        if (!(tf >= 0)) continue;
        const double pj[3] = {p[0] - tf * nfx, p[1] - tf * nfy, p[2] - tf * nfz};
        if (ib_point_in_triangle(pj, p1, p2, p3, nfx, nfy, nfz) == 1) {
            if (tf < dmin) dmin = tf, pmin[0] = pj[0], pmin[1] = pj[1], pmin[2] = pj[2], cell_min = e;
        } else {
            ib_dis_p_line(p, p1, p2, po, &d);
            if (d < dmin) dmin = d, pmin[0] = po[0], pmin[1] = po[1], pmin[2] = po[2], cell_min = e;
            ib_dis_p_line(p, p2, p3, po, &d);
            if (d < dmin) dmin = d, pmin[0] = po[0], pmin[1] = po[1], pmin[2] = po[2], cell_min = e;
            ib_dis_p_line(p, p3, p1, po, &d);
            if (d < dmin) dmin = d, pmin[0] = po[0], pmin[1] = po[1], pmin[2] = po[2], cell_min = e;
        }
    }
    if (cell_min == -100) return -1;
    nd->cell = cell_min;
    nd->pmin[0] = pmin[0], nd->pmin[1] = pmin[1], nd->pmin[2] = pmin[2];
    nd->d_s = dmin;
    double p1[3], p2[3], p3[3], w[3] = {nd->cs1, nd->cs2, nd->cs3};
    ib_vertex(s, s->nv1[cell_min], p1); ib_vertex(s, s->nv2[cell_min], p2); ib_vertex(s, s->nv3[cell_min], p3);
    ib_triangle_intp(pmin, p1, p2, p3, s->nf_x[cell_min], s->nf_y[cell_min], s->nf_z[cell_min], w);
    nd->cs1 = w[0], nd->cs2 = w[1], nd->cs3 = w[2];
    return 0;
}

/* ibm.c:1670-1677: the eight triangles of a 3x3 patch of cell centres; the
 * interception point's three cells are the vertices of the triangle hit
 * (ibm.c:1800-1847 and the five like switches) */
static const int IB_TRI[8][3] = {{0, 1, 4}, {1, 2, 4}, {2, 4, 5}, {4, 5, 8}, {4, 7, 8}, {4, 6, 7}, {3, 4, 6}, {3, 4, 0}};

/* patch point n (0..8) of candidate plane pl (0..5: i-1, i+1, j-1, j+1,
 * k-1, k+1) around cell (i, j, k), ibm.c:1785-1797 and the like blocks */
void vo_ib_patch_point(int pl, int n, int i, int j, int k, int *ii, int *jj, int *kk)
{
    const int a = n % 3 - 1, b = n / 3 - 1;   /* fastest, slowest offsets of the patch */
    const int s = (pl & 1) ? 1 : -1;
    if (pl < 2) *ii = i + s, *jj = j + a, *kk = k + b;
    else if (pl < 4) *ii = i + a, *jj = j + s, *kk = k + b;
    else *ii = i + a, *jj = j + b, *kk = k + s;
}

/* ibm.c:1659-1757 ICP on one patch: the first of the eight triangles that
 * the ray p + d n (d > 0) hits */
static int ib_icp(const double p[3], const double pc[9][3], const double n[3], vo_ib_node *nd)
{
    const double nfx = n[0], nfy = n[1], nfz = n[2];
    for (int t = 0; t < 8; t++) {
        const double *p1 = pc[IB_TRI[t][0]], *p2 = pc[IB_TRI[t][1]], *p3 = pc[IB_TRI[t][2]];
        const double dx1 = p[0] - p1[0], dy1 = p[1] - p1[1], dz1 = p[2] - p1[2];
        const double dx2 = p2[0] - p1[0], dy2 = p2[1] - p1[1], dz2 = p2[2] - p1[2];
        const double dx3 = p3[0] - p1[0], dy3 = p3[1] - p1[1], dz3 = p3[2] - p1[2];
        double d = (nfx * (dy2 * dz3 - dz2 * dy3) - nfy * (dx2 * dz3 - dz2 * dx3) + nfz * (dx2 * dy3 - dy2 * dx3));
        if (!(fabs(d) > 1.e-10)) continue;
        d = -(dx1 * (dy2 * dz3 - dz2 * dy3) - dy1 * (dx2 * dz3 - dz2 * dx3) + dz1 * (dx2 * dy3 - dy2 * dx3)) / d;
        if (!(d > 0)) continue;
        const double pint[3] = {p[0] + d * nfx, p[1] + d * nfy, p[2] + d * nfz};
        const double rx1 = p2[0] - p1[0], ry1 = p2[1] - p1[1], rz1 = p2[2] - p1[2];
        const double rx2 = p3[0] - p1[0], ry2 = p3[1] - p1[1], rz2 = p3[2] - p1[2];
        const double nfxt = ry1 * rz2 - rz1 * ry2, nfyt = -rx1 * rz2 + rz1 * rx2, nfzt = rx1 * ry2 - ry1 * rx2;
        if (ib_point_in_triangle(pint, p1, p2, p3, nfxt, nfyt, nfzt) < 0) continue;
        double w[3] = {nd->cr1, nd->cr2, nd->cr3};
        ib_triangle_intp(pint, p1, p2, p3, nfxt, nfyt, nfzt, w);
        nd->cr1 = w[0], nd->cr2 = w[1], nd->cr3 = w[2];
        nd->d_i = sqrt((pint[0] - p[0]) * (pint[0] - p[0]) + (pint[1] - p[1]) * (pint[1] - p[1]) +
                       (pint[2] - p[2]) * (pint[2] - p[2]));
        return t;
    }
    return -100;
}

/* ibm.c:1759-2250 InterceptionPoint: the six patches in order; then the
 * boundary fallback of ibm_search_advanced (ibm.c:413-493).  Returns -1 if
 * the node has neither (the reference prints "IBM Searching Fail!" and keeps
 * the previous node's values). */
static int ib_intercept(const vo_cfg *c, const vo_geom *g, const vo_ib_surface *s, vo_ib_node *nd)
{
    DIMS;
    const int i = nd->i, j = nd->j, k = nd->k;
    const double *p = &g->cent[3 * IX(k, j, i)];
    double n[3] = {p[0] - nd->pmin[0], p[1] - nd->pmin[1], p[2] - nd->pmin[2]};
    const double dr = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    n[0] /= dr, n[1] /= dr, n[2] /= dr;
    nd->imode = -100;
    for (int pl = 0; pl < 6; pl++) {
        int ip[9], jp[9], kp[9];
        double pc[9][3];
        for (int q = 0; q < 9; q++) {
            vo_ib_patch_point(pl, q, i, j, k, &ip[q], &jp[q], &kp[q]);
            const double *x = &g->cent[3 * IX(kp[q], jp[q], ip[q])];
            pc[q][0] = x[0], pc[q][1] = x[1], pc[q][2] = x[2];
        }
        const int t = ib_icp(p, pc, n, nd);
        nd->imode = t;
        if (t >= 0) {
            nd->i1 = ip[IB_TRI[t][0]], nd->j1 = jp[IB_TRI[t][0]], nd->k1 = kp[IB_TRI[t][0]];
            nd->i2 = ip[IB_TRI[t][1]], nd->j2 = jp[IB_TRI[t][1]], nd->k2 = kp[IB_TRI[t][1]];
            nd->i3 = ip[IB_TRI[t][2]], nd->j3 = jp[IB_TRI[t][2]], nd->k3 = kp[IB_TRI[t][2]];
            return 0;
        }
    }
    int ti, tj, tk;
    if (i == 1 || i == mx - 2 || j == 1 || j == my - 2) ti = i, tj = j, tk = s->nf_z[nd->cell] > 0 ? k + 1 : k - 1;
    else if (k == 1 || k == mz - 2) ti = i, tj = j + 1, tk = k;
    else return -1;
    const double *pt = &g->cent[3 * IX(tk, tj, ti)];
    nd->d_i = sqrt((p[0] - pt[0]) * (p[0] - pt[0]) + (p[1] - pt[1]) * (p[1] - pt[1]) + (p[2] - pt[2]) * (p[2] - pt[2]));
    nd->cr1 = 1., nd->cr2 = 0., nd->cr3 = 0.;
    nd->i1 = nd->i2 = nd->i3 = ti;
    nd->j1 = nd->j2 = nd->j3 = tj;
    nd->k1 = nd->k2 = nd->k3 = tk;
    return 0;
}

/* the body's IB nodes: interior cells with (int)(nvert+0.5) == 1, except
 * the planes of wall-function walls, which ibm.c:518-533 sets to 1 only
 * after the list is made */
static int ib_is_node(const vo_cfg *c, const double *nvert, int i, int j, int k)
{
    DIMS;
    if ((int)(nvert[IX(k, j, i)] + 0.5) != 1) return 0;
    const int *bt = c->bctype;
    if (((bt[0] == -1 || bt[0] == -2) && i == 1) || ((bt[1] == -1 || bt[1] == -2) && i == mx - 2)) return 0;
    if (((bt[2] == -1 || bt[2] == -2) && j == 1) || ((bt[3] == -1 || bt[3] == -2) && j == my - 2)) return 0;
    if (((bt[4] == -1 || bt[4] == -2) && k == 1) || ((bt[5] == -1 || bt[5] == -2) && k == mz - 2)) return 0;
    return 1;
}

int vo_ib_search(const vo_cfg *c, const vo_geom *g, const double *nvert, const vo_ib_surface *s, vo_ib_node *nodes,
                 int cap, int *n_nodes)
{
    DIMS;
    int rc = check_cfg(c);
    if (rc) return rc;
    if (s->n_elmt <= 0 || s->n_v <= 0) return -2;
    double *qv = malloc(3 * sizeof(double) * s->n_elmt), *rad = malloc(sizeof(double) * s->n_elmt);
    for (int e = 0; e < s->n_elmt; e++) vo_ib_bounding_sphere(s, e, &qv[3 * e], &rad[e]);
    int n = 0;
    rc = 0;
    for (int k = 1; k < mz - 1 && !rc; k++)
        for (int j = 1; j < my - 1 && !rc; j++)
            for (int i = 1; i < mx - 1 && !rc; i++) {
                if (!ib_is_node(c, nvert, i, j, k)) continue;
                if (n >= cap) { rc = -4; break; }
                vo_ib_node *nd = &nodes[n++];
                memset(nd, 0, sizeof *nd);
                nd->i = i, nd->j = j, nd->k = k;
                if (ib_nearest(s, qv, rad, &g->cent[3 * IX(k, j, i)], nd) || ib_intercept(c, g, s, nd)) rc = -3;
            }
    free(qv);
    free(rad);
    if (rc) return rc;
    /* AddIBMNode (variables.c:43-51) prepends: the list runs from the last
     * node found to the first */
    for (int a = 0, b = n - 1; a < b; a++, b--) {
        vo_ib_node t = nodes[a];
        nodes[a] = nodes[b];
        nodes[b] = t;
    }
    *n_nodes = n;
    return 0;
}

/* sediment_transport.c:18-68: the rough-wall law of the IB wall model */
static double hy_E(double utau, double ks, double nu)
{
    const double kplus = utau * ks / nu, kappa = 0.41, B = 5.5;
    double dB;
    if (kplus <= 2.25) dB = 0.0;
    else if (kplus > 2.25 && kplus < 90.) dB = (B - 8.5 + 1. / kappa * log(kplus)) * sin(0.4258 * (log(kplus) - 0.811));
    else dB = B - 8.5 + 1. / kappa * log(kplus);
    return exp(kappa * (B - dB));
}

double vo_hydset_u(double nu, double y, double utau, double ks)
{
    const double y0plus = 11.53, kappa = 0.41, yplus = utau * y / nu;
    if (yplus <= y0plus) return utau * yplus;
    return utau / kappa * log(hy_E(utau, ks, nu) * yplus);
}

static double hy_f(double nu, double u, double y, double utau0, double ks)
{
    const double y0plus = 11.53, kappa = 0.41, yplus = utau0 * y / nu;
    if (yplus <= y0plus) return utau0 * yplus - u;
    return utau0 * (1. / kappa * log(hy_E(utau0, ks, nu) * yplus)) - u;
}

double vo_hydset_utau(double nu, double u, double y, double guess, double ks)
{
    double utau = guess, utau0 = guess;
    for (int ii = 0; ii < 30; ii++) {
        const double eps = 1.e-7;
        const double df = (hy_f(nu, u, y, utau0 + eps, ks) - hy_f(nu, u, y, utau0 - eps, ks)) / (2. * eps);
        utau = utau0 - hy_f(nu, u, y, utau0, ks) / df;
        if (fabs(utau - utau0) < 1.e-7) break;
        utau0 = utau;
    }
    return utau;
}

/* wallfunction.c:221-231 u_loglaw_roughness: linear in u_tau */
double vo_roughlog_u(double y, double utau, double ks)
{
    return utau * (1. / WF_KAPPA * log(y / ks) + 8.5);
}

/* wallfunction.c:244-246, 264-269, 423-437: Newton on rl_u - u, slope by a
 * central difference of step 1e-7, stop below 1e-10, 30 steps at most */
double vo_roughlog_utau(double u, double y, double guess, double ks)
{
    double x = guess, x0 = guess;
    for (int it = 0; it < 30; it++) {
        const double h = 1.e-7;
        const double fp = vo_roughlog_u(y, x0 + h, ks) - u, fm = vo_roughlog_u(y, x0 - h, ks) - u;
        x = x0 - (vo_roughlog_u(y, x0, ks) - u) / ((fp - fm) / (2 * h));
        if (fabs(x0 - x) < 1.e-10) break;
        x0 = x;
    }
    return x;
}

int vo_ib_interpolate(const vo_cfg *c, const vo_geom *g, const double *nvert, const vo_ib_surface *s,
                      const vo_ib_node *nodes, int n, double *ucont, double *ucat, double *p, double *ustar,
                      int *count, double *shear, double *pressure)
{
    const vo_ib_opt o = {0, 0, 0.0, 0.0};
    return vo_ib_interpolate_opt(c, g, nvert, s, nodes, n, &o, ucont, ucat, p, ustar, count, shear, pressure);
}

int vo_ib_interpolate_opt(const vo_cfg *c, const vo_geom *g, const double *nvert, const vo_ib_surface *s,
                          const vo_ib_node *nodes, int n, const vo_ib_opt *opt, double *ucont, double *ucat, double *p,
                          double *ustar, int *count, double *shear, double *pressure)
{
    DIMS;
    int rc = check_cfg(c);
    if (rc) return rc;
    const size_t P = (size_t)mx * my * mz;
    /* reads see the fields as they were before the pass (lUcat, lP) */
    double *lucat = malloc(3 * P * sizeof(double)), *lp = malloc(P * sizeof(double));
    memcpy(lucat, ucat, 3 * P * sizeof(double));
    memcpy(lp, p, P * sizeof(double));
    for (int e = 0; e < s->n_elmt; e++) count[e] = 0, shear[e] = 0, pressure[e] = 0;
    for (int l = 0; l < n; l++) {
        const vo_ib_node *nd = &nodes[l];
        const int ni = nd->cell;
        const size_t id = IX(nd->k, nd->j, nd->i);
        const size_t q1 = IX(nd->k1, nd->j1, nd->i1), q2 = IX(nd->k2, nd->j2, nd->i2), q3 = IX(nd->k3, nd->j3, nd->i3);
        const double sb = nd->d_s, sc = sb + nd->d_i;
        const double sk1 = nd->cr1, sk2 = nd->cr2, sk3 = nd->cr3;
        const double cs1 = nd->cs1, cs2 = nd->cs2, cs3 = nd->cs3;
        const double nfx = s->nf_x[ni], nfy = s->nf_y[ni], nfz = s->nf_z[ni];
        double Ua[3] = {0, 0, 0}, Uc[3], Ub[3];
        if (ni >= 0 && s->u) {
            const double *u1 = &s->u[3 * s->nv1[ni]], *u2 = &s->u[3 * s->nv2[ni]], *u3 = &s->u[3 * s->nv3[ni]];
            for (int q = 0; q < 3; q++) Ua[q] = u1[q] * cs1 + u2[q] * cs2 + u3[q] * cs3;
        }
        for (int q = 0; q < 3; q++) Uc[q] = (lucat[3 * q1 + q] * sk1 + lucat[3 * q2 + q] * sk2 + lucat[3 * q3 + q] * sk3);
        if (!opt->moving && !c->inviscid && nvert[q1] + nvert[q2] + nvert[q3] > 0.1) {
            Ub[0] = Ub[1] = Ub[2] = 0;
            ustar[id] = 0;
        } else if (nvert[id] > 2.9) {
            Ub[0] = Ub[1] = Ub[2] = 0;
            ustar[id] = 0;
        } else {
            const double u_c = Uc[0] - Ua[0], v_c = Uc[1] - Ua[1], w_c = Uc[2] - Ua[2];
            const double un = u_c * nfx + v_c * nfy + w_c * nfz;
            double ut = u_c - un * nfx, vt = v_c - un * nfy, wt = w_c - un * nfz;
            if (opt->wallfunction == 2) {   /* wall_function_roughness_loglaw, wallfunction.c:71-95 */
                const double ut_mag = sqrt(ut * ut + vt * vt + wt * wt);
                ustar[id] = vo_roughlog_utau(ut_mag, sc, 0.01, opt->roughness);
                const double um = vo_roughlog_u(sb, ustar[id], opt->roughness);
                if (ut_mag > 1.e-10) ut *= um / ut_mag, vt *= um / ut_mag, wt *= um / ut_mag;
                else ut = vt = wt = 0;
                Ub[0] = ut;   /* the normal part is dropped (wallfunction.c:88-90) */
                Ub[1] = vt;
                Ub[2] = wt;
            } else if (opt->wallfunction) {   /* wall_function_roughness_a, sediment_transport.c:70-96 */
                const double ut_mag = sqrt(ut * ut + vt * vt + wt * wt);
                ustar[id] = vo_hydset_utau(1. / c->ren, ut_mag, sc, 0.01, opt->ks);
                const double um = vo_hydset_u(1. / c->ren, sb, ustar[id], opt->ks);
                if (ut_mag > 1.e-10) ut *= um / ut_mag, vt *= um / ut_mag, wt *= um / ut_mag;
                else ut = vt = wt = 0;
                Ub[0] = ut + sb / sc * un * nfx;
                Ub[1] = vt + sb / sc * un * nfy;
                Ub[2] = wt + sb / sc * un * nfz;
            } else if (c->inviscid) {   /* freeslip, wallfunction.c:152-166 */
                Ub[0] = ut + sb / sc * un * nfx;
                Ub[1] = vt + sb / sc * un * nfy;
                Ub[2] = wt + sb / sc * un * nfz;
            } else {             /* noslip, wallfunction.c:134-150 */
                const double ut_mag = sqrt(ut * ut + vt * vt + wt * wt);
                ustar[id] = sqrt(ut_mag / sc / c->ren);
                Ub[0] = sb / sc * u_c;
                Ub[1] = sb / sc * v_c;
                Ub[2] = sb / sc * w_c;
            }
            Ub[0] += Ua[0], Ub[1] += Ua[1], Ub[2] += Ua[2];
        }
        for (int q = 0; q < 3; q++) ucat[3 * id + q] = Ub[q];
        p[id] = (lp[q1] * sk1 + lp[q2] * sk2 + lp[q3] * sk3);
        count[ni]++;
        shear[ni] += ustar[id] * ustar[id];
        pressure[ni] += p[id];
    }
    for (int e = 0; e < s->n_elmt; e++)
        if (count[e]) shear[e] /= (double)count[e], pressure[e] /= (double)count[e];
    free(lucat);
    free(lp);
    /* fluxes of the faces of IB nodes from the new velocity, ibm.c:3619-3679 */
    const int *bt = c->bctype;
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                const size_t id = IX(k, j, i), ii = IX(k, j, i + 1), jj = IX(k, j + 1, i), kk = IX(k + 1, j, i);
                const double *u0 = &ucat[3 * id];
                double ucx, ucy, ucz;
                if ((int)(nvert[id] + 0.5) == 1) {
                    ucx = (u0[0] + ucat[3 * ii]) * 0.5;
                    ucy = (u0[1] + ucat[3 * ii + 1]) * 0.5;
                    ucz = (u0[2] + ucat[3 * ii + 2]) * 0.5;
                    VX(ucont, id) = (ucx * VX(g->icsi, id) + ucy * VY(g->icsi, id) + ucz * VZ(g->icsi, id)) * 1.0;
                    ucx = (ucat[3 * jj] + u0[0]) * 0.5;
                    ucy = (ucat[3 * jj + 1] + u0[1]) * 0.5;
                    ucz = (ucat[3 * jj + 2] + u0[2]) * 0.5;
                    VY(ucont, id) = (ucx * VX(g->jeta, id) + ucy * VY(g->jeta, id) + ucz * VZ(g->jeta, id)) * 1.0;
                    ucx = (ucat[3 * kk] + u0[0]) * 0.5;
                    ucy = (ucat[3 * kk + 1] + u0[1]) * 0.5;
                    ucz = (ucat[3 * kk + 2] + u0[2]) * 0.5;
                    VZ(ucont, id) = (ucx * VX(g->kzet, id) + ucy * VY(g->kzet, id) + ucz * VZ(g->kzet, id)) * 1.0;
                    if ((bt[0] == -1 || bt[0] == -2) && i == 1) VX(ucont, id) = 0;
                    if ((bt[2] == -1 || bt[2] == -2) && j == 1) VY(ucont, id) = 0;
                    if ((bt[4] == -1 || bt[4] == -2) && k == 1) VZ(ucont, id) = 0;
                }
                if ((int)(nvert[ii] + 0.5) == 1) {
                    ucx = (u0[0] + ucat[3 * ii]) * 0.5;
                    ucy = (u0[1] + ucat[3 * ii + 1]) * 0.5;
                    ucz = (u0[2] + ucat[3 * ii + 2]) * 0.5;
                    VX(ucont, id) = (ucx * VX(g->icsi, id) + ucy * VY(g->icsi, id) + ucz * VZ(g->icsi, id)) * 1.0;
                    if ((bt[1] == -1 || bt[1] == -2) && i == mx - 3) VX(ucont, id) = 0;
                }
                if ((int)(nvert[jj] + 0.5) == 1) {
                    ucx = (ucat[3 * jj] + u0[0]) * 0.5;
                    ucy = (ucat[3 * jj + 1] + u0[1]) * 0.5;
                    ucz = (ucat[3 * jj + 2] + u0[2]) * 0.5;
                    VY(ucont, id) = (ucx * VX(g->jeta, id) + ucy * VY(g->jeta, id) + ucz * VZ(g->jeta, id)) * 1.0;
                    if ((bt[3] == -1 || bt[3] == -2) && j == my - 3) VY(ucont, id) = 0;
                }
                if ((int)(nvert[kk] + 0.5) == 1) {
                    ucx = (ucat[3 * kk] + u0[0]) * 0.5;
                    ucy = (ucat[3 * kk + 1] + u0[1]) * 0.5;
                    ucz = (ucat[3 * kk + 2] + u0[2]) * 0.5;
                    VZ(ucont, id) = (ucx * VX(g->kzet, id) + ucy * VY(g->kzet, id) + ucz * VZ(g->kzet, id)) * 1.0;
                    if ((bt[5] == -1 || bt[5] == -2) && k == mz - 3) VZ(ucont, id) = 0;
                }
            }
    return vo_contra2cart(c, g, nvert, ucont, ucat);
}

/* ------------------------------------------------------------------------ */
/* ibm_search_advanced's marking (ibm.c:110-360): cells of a triangulated     */
/* body by ray casting against a 40^3 bin grid of the elements, then the     */
/* 3x3x3 dilation and the relabelling (ibm.c:319-341, 508-516)                */
/* ------------------------------------------------------------------------ */
#define IB_NC 40

/* compgeom.c:17-51 intsect_triangle (Moller-Trumbore, EPSILON 1e-8) */
static int ib_intsect(const double o[3], const double dir[3], const double v0[3], const double v1[3],
                      const double v2[3], double *t, double *u, double *v)
{
    double e1[3], e2[3], tv[3], pv[3], qv[3];
    for (int q = 0; q < 3; q++) e1[q] = v1[q] - v0[q], e2[q] = v2[q] - v0[q];
    pv[0] = dir[1] * e2[2] - dir[2] * e2[1];
    pv[1] = dir[2] * e2[0] - dir[0] * e2[2];
    pv[2] = dir[0] * e2[1] - dir[1] * e2[0];
    const double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
    if (det > -1e-8 && det < 1e-8) return 0;
    const double inv = 1.0 / det;
    for (int q = 0; q < 3; q++) tv[q] = o[q] - v0[q];
    *u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) * inv;
    if (*u < 0.0 || *u > 1.0) return 0;
    qv[0] = tv[1] * e1[2] - tv[2] * e1[1];
    qv[1] = tv[2] * e1[0] - tv[0] * e1[2];
    qv[2] = tv[0] * e1[1] - tv[1] * e1[0];
    *v = (dir[0] * qv[0] + dir[1] * qv[1] + dir[2] * qv[2]) * inv;
    if (*v < 0.0 || *u + *v > 1.0) return 0;
    *t = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) * inv;
    return 1;
}

/* the first rand() after srand(seed), scaled to [-0.5, 0.5) (ibm.c:66-68) */
double vo_ib_ray_jitter(int seed)
{
    srand(seed);
    return rand() / ((double)RAND_MAX + 1) - 0.5;
}

typedef struct {
    double xmin, ymin, zmin, xmax, ymax, zmax, dcx, dcy, dcz;
    int *head, *next, *node, n;   /* bin lists: head[bin] -> entry, next[entry], node[entry] */
} ib_bins;

/* ibm.c:135-209: the body's box widened by 0.05 and every element listed in
 * the bins its own box touches (insertnode keeps one entry per element) */
static void ib_make_bins(const vo_ib_surface *s, ib_bins *b)
{
    b->xmin = b->ymin = b->zmin = 1.e23;
    b->xmax = b->ymax = b->zmax = -1.e23;
    for (int v = 0; v < s->n_v; v++) {
        b->xmin = fmin(b->xmin, s->x_bp[v]); b->xmax = fmax(b->xmax, s->x_bp[v]);
        b->ymin = fmin(b->ymin, s->y_bp[v]); b->ymax = fmax(b->ymax, s->y_bp[v]);
        b->zmin = fmin(b->zmin, s->z_bp[v]); b->zmax = fmax(b->zmax, s->z_bp[v]);
    }
    b->xmin -= 0.05; b->xmax += 0.05;
    b->ymin -= 0.05; b->ymax += 0.05;
    b->zmin -= 0.05; b->zmax += 0.05;
    b->dcx = (b->xmax - b->xmin) / (IB_NC - 1.);
    b->dcy = (b->ymax - b->ymin) / (IB_NC - 1.);
    b->dcz = (b->zmax - b->zmin) / (IB_NC - 1.);
    const int nb = IB_NC * IB_NC * IB_NC;
    b->head = malloc(sizeof(int) * nb);
    for (int q = 0; q < nb; q++) b->head[q] = -1;
    int cap = 1024, n = 0;
    b->next = malloc(sizeof(int) * cap);
    b->node = malloc(sizeof(int) * cap);
    for (int e = 0; e < s->n_elmt; e++) {
        const int a = s->nv1[e], bb = s->nv2[e], c = s->nv3[e];
        const double x0 = fmin(fmin(s->x_bp[a], s->x_bp[bb]), s->x_bp[c]), x1 = fmax(fmax(s->x_bp[a], s->x_bp[bb]), s->x_bp[c]);
        const double y0 = fmin(fmin(s->y_bp[a], s->y_bp[bb]), s->y_bp[c]), y1 = fmax(fmax(s->y_bp[a], s->y_bp[bb]), s->y_bp[c]);
        const double z0 = fmin(fmin(s->z_bp[a], s->z_bp[bb]), s->z_bp[c]), z1 = fmax(fmax(s->z_bp[a], s->z_bp[bb]), s->z_bp[c]);
        int i0 = (int)floor((x0 - b->xmin) / b->dcx), i1 = (int)floor((x1 - b->xmin) / b->dcx) + 1;
        int j0 = (int)floor((y0 - b->ymin) / b->dcy), j1 = (int)floor((y1 - b->ymin) / b->dcy) + 1;
        int k0 = (int)floor((z0 - b->zmin) / b->dcz), k1 = (int)floor((z1 - b->zmin) / b->dcz) + 1;
        i0 = i0 < 0 ? 0 : i0; i1 = i1 > IB_NC ? IB_NC : i1;
        j0 = j0 < 0 ? 0 : j0; j1 = j1 > IB_NC ? IB_NC : j1;
        k0 = k0 < 0 ? 0 : k0; k1 = k1 > IB_NC ? IB_NC : k1;
        for (int k = k0; k < k1; k++)
            for (int j = j0; j < j1; j++)
                for (int i = i0; i < i1; i++) {
                    const int bin = (k * IB_NC + j) * IB_NC + i;
                    int dup = 0;
                    for (int q = b->head[bin]; q >= 0 && !dup; q = b->next[q]) dup = b->node[q] == e;
                    if (dup) continue;
                    if (n == cap) {
                        cap *= 2;
                        b->next = realloc(b->next, sizeof(int) * cap);
                        b->node = realloc(b->node, sizeof(int) * cap);
                    }
                    b->node[n] = e; b->next[n] = b->head[bin]; b->head[bin] = n; n++;
                }
    }
    b->n = n;
}

/* ibm.c:1068-1183 point_cell_advanced: 4 if the ray from p towards a
 * jittered point above the body crosses the surface an odd number of times,
 * else 0; a ray with two crossings within 1e-8 of each other (normals not
 * opposed) is cast again with the next seed */
static int ib_point_cell(const vo_ib_surface *s, const ib_bins *b, const double p[3], int ip, int jp, int kp,
                         char *searched)
{
    for (int seed = 1; seed < 100000; seed++) {
        const double xpc = b->dcx * (ip + 0.5) + b->xmin, ypc = b->dcy * (jp + 0.5) + b->ymin;
        const double sj = vo_ib_ray_jitter(seed);
        double dir[3] = {xpc + sj * b->dcx - p[0], ypc + sj * b->dcy - p[1], b->zmax + 0.2 - p[2]};
        const double len = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        dir[0] /= len, dir[1] /= len, dir[2] /= len;
        memset(searched, 0, s->n_elmt);
        double dt[1000], dn[1000][3];
        int nintp = 0, sing = 0;
        for (int k = kp; k < IB_NC && !sing; k++)
            for (int q = b->head[(k * IB_NC + jp) * IB_NC + ip]; q >= 0 && !sing; q = b->next[q]) {
                const int e = b->node[q];
                if (searched[e]) continue;
                searched[e] = 1;
                double v0[3], v1[3], v2[3], t, u, v;
                ib_vertex(s, s->nv1[e], v0); ib_vertex(s, s->nv2[e], v1); ib_vertex(s, s->nv3[e], v2);
                if (!(ib_intsect(p, dir, v0, v1, v2, &t, &u, &v) > 0 && t > 0)) continue;
                if (nintp == 1000) return -1;
                dt[nintp] = t;
                dn[nintp][0] = s->nf_x[e], dn[nintp][1] = s->nf_y[e], dn[nintp][2] = s->nf_z[e];
                nintp++;
                for (int m = 0; m < nintp - 1; m++) {
                    const double ndotn = dn[m][0] * s->nf_x[e] + dn[m][1] * s->nf_y[e] + dn[m][2] * s->nf_z[e];
                    if (fabs(t - dt[m]) < 1e-8 && ndotn > -0.97) sing = 1;
                }
            }
        if (!sing) return nintp % 2 ? 4 : 0;
    }
    return -1;
}

/* compgeom.c:301-370 ISLineTriangleIntp: does segment p1p2 cut element e */
static int ib_line_cuts(const vo_ib_surface *s, const double p1[3], const double p2[3], int e)
{
    const int n1 = s->nv1[e], n2 = s->nv2[e], n3 = s->nv3[e];
    const double xf1 = s->x_bp[n1], yf1 = s->y_bp[n1], zf1 = s->z_bp[n1];
    const double xf2 = s->x_bp[n2], yf2 = s->y_bp[n2], zf2 = s->z_bp[n2];
    const double xf3 = s->x_bp[n3], yf3 = s->y_bp[n3], zf3 = s->z_bp[n3];
    const double nfx = s->nf_x[e], nfy = s->nf_y[e], nfz = s->nf_z[e];
    const double tf1 = (p1[0] - xf1) * nfx + (p1[1] - yf1) * nfy + (p1[2] - zf1) * nfz;
    const double tf2 = (p2[0] - xf1) * nfx + (p2[1] - yf1) * nfy + (p2[2] - zf1) * nfz;
    if (!(tf1 * tf2 < 0)) return 0;
    const double dx1 = p1[0] - xf1, dy1 = p1[1] - yf1, dz1 = p1[2] - zf1;
    const double nftx = p2[0] - p1[0], nfty = p2[1] - p1[1], nftz = p2[2] - p1[2];
    const double dx2 = xf2 - xf1, dy2 = yf2 - yf1, dz2 = zf2 - zf1;
    const double dx3 = xf3 - xf1, dy3 = yf3 - yf1, dz3 = zf3 - zf1;
    const double d = -(dx1 * (dy2 * dz3 - dz2 * dy3) - dy1 * (dx2 * dz3 - dz2 * dx3) + dz1 * (dx2 * dy3 - dy2 * dx3)) /
                     (nftx * (dy2 * dz3 - dz2 * dy3) - nfty * (dx2 * dz3 - dz2 * dx3) + nftz * (dx2 * dy3 - dy2 * dx3));
    const double pint[3] = {p1[0] + d * nftx, p1[1] + d * nfty, p1[2] + d * nftz};
    if (!(d > 0 && d < 1.)) return 0;
    const double pa[3] = {xf1, yf1, zf1}, pb[3] = {xf2, yf2, zf2}, pc[3] = {xf3, yf3, zf3};
    return ib_point_in_triangle(pint, pa, pb, pc, nfx, nfy, nfz) == 1;
}

/* ibm.c:1186-1244 point_cell_thin: does a segment from p to p1..p4 cut an
 * element listed in the bins around (ip, jp, kp) */
static int ib_point_cell_thin(const vo_ib_surface *s, const ib_bins *b, const double *p, const double *const q[4],
                              int ip, int jp, int kp, char *searched)
{
    memset(searched, 0, s->n_elmt);
    const int ks = kp - 1 > 0 ? kp - 1 : 0, js = jp - 1 > 0 ? jp - 1 : 0, is = ip - 1 > 0 ? ip - 1 : 0;
    for (int k = ks; k < kp + 2 && k < IB_NC; k++)
        for (int j = js; j < jp + 2 && j < IB_NC; j++)
            for (int i = is; i < ip + 2 && i < IB_NC; i++)
                for (int m = b->head[(k * IB_NC + j) * IB_NC + i]; m >= 0; m = b->next[m]) {
                    const int e = b->node[m];
                    if (searched[e]) continue;
                    searched[e] = 1;
                    for (int w = 0; w < 4; w++)
                        if (ib_line_cuts(s, p, q[w], e)) return 2;
                }
    return 0;
}

int vo_nvert_body(const vo_cfg *c, const vo_geom *g, double *nvert, const vo_ib_surface *s, int reset)
{
    return vo_nvert_body_thin(c, g, nvert, s, reset, 0);
}

int vo_nvert_body_thin(const vo_cfg *c, const vo_geom *g, double *nvert, const vo_ib_surface *s, int reset, int thin)
{
    DIMS;
    int rc = check_cfg(c);
    if (rc) return rc;
    if (s->n_elmt <= 0 || s->n_v <= 0) return -2;
    ib_bins b;
    ib_make_bins(s, &b);
    char *searched = malloc(s->n_elmt);
    for (int k = 1; k < mz - 1 && !rc; k++)
        for (int j = 1; j < my - 1 && !rc; j++)
            for (int i = 1; i < mx - 1; i++) {
                const size_t id = IX(k, j, i);
                const double *p = &g->cent[3 * id];
                if (reset && nvert[id] < 5.) nvert[id] = 0;
                if (!(p[0] > b.xmin && p[0] < b.xmax && p[1] > b.ymin && p[1] < b.ymax && p[2] > b.zmin && p[2] < b.zmax))
                    continue;
                const int ic = (int)floor((p[0] - b.xmin) / b.dcx), jc = (int)floor((p[1] - b.ymin) / b.dcy),
                          kc = (int)floor((p[2] - b.zmin) / b.dcz);
                const int val = ib_point_cell(s, &b, p, ic, jc, kc, searched);
                if (val < 0) { rc = -3; break; }
                nvert[id] = nvert[id] > val ? nvert[id] : val;
                if (nvert[id] < 0) nvert[id] = 0;
            }
    /* thin bodies (ibm.c:281-317): a segment from a cell centre to its +i,
     * +j, +k and +ijk neighbours that cuts an element marks the eight cells
     * of that box with 2 where they are below 0.5 */
    if (!rc && thin)
        for (int k = 1; k < mz - 1; k++)
            for (int j = 1; j < my - 1; j++)
                for (int i = 1; i < mx - 1; i++) {
                    const double *p = &g->cent[3 * IX(k, j, i)];
                    if (!(p[0] > b.xmin && p[0] < b.xmax && p[1] > b.ymin && p[1] < b.ymax && p[2] > b.zmin &&
                          p[2] < b.zmax))
                        continue;
                    int open = 0;
                    for (int dk = 0; dk < 2; dk++)
                        for (int dj = 0; dj < 2; dj++)
                            for (int di = 0; di < 2; di++) open |= nvert[IX(k + dk, j + dj, i + di)] < 0.5;
                    if (!open) continue;
                    const int ic = (int)floor((p[0] - b.xmin) / b.dcx), jc = (int)floor((p[1] - b.ymin) / b.dcy),
                              kc = (int)floor((p[2] - b.zmin) / b.dcz);
                    const double *q[4] = {&g->cent[3 * IX(k, j, i + 1)], &g->cent[3 * IX(k, j + 1, i)],
                                          &g->cent[3 * IX(k + 1, j, i)], &g->cent[3 * IX(k + 1, j + 1, i + 1)]};
                    if (!ib_point_cell_thin(s, &b, p, q, ic, jc, kc, searched)) continue;
                    for (int dk = 0; dk < 2; dk++)
                        for (int dj = 0; dj < 2; dj++)
                            for (int di = 0; di < 2; di++)
                                if (nvert[IX(k + dk, j + dj, i + di)] < 0.5) nvert[IX(k + dk, j + dj, i + di)] = 2.;
                }
    free(searched);
    free(b.head); free(b.next); free(b.node);
    if (rc) return rc;
    /* the dilation and the relabelling, as for the sphere */
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                size_t id = IX(k, j, i);
                if (nvert[id] < 0) nvert[id] = 0;
                int ip = (i < mx - 1 ? (i + 1) : (i)), im = (i > 0 ? (i - 1) : (i));
                int jp = (j < my - 1 ? (j + 1) : (j)), jm = (j > 0 ? (j - 1) : (j));
                int kp = (k < mz - 1 ? (k + 1) : (k)), km = (k > 0 ? (k - 1) : (k));
                if ((int)(nvert[id] + 0.5) != 4) {
                    for (int kk = km; kk < kp + 1; kk++)
                        for (int jj = jm; jj < jp + 1; jj++)
                            for (int ii = im; ii < ip + 1; ii++)
                                if ((int)(nvert[IX(kk, jj, ii)] + 0.5) == 4)
                                    nvert[id] = nvert[id] > 2 ? nvert[id] : 2;
                }
            }
    for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                size_t id = IX(k, j, i);
                if ((int)(nvert[id] + 0.5) == 2) nvert[id] = 1;
                if ((int)(nvert[id] + 0.5) == 4) nvert[id] = 3;
            }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Open k boundaries: Calc_Inlet_Area (bcs.c:216-328), InflowFlux at a k = 0  */
/* inlet (bcs.c:329-706), OutflowFlux at a k = mz-1 outlet (bcs.c:801-870)   */
/* and FormBCS's outlet branch (bcs.c:2285-2418).  threshold = 0.1 (bcs.c:5) */
/* ------------------------------------------------------------------------ */
#define OB_THR 0.1

static double ob_norm(const double *v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

/* rhs.c:34-55 Contra2Cart_single */
void vo_contra2cart_single(const double *csi, const double *eta, const double *zet, const double *uc, double *u)
{
    const double det = csi[0] * (eta[1] * zet[2] - eta[2] * zet[1]) - csi[1] * (eta[0] * zet[2] - eta[2] * zet[0]) +
                       csi[2] * (eta[0] * zet[1] - eta[1] * zet[0]);
    const double det0 = uc[0] * (eta[1] * zet[2] - eta[2] * zet[1]) - uc[1] * (csi[1] * zet[2] - csi[2] * zet[1]) +
                        uc[2] * (csi[1] * eta[2] - csi[2] * eta[1]);
    const double det1 = -uc[0] * (eta[0] * zet[2] - eta[2] * zet[0]) + uc[1] * (csi[0] * zet[2] - csi[2] * zet[0]) -
                        uc[2] * (csi[0] * eta[2] - csi[2] * eta[0]);
    const double det2 = uc[0] * (eta[0] * zet[1] - eta[1] * zet[0]) - uc[1] * (csi[0] * zet[1] - csi[1] * zet[0]) +
                        uc[2] * (csi[0] * eta[1] - csi[1] * eta[0]);
    u[0] = det0 / det;
    u[1] = det1 / det;
    u[2] = det2 / det;
}

int vo_inlet_area(const vo_cfg *c, const vo_geom *g, const double *nvert, double *k_area, double *mean_k_area,
                  double *inlet_area)
{
    DIMS;
    if (c->k_periodic) return -2;
    for (int k = 0; k < mz; k++) k_area[k] = 0;
    for (int k = 1; k < mz - 1; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++)
                if (nvert[IX(k - 1, j, i)] + nvert[IX(k, j, i)] < 0.1 && j >= 1 && j <= my - 2 && i >= 1 && i <= mx - 2)
                    k_area[k] += ob_norm(&g->zet[3 * IX(k, j, i)]);
    k_area[0] = k_area[1];
    double m = 0;
    for (int k = 1; k <= mz - 2; k++) m += k_area[k];
    *mean_k_area = m / (double)(mz - 2);
    double a = 0;
    for (int j = 1; j < my - 1; j++)
        for (int i = 1; i < mx - 1; i++)
            if (nvert[IX(1, j, i)] < OB_THR) a += ob_norm(&g->zet[3 * IX(0, j, i)]);
    *inlet_area = a;
    return 0;
}

/* the inflow speed of InflowFlux's profiles (bcs.c:394-470) */
static double ob_uin(const vo_inflow_opt *o, double yc, double given)
{
    switch (o->inletprofile) {
    case 1:
    case 2: return o->inlet_flux < 0 ? 1. : o->inlet_flux / o->inlet_area;
    case 10: {
        const double delta = 0.45263, a = 5.99;
        if (yc >= delta) return 1.0;
        if (yc <= 0) return 0.0;
        return pow(yc / delta, 1. / a);
    }
    case 21: return 1.5 * 1.0 * yc * (2 - yc);
    default: return given;
    }
}

int vo_inflow_flux(const vo_cfg *c, const vo_geom *g, const double *nvert, const double *coor, const vo_inflow_opt *o,
                   const double *uin_plane, double *ucont, double *ucat, double *ubcs, double *flux_in)
{
    DIMS;
    if (c->k_periodic) return -2;
    const int k = 0;
#define CO(kk, jj, ii, q) coor[3 * IX(kk, jj, ii) + (q)]
    for (int j = 1; j < my - 1; j++)
        for (int i = 1; i < mx - 1; i++) {
            const double yc = (CO(k + 1, j, i, 1) + CO(k + 1, j - 1, i, 1) + CO(k + 1, j, i - 1, 1) + CO(k + 1, j - 1, i - 1, 1)) * 0.25;
            const size_t id = IX(k, j, i), up = IX(k + 1, j, i);
            const double area = ob_norm(&g->zet[3 * id]);
            const double uin = ob_uin(o, yc, uin_plane ? uin_plane[j * mx + i] : 0.);
            if (nvert[up] < OB_THR) {
                ucont[3 * id + 2] = uin * area;
                ucont[3 * id + 0] = 0;
                ucont[3 * id + 1] = 0;
                double u[3];
                vo_contra2cart_single(&g->csi[3 * id], &g->eta[3 * id], &g->zet[3 * id], &ucont[3 * id], u);
                for (int q = 0; q < 3; q++) ucat[3 * id + q] = -ucat[3 * up + q] + 2 * u[q];
                for (int q = 0; q < 3; q++) ubcs[3 * id + q] = ucat[3 * id + q];
            } else {
                ucat[3 * id + 2] = 0;
                ubcs[3 * id + 2] = 0;
                ucont[3 * id + 2] = 0;
            }
        }
#undef CO
    double f = 0;
    for (int j = 1; j < my - 1; j++)
        for (int i = 1; i < mx - 1; i++)
            if (nvert[IX(k + 1, j, i)] < OB_THR) f += ucont[3 * IX(k, j, i) + 2];
    *flux_in = f;
    return 0;
}

int vo_outflow_flux(const vo_cfg *c, const double *nvert, const double *ucont, double *flux_out)
{
    DIMS;
    if (c->k_periodic) return -2;
    const int k = mz - 2;
    double f = 0;
    for (int j = 1; j < my - 1; j++)
        for (int i = 1; i < mx - 1; i++)
            if (nvert[IX(k, j, i)] < OB_THR) f += ucont[3 * IX(k, j, i) + 2];
    *flux_out = f;
    return 0;
}

int vo_outlet_correction(const vo_cfg *c, const vo_geom *g, const double *nvert, double flux_in, int scale,
                         const double *ucat, double *ucont, double *ubcs, double *flux_out, double *area)
{
    DIMS;
    if (c->k_periodic) return -2;
    const int k = mz - 1;
    for (int j = 1; j < my - 1; j++)
        for (int i = 1; i < mx - 1; i++) {
            const size_t id = IX(k, j, i), in = IX(k - 1, j, i);
            ubcs[3 * id] = ucat[3 * in];
            ubcs[3 * id + 1] = ucat[3 * in + 1];
            if (nvert[in] < OB_THR) ubcs[3 * id + 2] = ucat[3 * in + 2];
            else ubcs[3 * id + 2] = ucat[3 * in + 2] * 0;
        }
    double fo = 0, a = 0;
    for (int j = 1; j < my - 1; j++)
        for (int i = 1; i < mx - 1; i++) {
            const size_t in = IX(k - 1, j, i);
            if (nvert[in] < OB_THR) {
                fo += ucont[3 * in + 2];
                a += ob_norm(&g->zet[3 * in]);
            }
        }
    if (scale)
        for (int j = 1; j < my - 1; j++)
            for (int i = 1; i < mx - 1; i++) {
                const size_t in = IX(k - 1, j, i);
                const double A = ob_norm(&g->zet[3 * in]);
                if (nvert[in] < OB_THR) ucont[3 * in + 2] += (flux_in - fo) * A / a;
            }
    *flux_out = fo;
    *area = a;
    return 0;
}

/* ibm.c:3285-3292 Contravariant_Reynolds_stress */
static double ib_crs(double uu, double uv, double uw, double vv, double vw, double ww, const double *a, const double *b)
{
    const double A = uu * a[0] * b[0] + vv * a[1] * b[1] + ww * a[2] * b[2] + uv * (a[0] * b[1] + a[1] * b[0]) +
                     uw * (a[0] * b[2] + a[2] * b[0]) + vw * (a[1] * b[2] + a[2] * b[1]);
    const double B = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]) * sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    return A / B;
}

/* ibm.c:3481-3521 with 3556-3576: the averaging statistics of
 * ibm_interpolation_advanced.  usum, u1sum (uv, vw, wu), u2sum (uu, vv, ww):
 * the running sums over N = ti - 1 steps.  Per element: the mean of
 * ustar_avg^2 (the wall model or noslip on the mean velocity) and of the
 * three contravariant Reynolds stresses of the nodes. */
int vo_ib_averages(const vo_cfg *c, const vo_geom *g, const vo_ib_surface *s, const vo_ib_node *nodes, int n,
                   const vo_ib_opt *opt, const double *usum, const double *u1sum, const double *u2sum, double N,
                   double *mean_shear, double *rs1, double *rs2, double *rs3)
{
    DIMS;
    int rc = check_cfg(c);
    if (rc) return rc;
    int *cnt = calloc(s->n_elmt, sizeof(int));
    for (int e = 0; e < s->n_elmt; e++) mean_shear[e] = rs1[e] = rs2[e] = rs3[e] = 0;
    for (int l = 0; l < n; l++) {
        const vo_ib_node *nd = &nodes[l];
        const int ni = nd->cell;
        const size_t id = IX(nd->k, nd->j, nd->i);
        const size_t q1 = IX(nd->k1, nd->j1, nd->i1), q2 = IX(nd->k2, nd->j2, nd->i2), q3 = IX(nd->k3, nd->j3, nd->i3);
        const double sb = nd->d_s, sc = sb + nd->d_i, sk1 = nd->cr1, sk2 = nd->cr2, sk3 = nd->cr3;
        const double nfx = s->nf_x[ni], nfy = s->nf_y[ni], nfz = s->nf_z[ni];
        double Ua[3] = {0, 0, 0}, Uc[3];
        if (ni >= 0 && s->u) {
            const double *u1 = &s->u[3 * s->nv1[ni]], *u2 = &s->u[3 * s->nv2[ni]], *u3 = &s->u[3 * s->nv3[ni]];
            for (int q = 0; q < 3; q++) Ua[q] = u1[q] * nd->cs1 + u2[q] * nd->cs2 + u3[q] * nd->cs3;
        }
        for (int q = 0; q < 3; q++) Uc[q] = (usum[3 * q1 + q] * sk1 + usum[3 * q2 + q] * sk2 + usum[3 * q3 + q] * sk3) / N;
        const double u_c = Uc[0] - Ua[0], v_c = Uc[1] - Ua[1], w_c = Uc[2] - Ua[2];
        const double un = u_c * nfx + v_c * nfy + w_c * nfz;
        const double ut = u_c - un * nfx, vt = v_c - un * nfy, wt = w_c - un * nfz;
        const double ut_mag = sqrt(ut * ut + vt * vt + wt * wt);
        double us;
        if (opt->wallfunction) us = vo_hydset_utau(1. / c->ren, ut_mag, sc, 0.01, opt->ks);
        else us = sqrt(ut_mag / sc / c->ren);
        double m[3], m2[3], m1[3];
        for (int q = 0; q < 3; q++) {
            m[q] = (usum[3 * q1 + q] * sk1 + usum[3 * q2 + q] * sk2 + usum[3 * q3 + q] * sk3) / N;
            m2[q] = (u2sum[3 * q1 + q] * sk1 + u2sum[3 * q2 + q] * sk2 + u2sum[3 * q3 + q] * sk3) / N;
            m1[q] = (u1sum[3 * q1 + q] * sk1 + u1sum[3 * q2 + q] * sk2 + u1sum[3 * q3 + q] * sk3) / N;
        }
        const double U = m[0], V = m[1], W = m[2];
        const double uu = m2[0] - U * U, vv = m2[1] - V * V, ww = m2[2] - W * W;
        const double uv = m1[0] - U * V, vw = m1[1] - V * W, uw = m1[2] - W * U;
        const double *cs = &g->csi[3 * id], *et = &g->eta[3 * id], *ze = &g->zet[3 * id];
        cnt[ni]++;
        mean_shear[ni] += us * us;
        rs1[ni] += ib_crs(uu, uv, uw, vv, vw, ww, cs, et);
        rs2[ni] += ib_crs(uu, uv, uw, vv, vw, ww, et, ze);
        rs3[ni] += ib_crs(uu, uv, uw, vv, vw, ww, cs, ze);
    }
    for (int e = 0; e < s->n_elmt; e++)
        if (cnt[e]) {
            mean_shear[e] /= (double)cnt[e];
            rs1[e] /= (double)cnt[e];
            rs2[e] /= (double)cnt[e];
            rs3[e] /= (double)cnt[e];
        }
    free(cnt);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* rhs.c:1578-1700 RungeKutta: the boundary work before the stages           */
/* (1604-1614), the four stages (1633-1662) and FormBCS with                 */
/* outflow_scale = 1 after them (1685-1686)                                  */
/* ------------------------------------------------------------------------ */
int vo_runge_kutta(const vo_cfg *c, const vo_geom *g, const double *nvert, const double *coor,
                   const vo_inflow_opt *inflow, const double *uin_plane, double *ubcs, const vo_ib_surface *s,
                   const vo_ib_node *nodes, int n_nodes, const vo_ib_opt *ib, double *p, double *ustar,
                   const double *f_eul, double *ucont, double *ucat, const double *nu_t, const double *dP, double *u0,
                   double *rhs, double *flux_in, double *flux_out)
{
    DIMS;
    const size_t P = (size_t)mx * my * mz;
    const int inlet = !c->k_periodic && c->bctype[4] == 5, outlet = !c->k_periodic && c->bctype[5] == 4;
    double fin = 0, fout = 0, area = 0;
    int rc = 0;
    if ((inlet || outlet) && !ubcs) return -2;
    if (inlet && (rc = vo_inflow_flux(c, g, nvert, coor, inflow, uin_plane, ucont, ucat, ubcs, &fin))) return rc;
    if (outlet && (rc = vo_outlet_correction(c, g, nvert, fin, 0, ucat, ucont, ubcs, &fout, &area))) return rc;
    if (c->immersed && s) {
        /* the element statistics the reconstruction also forms (not used here) */
        const vo_ib_opt dflt = {0, 0, 0.0, 0.0};
        int *cnt = calloc((size_t)s->n_elmt, sizeof(int));
        double *sh = calloc(2 * (size_t)s->n_elmt, sizeof(double));
        if (!cnt || !sh) { free(cnt); free(sh); return -1; }
        rc = vo_ib_interpolate_opt(c, g, nvert, s, nodes, n_nodes, ib ? ib : &dflt, ucont, ucat, p, ustar, cnt, sh,
                                   sh + s->n_elmt);
        free(cnt);
        free(sh);
        if (rc) return rc;
    }
    memcpy(u0, ucont, 3 * P * sizeof(double));
    const double alfa[4] = {0.25, 1 / 3., 0.5, 1.};
    for (int st = 0; st < 4; st++) {
        if ((rc = vo_contra2cart(c, g, nvert, ucont, ucat))) return rc;
        memset(rhs, 0, 3 * P * sizeof(double));
        if ((rc = vo_formfunction2(c, g, nvert, ucont, ucat, nu_t, rhs, 1.0))) return rc;
        const double a = alfa[st] * c->dt * c->st;
        for (size_t n = 0; n < 3 * P; n++) {
            rhs[n] += -1 * dP[n];
            if (f_eul) rhs[n] += 1. * f_eul[n];
            ucont[n] = a * rhs[n] + u0[n];
        }
        if ((rc = vo_contra2cart(c, g, nvert, ucont, ucat))) return rc;
    }
    if (outlet && (rc = vo_outlet_correction(c, g, nvert, fin, 1, ucat, ucont, ubcs, &fout, &area))) return rc;
    if (flux_in) *flux_in = fin;
    if (flux_out) *flux_out = fout;
    return 0;
}
