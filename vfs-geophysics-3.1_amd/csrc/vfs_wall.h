// The Cabot wall model behind the wall-function walls (bctype -1,
// wallfunction.c), device side.  The y+ integral table is built on the host
// (vfs_wall.hip) with the reference's arithmetic and read here; every
// expression keeps the reference's operation order.
#pragma once
#include "vfs_internal.h"

#define VFS_WF_KAPPA 0.41          // wallfunction.c:15
#define VFS_WF_INTERVAL 2          // interval_yp, wallfunction.c:288
#define VFS_WF_MAX_YP 1000000      // max_yp
#define VFS_WF_NYP (VFS_WF_MAX_YP / VFS_WF_INTERVAL)

// wallfunction.c:271-274 (nu_t / nu); pow(x, 2.0) of the reference is x * x
__host__ __device__ __forceinline__ double wf_nut(double yplus)
{
    const double e = 1. - exp(-yplus / 19.);
    return VFS_WF_KAPPA * yplus * (e * e);
}

// wallfunction.c:325-371 integrate_F: the table interpolated linearly up to
// y+ = 1e6, four Simpson 3/8 panels beyond.  The table index is clamped to
// the table (the reference reads outside it for y+ < 0).
__device__ __forceinline__ double wf_integrate_F(const double *__restrict__ buf, double nu, double utau, double yb)
{
    double val = 0;
    double ya_plus = 0 * utau / nu;
    const double yb_plus = yb * utau / nu;
    int ib = (int)(yb_plus / (double)VFS_WF_INTERVAL);
    const int N = 4;
    if (yb_plus <= (double)VFS_WF_MAX_YP) {
        ib = ib < 0 ? 0 : (ib > VFS_WF_NYP - 1 ? VFS_WF_NYP - 1 : ib);
        const double int_a = 0;
        const double int_b = (buf[ib + 1] - buf[ib]) / (double)(VFS_WF_INTERVAL) *
                                 (yb_plus - (double)(ib)*VFS_WF_INTERVAL) + buf[ib];
        return (int_b - int_a) / utau;
    }
    val = buf[VFS_WF_NYP];
    ya_plus = VFS_WF_MAX_YP;
    const double ydiff = yb_plus - ya_plus, dy = ydiff / (double)N;
    double E[N + 1];
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

// wallfunction.c:234-237, 252-256 and 389-404 (dpdn = 0): Newton on
// utau^2 * integrate_F(utau) = u with a central-difference derivative
__device__ __forceinline__ double wf_f(const double *buf, double nu, double u, double y, double utau)
{
    return utau * utau * wf_integrate_F(buf, nu, utau, y) - u;
}
__device__ __forceinline__ double wf_find_utau(const double *buf, double nu, double u, double y, double guess)
{
    double x = guess, x0 = guess;
    for (int i = 0; i < 30; i++) {
        const double eps = 1.e-7;
        const double df = (wf_f(buf, nu, u, y, x0 + eps) - wf_f(buf, nu, u, y, x0 - eps)) / (2 * eps);
        x = x0 - wf_f(buf, nu, u, y, x0) / df;
        if (fabs(x0 - x) < 1.e-10) break;
        x0 = x;
    }
    return x;
}

// wallfunction.c:18-43 wall_function with Ua = 0
__device__ __forceinline__ void wf_wall_function(const double *buf, double nu, double sc, double sb, const double Uc[3],
                                                 double Ub[3], double *ustar, double nx, double ny, double nz)
{
    const double u_c = Uc[0] - 0., v_c = Uc[1] - 0., w_c = Uc[2] - 0.;
    const double un = u_c * nx + v_c * ny + w_c * nz;
    double ut = u_c - un * nx, vt = v_c - un * ny, wt = w_c - un * nz;
    const double ut_mag = sqrt(ut * ut + vt * vt + wt * wt);
    *ustar = wf_find_utau(buf, nu, ut_mag, sc, 0.01);
    const double ut_mag_modeled = *ustar * *ustar * wf_integrate_F(buf, nu, *ustar, sb);
    if (ut_mag > 1.e-10) {
        ut *= ut_mag_modeled / ut_mag;
        vt *= ut_mag_modeled / ut_mag;
        wt *= ut_mag_modeled / ut_mag;
    } else {
        ut = vt = wt = 0;
    }
    Ub[0] = ut + sb / sc * un * nx;
    Ub[1] = vt + sb / sc * un * ny;
    Ub[2] = wt + sb / sc * un * nz;
    Ub[0] += 0.;
    Ub[1] += 0.;
    Ub[2] += 0.;
}

// rhs2.c:539-583 Calculate_Covariant_metrics + Calculate_normal: unit
// covariant base vector of direction dir (0 = csi, 1 = eta, 2 = zet)
__device__ __forceinline__ void wf_normal(const double *csi, const double *eta, const double *zet, int dir, double n[3])
{
    const double a11 = csi[0], a12 = csi[1], a13 = csi[2];
    const double a21 = eta[0], a22 = eta[1], a23 = eta[2];
    const double a31 = zet[0], a32 = zet[1], a33 = zet[2];
    const double det = a11 * (a33 * a22 - a32 * a23) - a21 * (a33 * a12 - a32 * a13) + a31 * (a23 * a12 - a22 * a13);
    double x, y, z;
    if (dir == 0) {
        x = (a33 * a22 - a32 * a23) / det; y = -(a33 * a21 - a31 * a23) / det; z = (a32 * a21 - a31 * a22) / det;
    } else if (dir == 1) {
        x = -(a33 * a12 - a32 * a13) / det; y = (a33 * a11 - a31 * a13) / det; z = -(a32 * a11 - a31 * a12) / det;
    } else {
        x = (a23 * a12 - a22 * a13) / det; y = -(a23 * a11 - a21 * a13) / det; z = (a22 * a11 - a21 * a12) / det;
    }
    const double sum = sqrt(x * x + y * y + z * z);
    x /= sum, y /= sum, z /= sum;
    n[0] = x; n[1] = y; n[2] = z;
}
