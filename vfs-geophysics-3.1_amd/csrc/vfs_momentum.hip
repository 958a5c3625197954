// Momentum right-hand side of the fractional step on gfx950:
// Contra2Cart_2 (rhs.c:57-741), Compute_eddy_viscosity_LES (les.c:1135-1252),
// Pressure_Gradient (momentum.c:197-440), IB_BC (momentum.c:1982-2195),
// Formfunction_2 (momentum.c:456-1979), FormFunction_SNES (2197-2301) and
// one RungeKutta stage (rhs.c:1637-1659).
//
// Every expression keeps the reference's operation order and the library is
// built with -ffp-contract=off, so results round exactly like the reference's
// scalar loops.  Formfunction_2 is the dominant kernel (53 words per cell of
// compulsory HBM traffic); it runs as the fused k-marching kernel in
// vfs_ff2_fused.hip, with the face/cell/projection passes below kept as the
// reference-structured path for the boundary variants it does not cover.
#include "vfs_internal.h"
#include "vfs_reduce.h"
#include "vfs_stencil.h"

// ---------------------------------------------------------------------------
// Contra2Cart_2
// ---------------------------------------------------------------------------
__global__ void k_c2c_interior(Dims d, Geom g, const double *__restrict__ nvert,
                               const double *__restrict__ ucont, double *__restrict__ ucat)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    if (i < 1 || j < 1 || k < 1 || i > d.mx - 2 || j > d.my - 2 || k > d.mz - 2) return;
    if (!(nvert[t] < 0.1)) return;
    double m00 = g.csi[3 * t], m01 = g.csi[3 * t + 1], m02 = g.csi[3 * t + 2];
    double m10 = g.eta[3 * t], m11 = g.eta[3 * t + 1], m12 = g.eta[3 * t + 2];
    double m20 = g.zet[3 * t], m21 = g.zet[3 * t + 1], m22 = g.zet[3 * t + 2];
    double q0 = 0.5 * (ucont[3 * (t - 1)] + ucont[3 * t]);
    double q1 = 0.5 * (ucont[3 * (t - d.mx) + 1] + ucont[3 * t + 1]);
    double q2 = 0.5 * (ucont[3 * (t - (long long)d.mx * d.my) + 2] + ucont[3 * t + 2]);
    double det = m00 * (m11 * m22 - m12 * m21) - m01 * (m10 * m22 - m12 * m20) + m02 * (m10 * m21 - m11 * m20);
    double det0 = q0 * (m11 * m22 - m12 * m21) - q1 * (m01 * m22 - m02 * m21) + q2 * (m01 * m12 - m02 * m11);
    double det1 = -q0 * (m10 * m22 - m12 * m20) + q1 * (m00 * m22 - m02 * m20) - q2 * (m00 * m12 - m02 * m10);
    double det2 = q0 * (m10 * m21 - m11 * m20) - q1 * (m00 * m21 - m01 * m20) + q2 * (m00 * m11 - m01 * m10);
    ucat[3 * t] = det0 / det;
    ucat[3 * t + 1] = det1 / det;
    ucat[3 * t + 2] = det2 / det;
}

// Ghost-point boundary rules of rhs.c:294-674 (no-slip mirror, solid ghosts).
// L is the snapshot the reference reads (lUcat); it may alias ucat when at
// most one direction is non-periodic (then no rule reads a point another
// rule writes).
__global__ void k_c2c_bc_ghost(Dims d, const double *__restrict__ nvert, const double *L, double *ucat)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int k, j, i;
    if (t >= shell_count(d) || !shell_point(d, t, k, j, i)) return;
    long long id = IX(d, k, j, i);
    if ((int)(nvert[id] + 0.1) == 3) { ucat[3 * id] = ucat[3 * id + 1] = ucat[3 * id + 2] = 0; return; }
    long long s = -1;
    if (i == 0 && d.bc[0] == 1 && (j != 0 && k != 0)) s = IX(d, k, j, i + 1);
    if (i == d.mx - 1 && d.bc[1] == 1 && (j != 0 && k != 0)) s = IX(d, k, j, i - 1);
    if (j == 0 && d.bc[2] == 1 && (i != 0 && k != 0)) s = IX(d, k, j + 1, i);
    if (j == d.my - 1 && d.bc[3] == 1 && (i != 0 && k != 0)) s = IX(d, k, j - 1, i);
    if (k == 0 && d.bc[4] == 1 && (i != 0 && j != 0)) s = IX(d, k + 1, j, i);
    if (k == d.mz - 1 && d.bc[5] == 1 && (i != 0 && j != 0)) s = IX(d, k - 1, j, i);
    if (s >= 0) {
        double x = L[3 * s], y = L[3 * s + 1], z = L[3 * s + 2];
        ucat[3 * id] = -1 * x; ucat[3 * id + 1] = -1 * y; ucat[3 * id + 2] = -1 * z;
    }
    // outlet (rhs.c:646-663): zero behind a solid last cell; the inlet ghost
    // is InflowFlux's (rhs.c:636-644 is disabled)
    if (d.bc[5] == 4 && k == d.mz - 1 && (i != 0 && i != d.mx - 1 && j != 0 && j != d.my - 1) &&
        nvert[IX(d, k - 1, j, i)] > 0.1)
        ucat[3 * id] = ucat[3 * id + 1] = ucat[3 * id + 2] = 0;
}

// interior part of the boundary loop: solid cells and the 101220 corner rule
__global__ void k_c2c_bc_interior(Dims d, const double *__restrict__ nvert, double *ucat)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    if (i < 1 || j < 1 || k < 1 || i > d.mx - 2 || j > d.my - 2 || k > d.mz - 2) return;
    bool z = (int)(nvert[t] + 0.1) == 3;
    if (i != 0 && i != d.mx - 1 && j != 0 && j != d.my - 2 && k != 0 && k != d.mz - 2) {
        if (d.bc[0] <= 1 && d.bc[2] <= 1 && (i == 1 && j == 1)) z = true;
        if (d.bc[1] <= 1 && d.bc[2] <= 1 && (i == d.mx - 2 && j == 1)) z = true;
        if (d.bc[0] <= 1 && d.bc[3] <= 1 && (i == 1 && j == d.my - 2)) z = true;
        if (d.bc[1] <= 1 && d.bc[3] <= 1 && (i == d.mx - 2 && j == d.my - 2)) z = true;
    }
    if (z) ucat[3 * t] = ucat[3 * t + 1] = ucat[3 * t + 2] = 0;
}

extern "C" int vfs_contra2cart(vfs_ctx *c, double *ucont, double *ucat)
{
    const Dims d = c->d;
    if (vfs_periodic_vec(c, ucont)) return -1;
    hipLaunchKernelGGL(k_c2c_interior, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d, c->g, c->nvert, ucont, ucat);
    VFS_CHECK_LAUNCH();
    if (vfs_periodic_vec(c, ucat)) return -1;
    const double *L = ucat;
    int nonper = !d.ip + !d.jp + !d.kp;
    if (nonper >= 2 || vfs_has_wallfn(d)) {
        VFS_CHECK_HIP(hipMemcpyAsync(c->snap, ucat, 3 * (size_t)d.P * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        L = c->snap;
    }
    long long ns = shell_count(d);
    hipLaunchKernelGGL(k_c2c_bc_ghost, dim3(grid_for(ns, 256)), dim3(256), 0, c->stream, d, c->nvert, L, ucat);
    VFS_CHECK_LAUNCH();
    if (vfs_c2c_wall(c, L, ucat)) return -1;   // wall-function cells (rhs.c:303-401)
    hipLaunchKernelGGL(k_c2c_bc_interior, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d, c->nvert, ucat);
    VFS_CHECK_LAUNCH();
    return vfs_periodic_vec(c, ucat);
}

// ---------------------------------------------------------------------------
// Compute_eddy_viscosity_LES
// ---------------------------------------------------------------------------
__global__ void k_les(Dims d, Geom g, const double *__restrict__ nvert, const double *__restrict__ ucat,
                      double cs, const double *__restrict__ csf, double *__restrict__ nu_t)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    if (i < 1 || j < 1 || k < 1 || i > d.mx - 2 || j > d.my - 2 || k > d.mz - 2) return;
    if (nvert[t] > 1.1) { nu_t[t] = 0; return; }
    double ajc = g.aj[t];
    double csi0 = g.csi[3 * t], csi1 = g.csi[3 * t + 1], csi2 = g.csi[3 * t + 2];
    double eta0 = g.eta[3 * t], eta1 = g.eta[3 * t + 1], eta2 = g.eta[3 * t + 2];
    double zet0 = g.zet[3 * t], zet1 = g.zet[3 * t + 1], zet2 = g.zet[3 * t + 2];
    double dc[3], de[3], dz[3];
    du_center(d, i, j, k, ucat, nvert, dc, de, dz);
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
    double filter = pow(1. / g.aj[t], 1. / 3.);
    nu_t[t] = (csf ? csf[t] : cs) * pow(filter, 2.0) * Sabs;
}

extern "C" int vfs_eddy_viscosity(vfs_ctx *c, const double *ucat, double cs, double *nu_t)
{
    const Dims d = c->d;
    hipLaunchKernelGGL(k_les, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d, c->g, c->nvert, ucat, cs,
                       (const double *)nullptr, nu_t);
    VFS_CHECK_LAUNCH();
    return vfs_periodic_sca(c, nu_t);
}

// the same with the per-cell Cs of the dynamic procedure (user->lCs)
extern "C" int vfs_eddy_viscosity_cs(vfs_ctx *c, const double *ucat, const double *cs, double *nu_t)
{
    VFS_REQUIRE(cs, "cs field required");
    const Dims d = c->d;
    hipLaunchKernelGGL(k_les, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d, c->g, c->nvert, ucat, 0.0, cs,
                       nu_t);
    VFS_CHECK_LAUNCH();
    return vfs_periodic_sca(c, nu_t);
}

// ---------------------------------------------------------------------------
// Pressure_Gradient
// ---------------------------------------------------------------------------
#define IB1(x) ((int)((x) + 0.5) == 1)
// kf: {mean_k_flux, mean_k_area} of the flux forcing (k-periodic without
// dpdz_set), fin the target flux (-flux)
__global__ void k_pgrad(Dims d, Geom g, const double *__restrict__ nvert, const double *__restrict__ p,
                        double *__restrict__ dp, const double *__restrict__ kf, double fin)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    if (i < 1 || j < 1 || k < 1 || i > d.mx - 2 || j > d.my - 2 || k > d.mz - 2) {
        dp[3 * t] = dp[3 * t + 1] = dp[3 * t + 2] = 0;
        return;
    }
#define PP(kk, jj, ii) p[IX(d, kk, jj, ii)]
#define NV(kk, jj, ii) nvert[IX(d, kk, jj, ii)]
    const double *ic = g.icsi + 3 * t, *ie = g.ieta + 3 * t, *iz = g.izet + 3 * t;
    const double *jc = g.jcsi + 3 * t, *je = g.jeta + 3 * t, *jz = g.jzet + 3 * t;
    const double *kc = g.kcsi + 3 * t, *ke = g.keta + 3 * t, *kz = g.kzet + 3 * t;
    double g11_i = (ic[0] * ic[0] + ic[1] * ic[1] + ic[2] * ic[2]);
    double g12_i = (ie[0] * ic[0] + ie[1] * ic[1] + ie[2] * ic[2]);
    double g13_i = (iz[0] * ic[0] + iz[1] * ic[1] + iz[2] * ic[2]);
    double g21_j = (jc[0] * je[0] + jc[1] * je[1] + jc[2] * je[2]);
    double g22_j = (je[0] * je[0] + je[1] * je[1] + je[2] * je[2]);
    double g23_j = (jz[0] * je[0] + jz[1] * je[1] + jz[2] * je[2]);
    double g31_k = (kc[0] * kz[0] + kc[1] * kz[1] + kc[2] * kz[2]);
    double g32_k = (ke[0] * kz[0] + ke[1] * kz[1] + ke[2] * kz[2]);
    double g33_k = (kz[0] * kz[0] + kz[1] * kz[1] + kz[2] * kz[2]);
    const double r1 = 1.0, r2 = 1.0, r3 = 1.0;
    const int ip = d.ip, jp = d.jp, kp = d.kp, mx = d.mx, my = d.my, mz = d.mz;
    double dpdc, dpde, dpdz, ox, oy, oz;

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
    ox = (dpdc * g11_i + dpde * g12_i + dpdz * g13_i) * g.iaj[t] / r1;

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
    oy = (dpdc * g21_j + dpde * g22_j + dpdz * g23_j) * g.jaj[t] / r2;

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
    if (kp) {   // momentum.c:382-398
        double dzz = 1. / g.kaj[t] / sqrt(kz[0] * kz[0] + kz[1] * kz[1] + kz[2] * kz[2]);
        if (d.dpdz_set) dpdz += d.mpg * dzz;
        else dpdz += dzz * (kf[0] - fin) / d.dt / kf[1];
    }
    oz = (dpdc * g31_k + dpde * g32_k + dpdz * g33_k) * g.kaj[t] / r3;

    if (i == 0 || NV(k, j, i) + NV(k, j, i + 1) > 0.1 || (!ip && i == mx - 2)) ox = 0;
    if (j == 0 || NV(k, j, i) + NV(k, j + 1, i) > 0.1 || (!jp && j == my - 2)) oy = 0;
    if (k == 0 || NV(k, j, i) + NV(k + 1, j, i) > 0.1 || (!kp && k == mz - 2)) oz = 0;
    dp[3 * t] = ox; dp[3 * t + 1] = oy; dp[3 * t + 2] = oz;
#undef PP
#undef NV
}

extern "C" int vfs_pressure_gradient(vfs_ctx *c, double *p, double *dP)
{
    const Dims d = c->d;
    VFS_REQUIRE(!d.kp || d.dpdz_set || c->kflux_ready,
                "k-periodic flux forcing: call vfs_calc_k_flux (Calc_k_Flux) before vfs_pressure_gradient");
    if (vfs_periodic_sca(c, p)) return -1;
    hipLaunchKernelGGL(k_pgrad, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d, c->g, c->nvert, p, dP,
                       c->kflux, c->inlet_flux);
    VFS_CHECK_LAUNCH();
    return 0;
}

// ---------------------------------------------------------------------------
// IB_BC
// ---------------------------------------------------------------------------
__global__ void k_ibbc(Dims d, Geom g, const double *__restrict__ nvert, const double *__restrict__ ucat,
                       double *__restrict__ ucont)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    if (i < 1 || j < 1 || k < 1 || i > d.mx - 2 || j > d.my - 2 || k > d.mz - 2) return;
    const long long sj = d.mx, sk = (long long)d.mx * d.my;
    const double f = 1.0;
    const double *ic = g.icsi + 3 * t, *je = g.jeta + 3 * t, *kz = g.kzet + 3 * t;
    double ucx, ucy, ucz;
    bool self = (int)nvert[t] == 1;
    if (self || (int)(nvert[t + 1]) == 1) {
        ucx = (ucat[3 * t] + ucat[3 * (t + 1)]) * 0.5;
        ucy = (ucat[3 * t + 1] + ucat[3 * (t + 1) + 1]) * 0.5;
        ucz = (ucat[3 * t + 2] + ucat[3 * (t + 1) + 2]) * 0.5;
        ucont[3 * t] = (ucx * ic[0] + ucy * ic[1] + ucz * ic[2]) * f;
    }
    if (self || (int)(nvert[t + sj]) == 1) {
        ucx = (ucat[3 * (t + sj)] + ucat[3 * t]) * 0.5;
        ucy = (ucat[3 * (t + sj) + 1] + ucat[3 * t + 1]) * 0.5;
        ucz = (ucat[3 * (t + sj) + 2] + ucat[3 * t + 2]) * 0.5;
        ucont[3 * t + 1] = (ucx * je[0] + ucy * je[1] + ucz * je[2]) * f;
    }
    if (self || (int)(nvert[t + sk]) == 1) {
        ucx = (ucat[3 * (t + sk)] + ucat[3 * t]) * 0.5;
        ucy = (ucat[3 * (t + sk) + 1] + ucat[3 * t + 1]) * 0.5;
        ucz = (ucat[3 * (t + sk) + 2] + ucat[3 * t + 2]) * 0.5;
        ucont[3 * t + 2] = (ucx * kz[0] + ucy * kz[1] + ucz * kz[2]) * f;
    }
}

// periodic face copies of IB_BC (momentum.c:2167-2181): one component each
__global__ void k_periodic_faces(Dims d, double *ucont)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int k, j, i;
    if (t >= shell_count(d) || !shell_point(d, t, k, j, i)) return;
    long long o = IX(d, k, j, i);
    if (d.ip && i == 0) ucont[3 * o] = ucont[3 * IX(d, k, j, d.mx - 2)];
    if (d.ip && i == d.mx - 1) ucont[3 * o] = ucont[3 * IX(d, k, j, 1)];
    if (d.jp && j == 0) ucont[3 * o + 1] = ucont[3 * IX(d, k, d.my - 2, i) + 1];
    if (d.jp && j == d.my - 1) ucont[3 * o + 1] = ucont[3 * IX(d, k, 1, i) + 1];
    if (d.kp && k == 0) ucont[3 * o + 2] = ucont[3 * IX(d, d.mz - 2, j, i) + 2];
    if (d.kp && k == d.mz - 1) ucont[3 * o + 2] = ucont[3 * IX(d, 1, j, i) + 2];
}

extern "C" int vfs_ib_bc(vfs_ctx *c, double *ucont, double *ucat)
{
    const Dims d = c->d;
    if (vfs_periodic_vec(c, ucat)) return -1;
    if (d.immersed) {
        hipLaunchKernelGGL(k_ibbc, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d, c->g, c->nvert, ucat, ucont);
        VFS_CHECK_LAUNCH();
    }
    if (vfs_ibbc_wall(c, ucont)) return -1;
    if (d.ip || d.jp || d.kp) {
        hipLaunchKernelGGL(k_periodic_faces, dim3(grid_for(shell_count(d), 256)), dim3(256), 0, c->stream, d, ucont);
        VFS_CHECK_LAUNCH();
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Formfunction_2, reference-structured passes
// ---------------------------------------------------------------------------
// face fluxes of all three directions at point (k,j,i): div and visc
__global__ void __launch_bounds__(256) k_ff2_faces(Dims d, Geom g, const double *__restrict__ nvert, const double *__restrict__ ucont,
                            const double *__restrict__ ucat, const double *__restrict__ nu_t, double *__restrict__ face)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    if (i > d.mx - 2 || j > d.my - 2 || k > d.mz - 2) return;
    const long long P = d.P;
    double D[3], V[3], A[3];
    if (j >= 1 && k >= 1) {
        face_flux<0>(d, g, nvert, ucont, ucat, nu_t, i, j, k, D, V, A);
#pragma unroll
        for (int q = 0; q < 3; q++) { face[(0 * 3 + q) * P + t] = D[q]; face[(3 * 3 + q) * P + t] = V[q]; }
        if (d.skew)
            for (int q = 0; q < 3; q++) face[(6 * 3 + q) * P + t] = A[q];
    }
    if (i >= 1 && k >= 1) {
        face_flux<1>(d, g, nvert, ucont, ucat, nu_t, i, j, k, D, V, A);
#pragma unroll
        for (int q = 0; q < 3; q++) { face[(1 * 3 + q) * P + t] = D[q]; face[(4 * 3 + q) * P + t] = V[q]; }
        if (d.skew)
            for (int q = 0; q < 3; q++) face[(7 * 3 + q) * P + t] = A[q];
    }
    if (i >= 1 && j >= 1) {
        face_flux<2>(d, g, nvert, ucont, ucat, nu_t, i, j, k, D, V, A);
#pragma unroll
        for (int q = 0; q < 3; q++) { face[(2 * 3 + q) * P + t] = D[q]; face[(5 * 3 + q) * P + t] = V[q]; }
        if (d.skew)
            for (int q = 0; q < 3; q++) face[(8 * 3 + q) * P + t] = A[q];
    }
}

// cell divergence with the 4th-order deferred correction (momentum.c:1506-1636)
__global__ void k_ff2_cells(Dims d, const double *__restrict__ nvert, const double *__restrict__ face, double *__restrict__ fp)
{
    long long t = (long long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    if (i < 1 || j < 1 || k < 1 || i > d.mx - 2 || j > d.my - 2 || k > d.mz - 2) return;
    double f[3];
    cell_fp(d, nvert, face, i, j, k, f);
    fp[3 * t] = f[0]; fp[3 * t + 1] = f[1]; fp[3 * t + 2] = f[2];
}

// projection onto faces, masks and boundary planes (momentum.c:1673-1900)
__global__ void k_ff2_rhs(Dims d, Geom g, const double *__restrict__ nvert, const double *__restrict__ fp,
                          double *__restrict__ rhs, double scale)
{
    long long t = (long long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    if (i < 1 || j < 1 || k < 1 || i > d.mx - 2 || j > d.my - 2 || k > d.mz - 2) {
        rhs[3 * t] = rhs[3 * t + 1] = rhs[3 * t + 2] = 0;
        return;
    }
    double r[3] = {rhs[3 * t], rhs[3 * t + 1], rhs[3 * t + 2]};
    project_rhs(d, g, nvert, fp, i, j, k, t, scale, r);
    rhs[3 * t] = r[0]; rhs[3 * t + 1] = r[1]; rhs[3 * t + 2] = r[2];
}

int vfs_ff2_passes(vfs_ctx *c, const double *ucont, const double *ucat, const double *nu_t, double *rhs, double scale)
{
    const Dims d = c->d;
    unsigned gb = grid_for(d.P, 256);
    hipLaunchKernelGGL(k_ff2_faces, dim3(gb), dim3(256), 0, c->stream, d, c->g, c->nvert, ucont, ucat, nu_t, c->face);
    VFS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_ff2_cells, dim3(gb), dim3(256), 0, c->stream, d, c->nvert, c->face, c->fp);
    VFS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_ff2_rhs, dim3(gb), dim3(256), 0, c->stream, d, c->g, c->nvert, c->fp, rhs, scale);
    VFS_CHECK_LAUNCH();
    return 0;
}

// cell divergence and projection of face fluxes already in c->face
int vfs_ff2_from_faces(vfs_ctx *c, double *rhs, double scale)
{
    const Dims d = c->d;
    unsigned gb = grid_for(d.P, 256);
    hipLaunchKernelGGL(k_ff2_cells, dim3(gb), dim3(256), 0, c->stream, d, c->nvert, c->face, c->fp);
    VFS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_ff2_rhs, dim3(gb), dim3(256), 0, c->stream, d, c->g, c->nvert, c->fp, rhs, scale);
    VFS_CHECK_LAUNCH();
    return 0;
}

// projection of fp already in c->fp
int vfs_ff2_rhs_only(vfs_ctx *c, double *rhs, double scale)
{
    const Dims d = c->d;
    hipLaunchKernelGGL(k_ff2_rhs, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d, c->g, c->nvert, c->fp, rhs, scale);
    VFS_CHECK_LAUNCH();
    return 0;
}

int vfs_ff2_fused(vfs_ctx *c, const double *ucont, const double *ucat, const double *nu_t, double *rhs, double scale);

extern "C" int vfs_formfunction2(vfs_ctx *c, double *ucont, const double *ucat, const double *nu_t,
                                 double *rhs, double scale)
{
    VFS_REQUIRE(nu_t || !c->d.les, "nu_t required when les is on");
    if (vfs_periodic_vec(c, ucont)) return -1;
    const char *path = getenv("VFS_FF2_PATH");
    if (path && path[0] == 'p') return vfs_ff2_passes(c, ucont, ucat, nu_t, rhs, scale);
    return vfs_ff2_fused(c, ucont, ucat, nu_t, rhs, scale);
}

// ---------------------------------------------------------------------------
// FormFunction_SNES, RK stage, Calc_Minimum_dt
// ---------------------------------------------------------------------------
// Ug = U with no-slip boundary fluxes zeroed (momentum.c:2224-2232); lucont = Ug;
// rhs = -1/dt * Ug + 1/dt * U_o (momentum.c:2263-2266)
__global__ void k_snes_init(Dims d, const double *__restrict__ U, const double *__restrict__ U_o,
                            double *__restrict__ lucont, double *__restrict__ rhs)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    double u[3] = {U[3 * t], U[3 * t + 1], U[3 * t + 2]};
    if ((i == 0 && d.bc[0] == 1) || (i == d.mx - 2 && d.bc[1] == 1)) u[0] = 0;
    if ((j == 0 && d.bc[2] == 1) || (j == d.my - 2 && d.bc[3] == 1)) u[1] = 0;
    if ((k == 0 && d.bc[4] == 1) || (k == d.mz - 2 && d.bc[5] == 1)) u[2] = 0;
    const double dt = d.dt;
#pragma unroll
    for (int q = 0; q < 3; q++) {
        lucont[3 * t + q] = u[q];
        double r = 0;
        r += -1. / dt * u[q];
        r += 1. / dt * U_o[3 * t + q];
        rhs[3 * t + q] = r;
    }
}

// rhs += 0.5 rhs_o; rhs -= dP; rhs += F_eul when a body force is attached
// (momentum.c:2277-2288)
__global__ void k_snes_final(long long n, const double *__restrict__ rhs_o, const double *__restrict__ dP,
                             const double *__restrict__ f_eul, double *rhs)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double r = rhs[t];
    r += 0.5 * rhs_o[t];
    r += -1 * dP[t];
    if (f_eul) r += 1. * f_eul[t];
    rhs[t] = r;
}

extern "C" int vfs_formfunction_snes(vfs_ctx *c, const double *U, const double *U_o, const double *rhs_o,
                                     const double *dP, const double *nu_t, double *lucont, double *ucat, double *rhs)
{
    const Dims d = c->d;
    hipLaunchKernelGGL(k_snes_init, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d, U, U_o, lucont, rhs);
    VFS_CHECK_LAUNCH();
    if (vfs_contra2cart(c, lucont, ucat)) return -1;
    if (vfs_ib_bc(c, lucont, ucat)) return -1;
    if (vfs_formfunction2(c, lucont, ucat, nu_t, rhs, 0.5)) return -1;
    hipLaunchKernelGGL(k_snes_final, dim3(grid_for(3 * d.P, 256)), dim3(256), 0, c->stream, 3 * d.P, rhs_o, dP,
                       c->body_force, rhs);
    VFS_CHECK_LAUNCH();
    return 0;
}

// rhs -= dP; ucont = a*rhs + u0  (rhs.c:1646-1650, VecAXPY then VecWAXPY)
__global__ void k_rk_update(long long n, double a, const double *__restrict__ dP, const double *__restrict__ u0,
                            double *__restrict__ rhs, double *__restrict__ ucont)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double r = rhs[t];
    r += -1 * dP[t];
    rhs[t] = r;
    ucont[t] = a * r + u0[t];
}

extern "C" int vfs_rk_stage(vfs_ctx *c, const double *u0, double *ucont, double *ucat, const double *nu_t,
                            const double *dP, double alfa, double *rhs)
{
    const Dims d = c->d;
    if (vfs_contra2cart(c, ucont, ucat)) return -1;
    VFS_CHECK_HIP(hipMemsetAsync(rhs, 0, 3 * (size_t)d.P * sizeof(double), c->stream));
    if (vfs_formfunction2(c, ucont, ucat, nu_t, rhs, 1.0)) return -1;
    double a = alfa * d.dt * d.st;
    hipLaunchKernelGGL(k_rk_update, dim3(grid_for(3 * d.P, 256)), dim3(256), 0, c->stream, 3 * d.P, a, dP, u0, rhs, ucont);
    VFS_CHECK_LAUNCH();
    return vfs_contra2cart(c, ucont, ucat);
}

__global__ void k_min_dt(Dims d, Geom g, const double *__restrict__ ucont, double *part)
{
    __shared__ double sh[4];
    double acc = -1.0e308;   // max of -ldt
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < d.P; t += (long long)gridDim.x * blockDim.x) {
        int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
        if (i < 1 || j < 1 || k < 1 || i > d.mx - 2 || j > d.my - 2 || k > d.mz - 2) continue;
        double a = g.aj[t];
        acc = fmax(acc, -fabs(1. / ucont[3 * t] / a));
        acc = fmax(acc, -fabs(1. / ucont[3 * t + 1] / a));
        acc = fmax(acc, -fabs(1. / ucont[3 * t + 2] / a));
    }
    double r = block_max<256>(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

extern "C" int vfs_min_dt(vfs_ctx *c, const double *ucont, double *dt_out)
{
    const Dims d = c->d;
    int nb = (int)grid_for(d.P, 256);
    if (nb > VFS_RED_BLOCKS) nb = VFS_RED_BLOCKS;
    hipLaunchKernelGGL(k_min_dt, dim3(nb), dim3(256), 0, c->stream, d, c->g, ucont, c->red);
    VFS_CHECK_LAUNCH();
    double *res = c->red + VFS_RED_SLOTS * VFS_RED_BLOCKS;
    if (vfs_reduce_finish(c, nb, 1, 1, res)) return -1;
    double h;
    if (vfs_read_scalars(c, res, &h, 1)) return -1;
    *dt_out = -h < 1.e7 ? -h : 1.e7;
    return 0;
}
