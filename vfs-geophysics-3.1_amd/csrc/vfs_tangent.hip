// Jacobian-vector and vector-Jacobian products of the implicit momentum
// residual F(U) of FormFunction_SNES (momentum.c:2197-2301):
//
//   F(U) = -(Mb U - U_o)/dt + 1/2 FF2(lucont(U), ucat(U)) + 1/2 rhs_o - dP
//
// where Mb zeroes the no-slip boundary fluxes (momentum.c:2224-2232),
// ucat(U) = Contra2Cart_2 (rhs.c:57-741) and lucont(U) = IB_BC (momentum.c:
// 1982-2195) applied to Mb U.  nu_t and the ucat of cells with nvert >= 0.1
// (IB nodes, solids) are frozen inputs, as they are inside the reference's
// SNES solve (Contra2Cart_2 only writes cells with nvert < 0.1, rhs.c:153).
// J = dF/dU is what the reference's MatMFFD approximates by a finite
// difference of FormFunction_SNES per Krylov iteration (implicitsolver.c:
// 4222-4236); here it is exact.
//
// JVP: the linear stages run the forward kernels on the tangent, and the
// face fluxes run in dual numbers (value, tangent).
// VJP: every stage is transposed by hand in reverse order (vfs_adjoint.h).
#include "vfs_adjoint.h"
#include "vfs_internal.h"

int vfs_ff2_from_faces(vfs_ctx *c, double *rhs, double scale);

static int ensure_ad(vfs_ctx *c)
{
    if (c->ad) return 0;
    VFS_CHECK_HIP(hipMalloc(&c->ad, 12 * (size_t)c->d.P * sizeof(double)));
    VFS_CHECK_HIP(hipMemsetAsync(c->ad, 0, 12 * (size_t)c->d.P * sizeof(double), c->stream));
    return 0;
}

// no-slip boundary flux mask of FormFunction_SNES (momentum.c:2224-2232)
__device__ __forceinline__ void noslip_mask(const Dims &d, int i, int j, int k, double u[3])
{
    if ((i == 0 && d.bc[0] == 1) || (i == d.mx - 2 && d.bc[1] == 1)) u[0] = 0;
    if ((j == 0 && d.bc[2] == 1) || (j == d.my - 2 && d.bc[3] == 1)) u[1] = 0;
    if ((k == 0 && d.bc[4] == 1) || (k == d.mz - 2 && d.bc[5] == 1)) u[2] = 0;
}

#define POINT_IJK(t)                                                                     \
    int i = (int)((t) % d.mx), j = (int)(((t) / d.mx) % d.my), k = (int)((t) / ((long long)d.mx * d.my))

// a = Mb U; optionally out = -1/dt * Mb U
__global__ void k_mask_init(Dims d, const double *__restrict__ U, double *__restrict__ a, double *__restrict__ out)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    POINT_IJK(t);
    double u[3] = {U[3 * t], U[3 * t + 1], U[3 * t + 2]};
    noslip_mask(d, i, j, k, u);
    const double dt = d.dt;
#pragma unroll
    for (int q = 0; q < 3; q++) {
        a[3 * t + q] = u[q];
        if (out) {
            double r = 0;
            r += -1. / dt * u[q];
            out[3 * t + q] = r;
        }
    }
}

// primal state of F at U: ad_u = lucont, ad_c = ucat (frozen cells from ucat_in)
static int primal_state(vfs_ctx *c, const double *U, const double *ucat_in, double *lu, double *uc)
{
    const Dims d = c->d;
    hipLaunchKernelGGL(k_mask_init, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d, U, lu, (double *)nullptr);
    VFS_CHECK_LAUNCH();
    VFS_CHECK_HIP(hipMemcpyAsync(uc, ucat_in, 3 * (size_t)d.P * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    if (vfs_contra2cart(c, lu, uc)) return -1;
    if (vfs_ib_bc(c, lu, uc)) return -1;
    return vfs_periodic_vec(c, lu);
}

// ---------------------------------------------------------------------------
// JVP
// ---------------------------------------------------------------------------
__global__ void k_ff2_faces_jvp(Dims d, Geom g, const double *__restrict__ nvert, const double *__restrict__ ucont,
                                const double *__restrict__ ucat, const double *__restrict__ tucont,
                                const double *__restrict__ tucat, const double *__restrict__ nu_t,
                                double *__restrict__ face)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    POINT_IJK(t);
    if (i > d.mx - 2 || j > d.my - 2 || k > d.mz - 2) return;
    const long long P = d.P;
    const DualVec U{ucat, tucat, d};
    Dual D[3], V[3];
    if (j >= 1 && k >= 1) {
        face_flux_t<0, Dual>(d, g, GlobalSca{nvert, d}, UConDual{ucont, tucont, 0}, U, nu_t, i, j, k, D, V);
#pragma unroll
        for (int q = 0; q < 3; q++) { face[(0 * 3 + q) * P + t] = D[q].d; face[(3 * 3 + q) * P + t] = V[q].d; }
    }
    if (i >= 1 && k >= 1) {
        face_flux_t<1, Dual>(d, g, GlobalSca{nvert, d}, UConDual{ucont, tucont, 1}, U, nu_t, i, j, k, D, V);
#pragma unroll
        for (int q = 0; q < 3; q++) { face[(1 * 3 + q) * P + t] = D[q].d; face[(4 * 3 + q) * P + t] = V[q].d; }
    }
    if (i >= 1 && j >= 1) {
        face_flux_t<2, Dual>(d, g, GlobalSca{nvert, d}, UConDual{ucont, tucont, 2}, U, nu_t, i, j, k, D, V);
#pragma unroll
        for (int q = 0; q < 3; q++) { face[(2 * 3 + q) * P + t] = D[q].d; face[(5 * 3 + q) * P + t] = V[q].d; }
    }
}

// Primal state of F at (U, ucat) in the context's workspace, reused by
// vfs_rhs_jvp_prepared until the next call.
int vfs_rhs_prepare(vfs_ctx *c, const double *U, const double *ucat)
{
    if (ensure_ad(c)) return -1;
    const size_t P = (size_t)c->d.P;
    return primal_state(c, U, ucat, c->ad, c->ad + 3 * P);
}

int vfs_rhs_jvp_prepared(vfs_ctx *c, const double *nu_t, const double *w, double *jw)
{
    const Dims d = c->d;
    const size_t P = (size_t)d.P;
    double *lu = c->ad, *uc = c->ad + 3 * P, *tu = c->ad + 6 * P, *tc = c->ad + 9 * P;
    // tangent of Mb U and of the -Mb U/dt term
    hipLaunchKernelGGL(k_mask_init, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d, w, tu, jw);
    VFS_CHECK_LAUNCH();
    // tangent of ucat: frozen cells stay 0
    VFS_CHECK_HIP(hipMemsetAsync(tc, 0, 3 * P * sizeof(double), c->stream));
    if (vfs_contra2cart(c, tu, tc)) return -1;
    if (vfs_ib_bc(c, tu, tc)) return -1;
    if (vfs_periodic_vec(c, tu)) return -1;
    hipLaunchKernelGGL(k_ff2_faces_jvp, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d, c->g, c->nvert, lu, uc,
                       tu, tc, nu_t, c->face);
    VFS_CHECK_LAUNCH();
    return vfs_ff2_from_faces(c, jw, 0.5);
}

// out = Mb in (no-slip boundary fluxes zeroed); in-place allowed
int vfs_noslip_mask(vfs_ctx *c, const double *in, double *out)
{
    hipLaunchKernelGGL(k_mask_init, dim3(grid_for(c->d.P, 256)), dim3(256), 0, c->stream, c->d, in, out,
                       (double *)nullptr);
    VFS_CHECK_LAUNCH();
    return 0;
}

extern "C" int vfs_rhs_jvp(vfs_ctx *c, const double *U, const double *ucat, const double *nu_t, const double *w,
                           double *jw)
{
    VFS_REQUIRE(nu_t || !c->d.les, "nu_t required when les is on");
    VFS_REQUIRE(!vfs_has_wallfn(c->d),
                "the derivatives do not cover the wall model of wall-function walls (bctype -1)");
    VFS_REQUIRE(!c->d.inviscid, "the derivatives do not cover the inviscid scheme (-inv)");
    VFS_REQUIRE(c->d.bc[4] != 5 && c->d.bc[5] != 4, "the derivatives do not cover open k boundaries (bctypes 5 and 4)");
    VFS_REQUIRE(U && ucat && w && jw, "null argument");
    VFS_REQUIRE(w != jw, "w and jw must not alias");
    if (vfs_rhs_prepare(c, U, ucat)) return -1;
    return vfs_rhs_jvp_prepared(c, nu_t, w, jw);
}

// ---------------------------------------------------------------------------
// VJP
// ---------------------------------------------------------------------------
// Transpose of the rhs projection with masks (k_ff2_rhs, momentum.c:1673-1900)
// and of the -Mb U/dt term: out = -vb/dt, fpb += scale * P^T vb.
__global__ void k_vjp_rhs(Dims d, Geom g, const double *__restrict__ nvert, const double *__restrict__ v,
                          double *__restrict__ fpb, double *__restrict__ out, double scale)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    POINT_IJK(t);
    if (i < 1 || j < 1 || k < 1 || i > d.mx - 2 || j > d.my - 2 || k > d.mz - 2) {
        out[3 * t] = out[3 * t + 1] = out[3 * t + 2] = 0;
        return;
    }
    double vb[3] = {v[3 * t], v[3 * t + 1], v[3 * t + 2]};
    const double n0 = nvert[t];
    if (n0 + nvert[t + 1] > 0.1 || (!d.ip && i == d.mx - 2)) vb[0] = 0;
    if (n0 + nvert[t + d.mx] > 0.1 || (!d.jp && j == d.my - 2)) vb[1] = 0;
    if (n0 + nvert[t + (long long)d.mx * d.my] > 0.1 || (!d.kp && k == d.mz - 2)) vb[2] = 0;
    const double dt = d.dt;
#pragma unroll
    for (int q = 0; q < 3; q++) out[3 * t + q] = -1. / dt * vb[q];
    const long long a = IX(d, k, j, wrap_cell(i + 1, d.mx, d.ip)), b = IX(d, k, wrap_cell(j + 1, d.my, d.jp), i),
                    cc = IX(d, wrap_cell(k + 1, d.mz, d.kp), j, i);
    const long long ga = t + 1, gb = t + d.mx, gc = t + (long long)d.mx * d.my;
    const double w0 = scale * 0.5 * g.iaj[t] * vb[0], w1 = scale * 0.5 * g.jaj[t] * vb[1],
                 w2 = scale * 0.5 * g.kaj[t] * vb[2];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        adj_add(fpb + 3 * t + q, g.csi[3 * t + q] * w0 + g.eta[3 * t + q] * w1 + g.zet[3 * t + q] * w2);
        adj_add(fpb + 3 * a + q, g.csi[3 * ga + q] * w0);
        adj_add(fpb + 3 * b + q, g.eta[3 * gb + q] * w1);
        adj_add(fpb + 3 * cc + q, g.zet[3 * gc + q] * w2);
    }
}

// Transpose of cell_fp (momentum.c:1506-1636)
__global__ void k_vjp_cells(Dims d, const double *__restrict__ nvert, const double *__restrict__ fpb,
                            double *__restrict__ face)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    POINT_IJK(t);
    if (i < 1 || j < 1 || k < 1 || i > d.mx - 2 || j > d.my - 2 || k > d.mz - 2) return;
    const long long P = d.P;
    const CellFaces c = cell_faces(d, nvert, i, j, k);
    const double cdv = d.second_order ? 1.0 : 9. / 8.;
#define FA(kind, q, kk, jj, ii, v) adj_add(face + (long long)(3 * (kind) + (q)) * P + IX(d, kk, jj, ii), (v))
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const double fb = fpb[3 * t + q];
        if (fb == 0.0) continue;
        FA(0, q, k, j, i, cdv * fb); FA(0, q, k, j, c.im, -cdv * fb);
        FA(1, q, k, j, i, cdv * fb); FA(1, q, k, c.jm, i, -cdv * fb);
        FA(2, q, k, j, i, cdv * fb); FA(2, q, c.km, j, i, -cdv * fb);
        if (!d.second_order) {
            const double wi = -1. / 8. * (1. / c.den_i) * fb, wj = -1. / 8. * (1. / c.den_j) * fb,
                         wk = -1. / 8. * (1. / c.den_k) * fb;
            FA(0, q, k, j, c.iR, wi); FA(0, q, k, j, c.iL, -wi);
            FA(1, q, k, c.jR, i, wj); FA(1, q, k, c.jL, i, -wj);
            FA(2, q, c.kR, j, i, wk); FA(2, q, c.kL, j, i, -wk);
        }
        FA(3, q, k, j, i, fb); FA(3, q, k, j, c.im, -fb);
        FA(4, q, k, j, i, fb); FA(4, q, k, c.jm, i, -fb);
        FA(5, q, k, j, i, fb); FA(5, q, c.km, j, i, -fb);
    }
#undef FA
}

// Transpose of the face fluxes (same face set as k_ff2_faces)
__global__ void k_vjp_faces(Dims d, Geom g, const double *__restrict__ nvert, const double *__restrict__ ucont,
                            const double *__restrict__ ucat, const double *__restrict__ nu_t,
                            const double *__restrict__ face, double *ucont_b, double *ucat_b)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    POINT_IJK(t);
    if (i > d.mx - 2 || j > d.my - 2 || k > d.mz - 2) return;
    const long long P = d.P;
    double Db[3], Vb[3];
    if (j >= 1 && k >= 1) {
        for (int q = 0; q < 3; q++) { Db[q] = face[(0 * 3 + q) * P + t]; Vb[q] = face[(3 * 3 + q) * P + t]; }
        face_flux_adj<0>(d, g, nvert, ucont, ucat, nu_t, i, j, k, Db, Vb, ucont_b, ucat_b);
    }
    if (i >= 1 && k >= 1) {
        for (int q = 0; q < 3; q++) { Db[q] = face[(1 * 3 + q) * P + t]; Vb[q] = face[(4 * 3 + q) * P + t]; }
        face_flux_adj<1>(d, g, nvert, ucont, ucat, nu_t, i, j, k, Db, Vb, ucont_b, ucat_b);
    }
    if (i >= 1 && j >= 1) {
        for (int q = 0; q < 3; q++) { Db[q] = face[(2 * 3 + q) * P + t]; Vb[q] = face[(5 * 3 + q) * P + t]; }
        face_flux_adj<2>(d, g, nvert, ucont, ucat, nu_t, i, j, k, Db, Vb, ucont_b, ucat_b);
    }
}

// Transpose of the periodic ghost copy of a vector field (k_periodic_vec):
// ghost adjoints move onto their sources; sources are never ghosts.
__global__ void k_periodic_vec_T(Dims d, double *v)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int k, j, i;
    if (t >= shell_count(d) || !shell_point(d, t, k, j, i)) return;
    int a, b, c;
    if (!per_src(d, i, j, k, a, b, c)) return;
    long long s = IX(d, c, b, a), o = IX(d, k, j, i);
#pragma unroll
    for (int q = 0; q < 3; q++) {
        adj_add(v + 3 * s + q, v[3 * o + q]);
        v[3 * o + q] = 0;
    }
}

static int periodic_vec_T(vfs_ctx *c, double *v)
{
    if (!(c->d.ip || c->d.jp || c->d.kp)) return 0;
    long long n = shell_count(c->d);
    hipLaunchKernelGGL(k_periodic_vec_T, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, c->d, v);
    VFS_CHECK_LAUNCH();
    return 0;
}

// Transpose of the periodic face copies of IB_BC (momentum.c:2167-2181)
__global__ void k_periodic_faces_T(Dims d, double *u)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int k, j, i;
    if (t >= shell_count(d) || !shell_point(d, t, k, j, i)) return;
    long long o = IX(d, k, j, i);
#define MOVE(q, src)                                                  \
    {                                                                 \
        adj_add(u + 3 * (src) + (q), u[3 * o + (q)]);                 \
        u[3 * o + (q)] = 0;                                           \
    }
    if (d.ip && i == 0) MOVE(0, IX(d, k, j, d.mx - 2));
    if (d.ip && i == d.mx - 1) MOVE(0, IX(d, k, j, 1));
    if (d.jp && j == 0) MOVE(1, IX(d, k, d.my - 2, i));
    if (d.jp && j == d.my - 1) MOVE(1, IX(d, k, 1, i));
    if (d.kp && k == 0) MOVE(2, IX(d, d.mz - 2, j, i));
    if (d.kp && k == d.mz - 1) MOVE(2, IX(d, 1, j, i));
#undef MOVE
}

// Transpose of the IB face rule of IB_BC (momentum.c:2088-2124)
__global__ void k_ibbc_T(Dims d, Geom g, const double *__restrict__ nvert, double *ucont_b, double *ucat_b)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    POINT_IJK(t);
    if (i < 1 || j < 1 || k < 1 || i > d.mx - 2 || j > d.my - 2 || k > d.mz - 2) return;
    const long long sj = d.mx, sk = (long long)d.mx * d.my;
    const bool self = (int)nvert[t] == 1;
    const long long nb[3] = {t + 1, t + sj, t + sk};
    const double *mt[3] = {g.icsi + 3 * t, g.jeta + 3 * t, g.kzet + 3 * t};
#pragma unroll
    for (int dir = 0; dir < 3; dir++) {
        if (!(self || (int)(nvert[nb[dir]]) == 1)) continue;
        const double val = ucont_b[3 * t + dir];
        ucont_b[3 * t + dir] = 0;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const double w = 0.5 * mt[dir][q] * val;
            adj_add(ucat_b + 3 * t + q, w);
            adj_add(ucat_b + 3 * nb[dir] + q, w);
        }
    }
}

// Transpose of k_c2c_bc_interior: zero the adjoint where ucat was zeroed
__global__ void k_c2c_bc_interior_T(Dims d, const double *__restrict__ nvert, double *ub)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    POINT_IJK(t);
    if (i < 1 || j < 1 || k < 1 || i > d.mx - 2 || j > d.my - 2 || k > d.mz - 2) return;
    bool z = (int)(nvert[t] + 0.1) == 3;
    if (i != 0 && i != d.mx - 1 && j != 0 && j != d.my - 2 && k != 0 && k != d.mz - 2) {
        if (d.bc[0] <= 1 && d.bc[2] <= 1 && (i == 1 && j == 1)) z = true;
        if (d.bc[1] <= 1 && d.bc[2] <= 1 && (i == d.mx - 2 && j == 1)) z = true;
        if (d.bc[0] <= 1 && d.bc[3] <= 1 && (i == 1 && j == d.my - 2)) z = true;
        if (d.bc[1] <= 1 && d.bc[3] <= 1 && (i == d.mx - 2 && j == d.my - 2)) z = true;
    }
    if (z) ub[3 * t] = ub[3 * t + 1] = ub[3 * t + 2] = 0;
}

// ghost rule of k_c2c_bc_ghost: -1 solid ghost (zeroed), >= 0 mirror source, -2 untouched
__device__ __forceinline__ long long c2c_ghost_rule(const Dims &d, const double *__restrict__ nvert, int i, int j, int k)
{
    long long id = IX(d, k, j, i);
    if ((int)(nvert[id] + 0.1) == 3) return -1;
    long long s = -2;
    if (i == 0 && d.bc[0] == 1 && (j != 0 && k != 0)) s = IX(d, k, j, i + 1);
    if (i == d.mx - 1 && d.bc[1] == 1 && (j != 0 && k != 0)) s = IX(d, k, j, i - 1);
    if (j == 0 && d.bc[2] == 1 && (i != 0 && k != 0)) s = IX(d, k, j + 1, i);
    if (j == d.my - 1 && d.bc[3] == 1 && (i != 0 && k != 0)) s = IX(d, k, j - 1, i);
    if (k == 0 && d.bc[4] == 1 && (i != 0 && j != 0)) s = IX(d, k + 1, j, i);
    if (k == d.mz - 1 && d.bc[5] == 1 && (i != 0 && j != 0)) s = IX(d, k - 1, j, i);
    return s;
}

__global__ void k_c2c_bc_ghost_T_zero(Dims d, const double *__restrict__ nvert, double *ub)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int k, j, i;
    if (t >= shell_count(d) || !shell_point(d, t, k, j, i)) return;
    if (c2c_ghost_rule(d, nvert, i, j, k) == -2) return;
    long long id = IX(d, k, j, i);
    ub[3 * id] = ub[3 * id + 1] = ub[3 * id + 2] = 0;
}

__global__ void k_c2c_bc_ghost_T_add(Dims d, const double *__restrict__ nvert, const double *__restrict__ snap, double *ub)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int k, j, i;
    if (t >= shell_count(d) || !shell_point(d, t, k, j, i)) return;
    long long s = c2c_ghost_rule(d, nvert, i, j, k);
    if (s < 0) return;
    long long id = IX(d, k, j, i);
#pragma unroll
    for (int q = 0; q < 3; q++) adj_add(ub + 3 * s + q, -1 * snap[3 * id + q]);
}

// Transpose of the Cramer solve of k_c2c_interior (rhs.c:203-231)
__global__ void k_c2c_interior_T(Dims d, Geom g, const double *__restrict__ nvert, const double *__restrict__ ub,
                                 double *cb)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    POINT_IJK(t);
    if (i < 1 || j < 1 || k < 1 || i > d.mx - 2 || j > d.my - 2 || k > d.mz - 2) return;
    if (!(nvert[t] < 0.1)) return;
    const double u0 = ub[3 * t], u1 = ub[3 * t + 1], u2 = ub[3 * t + 2];
    if (u0 == 0.0 && u1 == 0.0 && u2 == 0.0) return;
    double m00 = g.csi[3 * t], m01 = g.csi[3 * t + 1], m02 = g.csi[3 * t + 2];
    double m10 = g.eta[3 * t], m11 = g.eta[3 * t + 1], m12 = g.eta[3 * t + 2];
    double m20 = g.zet[3 * t], m21 = g.zet[3 * t + 1], m22 = g.zet[3 * t + 2];
    double det = m00 * (m11 * m22 - m12 * m21) - m01 * (m10 * m22 - m12 * m20) + m02 * (m10 * m21 - m11 * m20);
    // ucat_r = sum_c inv[r][c] q_c, so qb_c = sum_r inv[r][c] ub_r
    const double i00 = m11 * m22 - m12 * m21, i01 = -(m01 * m22 - m02 * m21), i02 = m01 * m12 - m02 * m11;
    const double i10 = -(m10 * m22 - m12 * m20), i11 = m00 * m22 - m02 * m20, i12 = -(m00 * m12 - m02 * m10);
    const double i20 = m10 * m21 - m11 * m20, i21 = -(m00 * m21 - m01 * m20), i22 = m00 * m11 - m01 * m10;
    const double q0 = (i00 * u0 + i10 * u1 + i20 * u2) / det;
    const double q1 = (i01 * u0 + i11 * u1 + i21 * u2) / det;
    const double q2 = (i02 * u0 + i12 * u1 + i22 * u2) / det;
    adj_add(cb + 3 * (t - 1), 0.5 * q0);
    adj_add(cb + 3 * t, 0.5 * q0);
    adj_add(cb + 3 * (t - d.mx) + 1, 0.5 * q1);
    adj_add(cb + 3 * t + 1, 0.5 * q1);
    adj_add(cb + 3 * (t - (long long)d.mx * d.my) + 2, 0.5 * q2);
    adj_add(cb + 3 * t + 2, 0.5 * q2);
}

// out = Mb (out + cb)
__global__ void k_vjp_final(Dims d, const double *__restrict__ cb, double *__restrict__ out)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    POINT_IJK(t);
    double u[3] = {out[3 * t] + cb[3 * t], out[3 * t + 1] + cb[3 * t + 1], out[3 * t + 2] + cb[3 * t + 2]};
    noslip_mask(d, i, j, k, u);
    out[3 * t] = u[0]; out[3 * t + 1] = u[1]; out[3 * t + 2] = u[2];
}

extern "C" int vfs_rhs_vjp(vfs_ctx *c, const double *U, const double *ucat, const double *nu_t, const double *v,
                           double *jtv)
{
    const Dims d = c->d;
    VFS_REQUIRE(nu_t || !d.les, "nu_t required when les is on");
    VFS_REQUIRE(!vfs_has_wallfn(d),
                "the derivatives do not cover the wall model of wall-function walls (bctype -1)");
    VFS_REQUIRE(!d.inviscid, "the derivatives do not cover the inviscid scheme (-inv)");
    VFS_REQUIRE(d.bc[4] != 5 && d.bc[5] != 4, "the derivatives do not cover open k boundaries (bctypes 5 and 4)");
    VFS_REQUIRE(U && ucat && v && jtv, "null argument");
    VFS_REQUIRE(v != jtv, "v and jtv must not alias");
    if (ensure_ad(c)) return -1;
    const size_t P = (size_t)d.P;
    double *lu = c->ad, *uc = c->ad + 3 * P, *cb = c->ad + 6 * P, *ub = c->ad + 9 * P;
    if (primal_state(c, U, ucat, lu, uc)) return -1;
    const dim3 gp(grid_for(d.P, 256)), bk(256);
    const unsigned gs = grid_for(shell_count(d), 256);
    VFS_CHECK_HIP(hipMemsetAsync(c->fp, 0, 3 * P * sizeof(double), c->stream));
    VFS_CHECK_HIP(hipMemsetAsync(c->face, 0, 18 * P * sizeof(double), c->stream));
    VFS_CHECK_HIP(hipMemsetAsync(cb, 0, 6 * P * sizeof(double), c->stream));   // cb and ub
    hipLaunchKernelGGL(k_vjp_rhs, gp, bk, 0, c->stream, d, c->g, c->nvert, v, c->fp, jtv, 0.5);
    VFS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_vjp_cells, gp, bk, 0, c->stream, d, c->nvert, c->fp, c->face);
    VFS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_vjp_faces, gp, bk, 0, c->stream, d, c->g, c->nvert, lu, uc, nu_t, c->face, cb, ub);
    VFS_CHECK_LAUNCH();
    // periodic copy of lucont at the start of Formfunction_2
    if (periodic_vec_T(c, cb)) return -1;
    // IB_BC: periodic face copies, then the IB face rule
    if (d.ip || d.jp || d.kp) {
        hipLaunchKernelGGL(k_periodic_faces_T, dim3(gs), bk, 0, c->stream, d, cb);
        VFS_CHECK_LAUNCH();
    }
    if (d.immersed) {
        hipLaunchKernelGGL(k_ibbc_T, gp, bk, 0, c->stream, d, c->g, c->nvert, cb, ub);
        VFS_CHECK_LAUNCH();
    }
    // Contra2Cart_2 in reverse: periodic copy, interior zeroing, ghost rules,
    // periodic copy, Cramer solve, periodic copy of ucont
    if (periodic_vec_T(c, ub)) return -1;
    hipLaunchKernelGGL(k_c2c_bc_interior_T, gp, bk, 0, c->stream, d, c->nvert, ub);
    VFS_CHECK_LAUNCH();
    VFS_CHECK_HIP(hipMemcpyAsync(c->snap, ub, 3 * P * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    hipLaunchKernelGGL(k_c2c_bc_ghost_T_zero, dim3(gs), bk, 0, c->stream, d, c->nvert, ub);
    VFS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_c2c_bc_ghost_T_add, dim3(gs), bk, 0, c->stream, d, c->nvert, c->snap, ub);
    VFS_CHECK_LAUNCH();
    if (periodic_vec_T(c, ub)) return -1;
    hipLaunchKernelGGL(k_c2c_interior_T, gp, bk, 0, c->stream, d, c->g, c->nvert, ub, cb);
    VFS_CHECK_LAUNCH();
    if (periodic_vec_T(c, cb)) return -1;
    hipLaunchKernelGGL(k_vjp_final, gp, bk, 0, c->stream, d, cb, jtv);
    VFS_CHECK_LAUNCH();
    return 0;
}
