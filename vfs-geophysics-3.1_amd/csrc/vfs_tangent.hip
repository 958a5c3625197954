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
// VJP: every stage is transposed by hand in reverse order (vfs_adjoint.h),
// each as a gather with a fixed summation order.
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
    long long t = (long long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
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
__global__ void __launch_bounds__(256) k_ff2_faces_jvp(Dims d, Geom g, const double *__restrict__ nvert, const double *__restrict__ ucont,
                                const double *__restrict__ ucat, const double *__restrict__ tucont,
                                const double *__restrict__ tucat, const double *__restrict__ nu_t,
                                double *__restrict__ face)
{
    long long t = (long long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
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
    VFS_REQUIRE(!c->d.skew && !c->d.clark, "the derivatives do not cover -skew or -clark");
    VFS_REQUIRE(c->d.bc[4] != 5 && c->d.bc[5] != 4, "the derivatives do not cover open k boundaries (bctypes 5 and 4)");
    VFS_REQUIRE(U && ucat && w && jw, "null argument");
    VFS_REQUIRE(w != jw, "w and jw must not alias");
    if (vfs_rhs_prepare(c, U, ucat)) return -1;
    return vfs_rhs_jvp_prepared(c, nu_t, w, jw);
}

// ---------------------------------------------------------------------------
// VJP
// ---------------------------------------------------------------------------
// Every transpose is a gather: each adjoint element sums, in a fixed order,
// the contributions of the forward outputs that read it, so no two threads
// write one element, no atomics are used and J^T v is the same bits on every
// run.

// v at cell t, component dir, with the face masks of the projection
// (momentum.c:1795-1803)
__device__ __forceinline__ double vb_masked(const Dims &d, const double *__restrict__ nvert,
                                            const double *__restrict__ v, long long t, int i, int j, int k, int dir)
{
    const long long s = dir == 0 ? 1 : dir == 1 ? (long long)d.mx : (long long)d.mx * d.my;
    const int c = dir == 0 ? i : dir == 1 ? j : k;
    const int m = dir == 0 ? d.mx : dir == 1 ? d.my : d.mz;
    const int per = dir == 0 ? d.ip : dir == 1 ? d.jp : d.kp;
    if (nvert[t] + nvert[t + s] > 0.1 || (!per && c == m - 2)) return 0.0;
    return v[3 * t + dir];
}

// Transpose of the rhs projection with masks (k_ff2_rhs, momentum.c:
// 1673-1900) and of the -Mb U/dt term: out = -vb/dt and fpb = scale P^T vb.
// A cell's fp is read by its own three faces and by the i / j / k face of the
// cell before it (across a periodic seam, the last cell), with that cell's
// metric at its +1 neighbour.
__global__ void k_vjp_rhs(Dims d, Geom g, const double *__restrict__ nvert, const double *__restrict__ v,
                          double *__restrict__ fpb, double *__restrict__ out, double scale)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    POINT_IJK(t);
    if (i < 1 || j < 1 || k < 1 || i > d.mx - 2 || j > d.my - 2 || k > d.mz - 2) {
#pragma unroll
        for (int q = 0; q < 3; q++) out[3 * t + q] = fpb[3 * t + q] = 0;
        return;
    }
    double vb[3];
#pragma unroll
    for (int dir = 0; dir < 3; dir++) vb[dir] = vb_masked(d, nvert, v, t, i, j, k, dir);
    const double dt = d.dt;
#pragma unroll
    for (int q = 0; q < 3; q++) out[3 * t + q] = -1. / dt * vb[q];
    const double w0 = scale * 0.5 * g.iaj[t] * vb[0], w1 = scale * 0.5 * g.jaj[t] * vb[1],
                 w2 = scale * 0.5 * g.kaj[t] * vb[2];
    double f[3];
#pragma unroll
    for (int q = 0; q < 3; q++) f[q] = g.csi[3 * t + q] * w0 + g.eta[3 * t + q] * w1 + g.zet[3 * t + q] * w2;
    const int ib = i > 1 ? i - 1 : (d.ip ? d.mx - 2 : 0);
    if (ib >= 1) {
        const long long c = IX(d, k, j, ib);
        const double w = scale * 0.5 * g.iaj[c] * vb_masked(d, nvert, v, c, ib, j, k, 0);
#pragma unroll
        for (int q = 0; q < 3; q++) f[q] += g.csi[3 * (c + 1) + q] * w;
    }
    const int jb = j > 1 ? j - 1 : (d.jp ? d.my - 2 : 0);
    if (jb >= 1) {
        const long long c = IX(d, k, jb, i);
        const double w = scale * 0.5 * g.jaj[c] * vb_masked(d, nvert, v, c, i, jb, k, 1);
#pragma unroll
        for (int q = 0; q < 3; q++) f[q] += g.eta[3 * (c + d.mx) + q] * w;
    }
    const int kb = k > 1 ? k - 1 : (d.kp ? d.mz - 2 : 0);
    if (kb >= 1) {
        const long long c = IX(d, kb, j, i);
        const double w = scale * 0.5 * g.kaj[c] * vb_masked(d, nvert, v, c, i, j, kb, 2);
#pragma unroll
        for (int q = 0; q < 3; q++) f[q] += g.zet[3 * (c + (long long)d.mx * d.my) + q] * w;
    }
#pragma unroll
    for (int q = 0; q < 3; q++) fpb[3 * t + q] = f[q];
}

// Transpose of cell_fp (momentum.c:1506-1636): the adjoint of each face flux
// from the cells whose divergence reads it.  Along its direction a face f is
// the own face of cell f, the upwind face of cell f+1 and a far face of the
// 4th-order correction of cells f-1 .. f+2 (periodic-wrapped).
__global__ void k_vjp_cells(Dims d, const double *__restrict__ nvert, const double *__restrict__ fpb,
                            double *__restrict__ face)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    POINT_IJK(t);
    const long long P = d.P;
    const double cdv = d.second_order ? 1.0 : 9. / 8.;
#pragma unroll
    for (int dir = 0; dir < 3; dir++) {
        double D[3] = {0, 0, 0}, Vv[3] = {0, 0, 0};
        const int f = dir == 0 ? i : dir == 1 ? j : k;
        const int m = dir == 0 ? d.mx : dir == 1 ? d.my : d.mz;
        const int per = dir == 0 ? d.ip : dir == 1 ? d.jp : d.kp;
        const bool row = (dir == 0 || (i >= 1 && i <= d.mx - 2)) && (dir == 1 || (j >= 1 && j <= d.my - 2)) &&
                         (dir == 2 || (k >= 1 && k <= d.mz - 2));
        if (row && f <= m - 2) {
            for (int o = -1; o <= 2; o++) {
                int cn = f + o;
                if (cn < 1 || cn > m - 2) {
                    if (!per) continue;
                    cn = cn < 1 ? cn + (m - 2) : cn - (m - 2);
                }
                const int ci = dir == 0 ? cn : i, cj = dir == 1 ? cn : j, ck = dir == 2 ? cn : k;
                const CellFaces cf = cell_faces(d, nvert, ci, cj, ck);
                const int im = dir == 0 ? cf.im : dir == 1 ? cf.jm : cf.km;
                const int fR = dir == 0 ? cf.iR : dir == 1 ? cf.jR : cf.kR;
                const int fL = dir == 0 ? cf.iL : dir == 1 ? cf.jL : cf.kL;
                const double den = dir == 0 ? cf.den_i : dir == 1 ? cf.den_j : cf.den_k;
                const long long c = IX(d, ck, cj, ci);
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    const double fb = fpb[3 * c + q];
                    if (cn == f) { D[q] += cdv * fb; Vv[q] += fb; }
                    if (im == f) { D[q] += -cdv * fb; Vv[q] += -fb; }
                    if (!d.second_order) {
                        const double w = -1. / 8. * (1. / den) * fb;
                        if (fR == f) D[q] += w;
                        if (fL == f) D[q] += -w;
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 3; q++) {
            face[(long long)(3 * dir + q) * P + t] = D[q];
            face[(long long)(3 * (3 + dir) + q) * P + t] = Vv[q];
        }
    }
}

// The adjoints of the DIR faces by stencil position (FaceAdj, slot-major:
// ws[s * P + t]), zero at points that hold no such face.  The faces are
// those of k_ff2_faces.
template <int DIR>
__global__ void k_vjp_faces(Dims d, Geom g, const double *__restrict__ nvert, const double *__restrict__ ucont,
                            const double *__restrict__ ucat, const double *__restrict__ nu_t,
                            const double *__restrict__ face, double *__restrict__ ws)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    POINT_IJK(t);
    const long long P = d.P;
    FaceAdj F;
    const bool isface = i <= d.mx - 2 && j <= d.my - 2 && k <= d.mz - 2 && (DIR == 0 || i >= 1) &&
                        (DIR == 1 || j >= 1) && (DIR == 2 || k >= 1);
    double Db[3] = {0, 0, 0}, Vb[3] = {0, 0, 0};
    bool any = false;
    if (isface) {
#pragma unroll
        for (int q = 0; q < 3; q++) {
            Db[q] = face[(long long)(DIR * 3 + q) * P + t];
            Vb[q] = face[(long long)((3 + DIR) * 3 + q) * P + t];
            any = any || Db[q] != 0.0 || Vb[q] != 0.0;
        }
    }
    if (any) {
        face_flux_adj<DIR>(d, g, nvert, ucont, ucat, nu_t, i, j, k, Db, Vb, F);
    } else {
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int q = 0; q < 3; q++) F.S[r][q] = 0;
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int s = 0; s < 2; s++)
#pragma unroll
                for (int q = 0; q < 3; q++) F.X[a][s][q] = 0;
        F.uself = F.usrc = 0;
    }
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int q = 0; q < 3; q++) ws[(long long)(3 * r + q) * P + t] = F.S[r][q];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int s = 0; s < 2; s++)
#pragma unroll
            for (int q = 0; q < 3; q++) ws[(long long)(12 + (2 * a + s) * 3 + q) * P + t] = F.X[a][s][q];
    ws[24 * P + t] = F.uself;
    ws[25 * P + t] = F.usrc;
}

// ucat_b (+)= the DIR-face adjoints at this point, cb[DIR] = the ucont
// adjoint of the DIR face here.  Order: the along-DIR slots of the faces
// x+1, x, x-1, x-2 (x: this point's DIR index), the wrapped slots of a
// periodic direction, then per transverse direction the faces of the row
// above and below, each at positions n and n+1.
template <int DIR>
__global__ void k_vjp_gather(Dims d, const double *__restrict__ ws, double *__restrict__ ub,
                             double *__restrict__ cb, int accumulate)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    POINT_IJK(t);
    const long long P = d.P;
    const int x = DIR == 0 ? i : DIR == 1 ? j : k;
    const int m = DIR == 0 ? d.mx : DIR == 1 ? d.my : d.mz;
    const int per = DIR == 0 ? d.ip : DIR == 1 ? d.jp : d.kp;
    const long long st = DIR == 0 ? 1 : DIR == 1 ? (long long)d.mx : (long long)d.mx * d.my;
    double acc[3];
#pragma unroll
    for (int q = 0; q < 3; q++) acc[q] = accumulate ? ub[3 * t + q] : 0.0;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int n = x + 1 - r;
        if (n < 0 || n > m - 1) continue;
        const long long f = t + (long long)(n - x) * st;
#pragma unroll
        for (int q = 0; q < 3; q++) acc[q] += ws[(long long)(3 * r + q) * P + f];
    }
    if (per && x == m - 3) {
        const long long f = t - (long long)x * st;   // face 0: its slot 0 is cell m-3
#pragma unroll
        for (int q = 0; q < 3; q++) acc[q] += ws[(long long)q * P + f];
    }
    if (per && x == 2) {
        const long long f = t + (long long)(m - 2 - x) * st;   // face m-2: its slot 3 is cell 2
#pragma unroll
        for (int q = 0; q < 3; q++) acc[q] += ws[(long long)(9 + q) * P + f];
    }
#pragma unroll
    for (int tt = 0; tt < 2; tt++) {
        const int td = DIR == 0 ? 1 + tt : DIR == 1 ? 2 * tt : tt;
        const int y = td == 0 ? i : td == 1 ? j : k;
        const int my_ = td == 0 ? d.mx : td == 1 ? d.my : d.mz;
        const long long sy = td == 0 ? 1 : td == 1 ? (long long)d.mx : (long long)d.mx * d.my;
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const int yf = s == 0 ? y + 1 : y - 1;   // the face row this point is the -/+ neighbour of
            if (yf < 0 || yf > my_ - 1) continue;
#pragma unroll
            for (int sn = 0; sn < 2; sn++) {
                const int n = x - sn;
                if (n < 0) continue;
                const long long f = t + (long long)(yf - y) * sy - (long long)sn * st;
#pragma unroll
                for (int q = 0; q < 3; q++) acc[q] += ws[(long long)(12 + (2 * tt + s) * 3 + q) * P + f];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 3; q++) ub[3 * t + q] = acc[q];
    double u = ws[24 * P + t];
    if (per && x == m - 2) u += ws[25 * P + (t - (long long)x * st)];
    cb[3 * t + DIR] = u;
}

// Transpose of the periodic ghost copy of a vector field (k_periodic_vec)
// along one direction: each ghost plane's adjoint moves onto the plane it
// copies.  The copy of k_periodic_vec equals the copies along i, then j, then
// k, so its transpose is these moves along k, then j, then i; each move has
// one source per destination.
__global__ void k_periodic_vec_T_dir(Dims d, int dir, double *v)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long A = dir == 0 ? (long long)d.my * d.mz : dir == 1 ? (long long)d.mx * d.mz : (long long)d.mx * d.my;
    if (t >= 2 * A) return;
    const bool hi = t >= A;
    const long long r = hi ? t - A : t;
    int i, j, k;
    if (dir == 0) { j = (int)(r % d.my); k = (int)(r / d.my); i = hi ? d.mx - 1 : 0; }
    else if (dir == 1) { i = (int)(r % d.mx); k = (int)(r / d.mx); j = hi ? d.my - 1 : 0; }
    else { i = (int)(r % d.mx); j = (int)(r / d.mx); k = hi ? d.mz - 1 : 0; }
    const int src = hi ? 1 : (dir == 0 ? d.mx : dir == 1 ? d.my : d.mz) - 2;
    const long long o = IX(d, k, j, i);
    const long long s = dir == 0 ? IX(d, k, j, src) : dir == 1 ? IX(d, k, src, i) : IX(d, src, j, i);
#pragma unroll
    for (int q = 0; q < 3; q++) {
        v[3 * s + q] += v[3 * o + q];
        v[3 * o + q] = 0;
    }
}

static int periodic_vec_T(vfs_ctx *c, double *v)
{
    const Dims d = c->d;
    const int per[3] = {d.ip, d.jp, d.kp};
    for (int dir = 2; dir >= 0; dir--) {
        if (!per[dir]) continue;
        const long long A = dir == 0 ? (long long)d.my * d.mz : dir == 1 ? (long long)d.mx * d.mz : (long long)d.mx * d.my;
        hipLaunchKernelGGL(k_periodic_vec_T_dir, dim3(grid_for(2 * A, 256)), dim3(256), 0, c->stream, d, dir, v);
        VFS_CHECK_LAUNCH();
    }
    return 0;
}

// Transpose of the periodic face copies of IB_BC (momentum.c:2167-2181); each
// face source has one ghost copy
__global__ void k_periodic_faces_T(Dims d, double *u)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int k, j, i;
    if (t >= shell_count(d) || !shell_point(d, t, k, j, i)) return;
    long long o = IX(d, k, j, i);
#define MOVE(q, src)                                                  \
    {                                                                 \
        u[3 * (src) + (q)] += u[3 * o + (q)];                         \
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

// Transpose of the IB face rule of IB_BC (momentum.c:2088-2124), in two
// passes: w = the ucont adjoint of the IB faces (then zeroed there); ucat_b
// gathers 0.5 * metric * w of its own faces and of those of its -1
// neighbours.
__global__ void k_ibbc_T_take(Dims d, const double *__restrict__ nvert, double *cb, double *__restrict__ w)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    POINT_IJK(t);
    const bool in = i >= 1 && j >= 1 && k >= 1 && i <= d.mx - 2 && j <= d.my - 2 && k <= d.mz - 2;
    const long long nb[3] = {t + 1, t + d.mx, t + (long long)d.mx * d.my};
    const bool self = in && (int)nvert[t] == 1;
#pragma unroll
    for (int dir = 0; dir < 3; dir++) {
        double val = 0;
        if (in && (self || (int)(nvert[nb[dir]]) == 1)) {
            val = cb[3 * t + dir];
            cb[3 * t + dir] = 0;
        }
        w[3 * t + dir] = val;
    }
}

__global__ void k_ibbc_T_give(Dims d, Geom g, const double *__restrict__ w, double *__restrict__ ub)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    POINT_IJK(t);
    const long long sd[3] = {1, d.mx, (long long)d.mx * d.my};
    const int cd[3] = {i, j, k};
    const double *mt[3] = {g.icsi, g.jeta, g.kzet};
    double acc[3] = {ub[3 * t], ub[3 * t + 1], ub[3 * t + 2]};
#pragma unroll
    for (int dir = 0; dir < 3; dir++) {
        const double val = w[3 * t + dir];
#pragma unroll
        for (int q = 0; q < 3; q++) acc[q] += 0.5 * mt[dir][3 * t + q] * val;
    }
#pragma unroll
    for (int dir = 0; dir < 3; dir++) {
        if (cd[dir] < 1) continue;
        const long long c = t - sd[dir];
        const double val = w[3 * c + dir];
#pragma unroll
        for (int q = 0; q < 3; q++) acc[q] += 0.5 * mt[dir][3 * c + q] * val;
    }
#pragma unroll
    for (int q = 0; q < 3; q++) ub[3 * t + q] = acc[q];
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

// ub at a mirror source += -1 * (the adjoint of each ghost mirroring it, from
// the snapshot), the ghosts taken in the order -i, +i, -j, +j, -k, +k
__global__ void k_c2c_bc_ghost_T_gather(Dims d, const double *__restrict__ nvert, const double *__restrict__ snap,
                                        double *__restrict__ ub)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    POINT_IJK(t);
    double acc[3] = {0, 0, 0};
    bool any = false;
#pragma unroll
    for (int e = 0; e < 6; e++) {
        const int gi = i + (e == 0 ? -1 : e == 1 ? 1 : 0), gj = j + (e == 2 ? -1 : e == 3 ? 1 : 0),
                  gk = k + (e == 4 ? -1 : e == 5 ? 1 : 0);
        if (gi < 0 || gj < 0 || gk < 0 || gi > d.mx - 1 || gj > d.my - 1 || gk > d.mz - 1) continue;
        if (gi > 0 && gj > 0 && gk > 0 && gi < d.mx - 1 && gj < d.my - 1 && gk < d.mz - 1) continue;
        if (c2c_ghost_rule(d, nvert, gi, gj, gk) != t) continue;
        const long long g = IX(d, gk, gj, gi);
        if (!any) {
#pragma unroll
            for (int q = 0; q < 3; q++) acc[q] = ub[3 * t + q];
            any = true;
        }
#pragma unroll
        for (int q = 0; q < 3; q++) acc[q] += -1 * snap[3 * g + q];
    }
    if (any) {
#pragma unroll
        for (int q = 0; q < 3; q++) ub[3 * t + q] = acc[q];
    }
}

// Transpose of the Cramer solve of k_c2c_interior (rhs.c:203-231), in two
// passes: qb = M^-T ub at the fluid cells; the ucont adjoint of each face
// gathers 0.5 qb of the two cells that average it.
__global__ void k_c2c_interior_T_q(Dims d, Geom g, const double *__restrict__ nvert, const double *__restrict__ ub,
                                   double *__restrict__ qb)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    POINT_IJK(t);
    qb[3 * t] = qb[3 * t + 1] = qb[3 * t + 2] = 0;
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
    qb[3 * t] = (i00 * u0 + i10 * u1 + i20 * u2) / det;
    qb[3 * t + 1] = (i01 * u0 + i11 * u1 + i21 * u2) / det;
    qb[3 * t + 2] = (i02 * u0 + i12 * u1 + i22 * u2) / det;
}

__global__ void k_c2c_interior_T_faces(Dims d, const double *__restrict__ qb, double *__restrict__ cb)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    POINT_IJK(t);
    const long long sd[3] = {1, d.mx, (long long)d.mx * d.my};
    const int cd[3] = {i, j, k}, md[3] = {d.mx, d.my, d.mz};
#pragma unroll
    for (int dir = 0; dir < 3; dir++) {
        double a = cb[3 * t + dir];
        a += 0.5 * qb[3 * t + dir];
        if (cd[dir] + 1 <= md[dir] - 1) a += 0.5 * qb[3 * (t + sd[dir]) + dir];
        cb[3 * t + dir] = a;
    }
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
    VFS_REQUIRE(!d.skew && !d.clark, "the derivatives do not cover -skew or -clark");
    VFS_REQUIRE(d.bc[4] != 5 && d.bc[5] != 4, "the derivatives do not cover open k boundaries (bctypes 5 and 4)");
    VFS_REQUIRE(U && ucat && v && jtv, "null argument");
    VFS_REQUIRE(v != jtv, "v and jtv must not alias");
    if (ensure_ad(c)) return -1;
    const size_t P = (size_t)d.P;
    if (!c->vjp_ws) VFS_CHECK_HIP(hipMalloc(&c->vjp_ws, FACE_ADJ_SLOTS * P * sizeof(double)));
    double *lu = c->ad, *uc = c->ad + 3 * P, *cb = c->ad + 6 * P, *ub = c->ad + 9 * P, *ws = c->vjp_ws;
    if (primal_state(c, U, ucat, lu, uc)) return -1;
    const dim3 gp(grid_for(d.P, 256)), bk(256);
    const unsigned gs = grid_for(shell_count(d), 256);
    // projection and divergence (every element written)
    hipLaunchKernelGGL(k_vjp_rhs, gp, bk, 0, c->stream, d, c->g, c->nvert, v, c->fp, jtv, 0.5);
    VFS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_vjp_cells, gp, bk, 0, c->stream, d, c->nvert, c->fp, c->face);
    VFS_CHECK_LAUNCH();
    // face fluxes, one direction at a time: ub = sum over i, j, k faces,
    // cb[dir] = the ucont adjoint of the dir faces
    hipLaunchKernelGGL(k_vjp_faces<0>, gp, bk, 0, c->stream, d, c->g, c->nvert, lu, uc, nu_t, c->face, ws);
    VFS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_vjp_gather<0>, gp, bk, 0, c->stream, d, ws, ub, cb, 0);
    VFS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_vjp_faces<1>, gp, bk, 0, c->stream, d, c->g, c->nvert, lu, uc, nu_t, c->face, ws);
    VFS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_vjp_gather<1>, gp, bk, 0, c->stream, d, ws, ub, cb, 1);
    VFS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_vjp_faces<2>, gp, bk, 0, c->stream, d, c->g, c->nvert, lu, uc, nu_t, c->face, ws);
    VFS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_vjp_gather<2>, gp, bk, 0, c->stream, d, ws, ub, cb, 1);
    VFS_CHECK_LAUNCH();
    // periodic copy of lucont at the start of Formfunction_2
    if (periodic_vec_T(c, cb)) return -1;
    // IB_BC: periodic face copies, then the IB face rule
    if (d.ip || d.jp || d.kp) {
        hipLaunchKernelGGL(k_periodic_faces_T, dim3(gs), bk, 0, c->stream, d, cb);
        VFS_CHECK_LAUNCH();
    }
    if (d.immersed) {
        hipLaunchKernelGGL(k_ibbc_T_take, gp, bk, 0, c->stream, d, c->nvert, cb, ws);
        VFS_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_ibbc_T_give, gp, bk, 0, c->stream, d, c->g, ws, ub);
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
    hipLaunchKernelGGL(k_c2c_bc_ghost_T_gather, gp, bk, 0, c->stream, d, c->nvert, c->snap, ub);
    VFS_CHECK_LAUNCH();
    if (periodic_vec_T(c, ub)) return -1;
    hipLaunchKernelGGL(k_c2c_interior_T_q, gp, bk, 0, c->stream, d, c->g, c->nvert, ub, ws);
    VFS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_c2c_interior_T_faces, gp, bk, 0, c->stream, d, ws, cb);
    VFS_CHECK_LAUNCH();
    if (periodic_vec_T(c, cb)) return -1;
    hipLaunchKernelGGL(k_vjp_final, gp, bk, 0, c->stream, d, cb, jtv);
    VFS_CHECK_LAUNCH();
    return 0;
}
