// Matrix-free Newton-Krylov solve of the implicit momentum equation
// F(U) = 0, F = FormFunction_SNES (momentum.c:2197-2301), in the structure of
// Implicit_MatrixFree (implicitsolver.c:4196-4441): Newton outer iteration
// with relative tolerance imp_free_tol on ||F|| (SNESSetTolerances, 4247),
// unpreconditioned restarted GMRES on J dU = -F (KSPGMRES + PCNONE, 4258-4271)
// with Eisenstat-Walker forcing terms (SNESKSPSetUseEW, 4244-4245), and the
// no-slip boundary fluxes zeroed in the result (4300-4320).
//
// Differences from the reference, by design: J w is the exact JVP of F
// (vfs_rhs_jvp) instead of PETSc's MatMFFD finite difference, and the
// globalisation is a backtracking line search on ||F|| instead of PETSc
// 3.1's SNESTR trust region.  The iterate sequence of PETSc is therefore
// not reproduced; the converged U is (both solve F(U) = 0 to the tolerance).
#include <cmath>
#include <vector>

#include "vfs_internal.h"
#include "vfs_reduce.h"

int vfs_rhs_jvp_prepared(vfs_ctx *c, const double *nu_t, const double *w, double *jw);
int vfs_rhs_prepare(vfs_ctx *c, const double *U, const double *ucat);
int vfs_noslip_mask(vfs_ctx *c, const double *in, double *out);

struct NKWork {
    int m = 0;
    double *V = nullptr;   // (m+1) x n Krylov basis
    double *F = nullptr, *dU = nullptr, *Ut = nullptr, *lu = nullptr, *ut = nullptr, *w = nullptr;
};

static NKWork *nk_of(vfs_ctx *c) { return reinterpret_cast<NKWork *>(c->nk); }

void vfs_implicit_destroy(vfs_ctx *c)
{
    NKWork *k = nk_of(c);
    if (!k) return;
    hipFree(k->V); hipFree(k->F);
    delete k;
    c->nk = nullptr;
}

static int nk_alloc(vfs_ctx *c, int m)
{
    NKWork *k = nk_of(c);
    if (k && k->m >= m) return 0;
    vfs_implicit_destroy(c);
    k = new NKWork();
    const size_t n = 3 * (size_t)c->d.P;
    if (hipMalloc(&k->V, (size_t)(m + 1) * n * sizeof(double)) != hipSuccess ||
        hipMalloc(&k->F, 6 * n * sizeof(double)) != hipSuccess) {
        hipFree(k->V);
        delete k;
        vfs_set_error("hipMalloc Newton-Krylov workspace");
        return -1;
    }
    k->m = m;
    k->dU = k->F + n; k->Ut = k->F + 2 * n; k->lu = k->F + 3 * n; k->ut = k->F + 4 * n; k->w = k->F + 5 * n;
    c->nk = k;
    return 0;
}

// dots of w with V_0..V_{nv-1}: partial sums per block into red slots
__global__ void k_nk_multidot(long long n, const double *__restrict__ V, int nv, const double *__restrict__ w,
                              double *__restrict__ part)
{
    __shared__ double sh[4];
    for (int v = 0; v < nv; v++) {
        const double *a = V + (long long)v * n;
        double acc = 0;
        for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
            acc += a[t] * w[t];
        double r = block_sum<256>(acc, sh);
        if (threadIdx.x == 0) part[(long long)v * VFS_RED_BLOCKS + blockIdx.x] = r;
    }
}

// w -= sum_v h[v] V_v (h on the device)
__global__ void k_nk_multiaxpy(long long n, const double *__restrict__ V, int nv, const double *__restrict__ h,
                               double *__restrict__ w)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double s = w[t];
    for (int v = 0; v < nv; v++) s -= h[v] * V[(long long)v * n + t];
    w[t] = s;
}

// y = a*x + b*y
__global__ void k_nk_axpby(long long n, double a, const double *__restrict__ x, double b, double *__restrict__ y)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    y[t] = a * x[t] + b * y[t];
}

// y = x + a*z
__global__ void k_nk_waxpy(long long n, const double *__restrict__ x, double a, const double *__restrict__ z,
                           double *__restrict__ y)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    y[t] = x[t] + a * z[t];
}

static const int NK_BLOCKS = 1024;

// host copies of dots of w with nv basis vectors
static int multidot(vfs_ctx *c, long long n, const double *V, int nv, const double *w, double *out)
{
    hipLaunchKernelGGL(k_nk_multidot, dim3(NK_BLOCKS), dim3(256), 0, c->stream, n, V, nv, w, c->red);
    VFS_CHECK_LAUNCH();
    double *res = c->red + (long long)VFS_RED_SLOTS * VFS_RED_BLOCKS;
    if (vfs_reduce_finish(c, NK_BLOCKS, nv, 0, res)) return -1;
    return vfs_read_scalars(c, res, out, nv);
}

static int norm2(vfs_ctx *c, long long n, const double *x, double *out)
{
    double s;
    if (multidot(c, n, x, 1, x, &s)) return -1;
    *out = std::sqrt(s);
    return 0;
}

static int axpby(vfs_ctx *c, long long n, double a, const double *x, double b, double *y)
{
    hipLaunchKernelGGL(k_nk_axpby, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n, a, x, b, y);
    VFS_CHECK_LAUNCH();
    return 0;
}

// coefficients for k_nk_multiaxpy go through the last result slots
static int sub_combination(vfs_ctx *c, long long n, const double *V, int nv, const double *h, double *w)
{
    double *dh = c->red + (long long)VFS_RED_SLOTS * VFS_RED_BLOCKS + 128;
    VFS_CHECK_HIP(hipMemcpyAsync(dh, h, nv * sizeof(double), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_nk_multiaxpy, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n, V, nv, dh, w);
    VFS_CHECK_LAUNCH();
    // h lives in host memory owned by the caller: wait before it changes
    VFS_CHECK_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

// Restarted GMRES(m) without preconditioner on J x = b, x0 = 0, to
// ||b - J x|| <= tol.  Classical Gram-Schmidt with one full
// re-orthogonalisation pass.  Returns the iteration count in *its.
static int gmres_mf(vfs_ctx *c, const double *nu_t, const double *b, double *x, double tol, int m, int maxit,
                    int *its, double *res_out)
{
    NKWork *k = nk_of(c);
    const long long n = 3 * c->d.P;
    VFS_CHECK_HIP(hipMemsetAsync(x, 0, n * sizeof(double), c->stream));
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m), h(m + 1), h2(m + 1);
    int total = 0;
    double beta = 0;
    bool first = true;
    while (true) {
        // r = b - J x into V_0
        double *V0 = k->V;
        if (first) {
            VFS_CHECK_HIP(hipMemcpyAsync(V0, b, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        } else {
            if (vfs_rhs_jvp_prepared(c, nu_t, x, V0)) return -1;
            if (axpby(c, n, 1.0, b, -1.0, V0)) return -1;
        }
        first = false;
        if (norm2(c, n, V0, &beta)) return -1;
        if (beta <= tol || total >= maxit) break;
        if (axpby(c, n, 0.0, V0, 1.0 / beta, V0)) return -1;
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int j = 0;
        double res = beta;
        for (; j < m && total < maxit; j++) {
            double *w = k->V + (long long)(j + 1) * n;
            if (vfs_rhs_jvp_prepared(c, nu_t, k->V + (long long)j * n, w)) return -1;
            if (multidot(c, n, k->V, j + 1, w, h.data())) return -1;
            if (sub_combination(c, n, k->V, j + 1, h.data(), w)) return -1;
            if (multidot(c, n, k->V, j + 1, w, h2.data())) return -1;
            if (sub_combination(c, n, k->V, j + 1, h2.data(), w)) return -1;
            for (int i = 0; i <= j; i++) h[i] += h2[i];
            double hn;
            if (norm2(c, n, w, &hn)) return -1;
            h[j + 1] = hn;
            if (hn > 0) {
                if (axpby(c, n, 0.0, w, 1.0 / hn, w)) return -1;
            }
            for (int i = 0; i < j; i++) {
                double a = cs[i] * h[i] + sn[i] * h[i + 1];
                h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
                h[i] = a;
            }
            double den = std::hypot(h[j], h[j + 1]);
            cs[j] = den > 0 ? h[j] / den : 1.0;
            sn[j] = den > 0 ? h[j + 1] / den : 0.0;
            h[j] = den;
            h[j + 1] = 0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            for (int i = 0; i <= j; i++) H[(size_t)i * m + j] = h[i];
            total++;
            res = std::fabs(g[j + 1]);
            if (res <= tol || hn == 0) { j++; break; }
        }
        // back substitution and update x += V y
        for (int i = j - 1; i >= 0; i--) {
            double s = g[i];
            for (int l = i + 1; l < j; l++) s -= H[(size_t)i * m + l] * y[l];
            y[i] = s / H[(size_t)i * m + i];
        }
        for (int i = 0; i < j; i++) h[i] = -y[i];
        if (sub_combination(c, n, k->V, j, h.data(), x)) return -1;
        if (res <= tol || total >= maxit) {
            beta = res;
            break;
        }
    }
    *its = total;
    *res_out = beta;
    return 0;
}

extern "C" int vfs_implicit_matrixfree(vfs_ctx *c, double *U, const double *U_o, const double *rhs_o,
                                       const double *dP, const double *nu_t, double *ucat, double rtol,
                                       int max_newton, int restart, int max_krylov, int *newton_its,
                                       int *krylov_its, double *fnorm)
{
    const Dims d = c->d;
    VFS_REQUIRE(U && U_o && rhs_o && dP && ucat, "null argument");
    VFS_REQUIRE(nu_t || !d.les, "nu_t required when les is on");
    VFS_REQUIRE(!vfs_has_wallfn(d),
                "the derivatives do not cover the wall model of wall-function walls (bctype -1)");
    VFS_REQUIRE(!d.inviscid, "the derivatives do not cover the inviscid scheme (-inv)");
    VFS_REQUIRE(!d.skew && !d.clark, "the derivatives do not cover -skew or -clark");
    VFS_REQUIRE(d.bc[4] != 5 && d.bc[5] != 4, "the derivatives do not cover open k boundaries (bctypes 5 and 4)");
    VFS_REQUIRE(restart >= 1 && restart <= VFS_RED_SLOTS - 4, "restart must be in [1, 60]");
    VFS_REQUIRE(rtol > 0 && max_newton >= 0 && max_krylov >= 1, "bad tolerance or iteration limits");
    if (nk_alloc(c, restart)) return -1;
    NKWork *k = nk_of(c);
    const long long n = 3 * d.P;
    // F(U), ||F0||
    if (vfs_formfunction_snes(c, U, U_o, rhs_o, dP, nu_t, k->lu, ucat, k->F)) return -1;
    double f0, f;
    if (norm2(c, n, k->F, &f0)) return -1;
    f = f0;
    // Eisenstat-Walker choice 2 (gamma 0.9, alpha (1+sqrt 5)/2, eta_0 0.3, eta_max 0.9)
    const double gamma = 0.9, alpha = 0.5 * (1.0 + std::sqrt(5.0)), eta_max = 0.9;
    double eta = 0.3, fprev = f0;
    int it = 0, kits = 0;
    for (; it < max_newton && f > rtol * f0; it++) {
        if (it > 0) {
            double e = gamma * std::pow(f / fprev, alpha);
            double safe = gamma * std::pow(eta, alpha);
            if (safe > 0.1) e = std::fmax(e, safe);
            eta = std::fmin(e, eta_max);
        }
        // do not solve far past the Newton tolerance
        eta = std::fmax(eta, 0.5 * rtol * f0 / f);
        if (vfs_rhs_prepare(c, U, ucat)) return -1;
        if (axpby(c, n, 0.0, k->F, -1.0, k->w)) return -1;   // w = -F
        int li = 0;
        double lres = 0;
        if (gmres_mf(c, nu_t, k->w, k->dU, eta * f, restart, max_krylov, &li, &lres)) return -1;
        kits += li;
        // backtracking on ||F||
        double lam = 1.0, ft = 0;
        int tries = 0;
        for (;; tries++) {
            hipLaunchKernelGGL(k_nk_waxpy, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n, U, lam, k->dU, k->Ut);
            VFS_CHECK_LAUNCH();
            VFS_CHECK_HIP(hipMemcpyAsync(k->ut, ucat, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            if (vfs_formfunction_snes(c, k->Ut, U_o, rhs_o, dP, nu_t, k->lu, k->ut, k->w)) return -1;
            if (norm2(c, n, k->w, &ft)) return -1;
            if (ft < (1.0 - 1e-4 * lam) * f || tries >= 10) break;
            lam *= 0.5;
        }
        VFS_CHECK_HIP(hipMemcpyAsync(U, k->Ut, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        VFS_CHECK_HIP(hipMemcpyAsync(ucat, k->ut, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        VFS_CHECK_HIP(hipMemcpyAsync(k->F, k->w, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        fprev = f;
        f = ft;
    }
    // the returned U carries zero no-slip boundary fluxes (implicitsolver.c:4300-4320)
    if (vfs_noslip_mask(c, U, U)) return -1;
    VFS_CHECK_HIP(hipStreamSynchronize(c->stream));
    if (newton_its) *newton_its = it;
    if (krylov_its) *krylov_its = kits;
    if (fnorm) *fnorm = f;
    return 0;
}
