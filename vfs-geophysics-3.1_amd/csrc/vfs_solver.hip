// Pressure Poisson solve on gfx950: restarted GMRES, right-preconditioned by
// one V(1,1)-cycle of an aggregation multigrid, in FP64 or FP32 inside an
// FP64 defect-correction loop.  Replaces HYPRE GMRES(51) + BoomerAMG of
// PoissonSolver_Hypre (poisson_hypre.c:26-89, 498-630); parity with the
// reference is on the converged solution (true FP64 residual <= rtol, then
// the mean-free normalisation of Remove_Nullspace, poisson_hypre.c:194-255).
//
// Multigrid: 2x2x2 piecewise-constant aggregation with Galerkin coarse
// operators (27-point), damped j-line Jacobi smoothing (a tridiagonal Thomas
// solve along every wall-normal grid line: the tanh-stretched direction is
// where point smoothers stall), and a dense inverse of the regularised
// coarsest operator.  The constant vector spans the null space and is the
// left null vector of every level (the flux form makes column sums vanish),
// so the regularised coarse solve returns the mean-free correction.
#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstring>
#include <vector>

#include "vfs_internal.h"
#include "vfs_poisson.h"
#include "vfs_reduce.h"

// Scale of the prolongated coarse correction (over-correction: the
// piecewise-constant prolongation of an aggregate undershoots the smooth
// error), for 2x2x2 aggregates and for 2x2x1 ones (k kept).  Measured with
// GMRES to 1e-8: 1.4 / 1.2 take 22 iterations at 256x128x128 (25 at 1 / 1),
// 23 at 512x256x256 (31) and 41 on the 768x384x48 turbine slab (59).  The
// hierarchy loses stability above them: 1.9 on 2x2x2 aggregates, and 1.3 on
// 2x2x1 ones with 1.4 on the others, take more iterations than 1 / 1.
#ifndef VFS_MG_OMEGA
#define VFS_MG_OMEGA 0.9    // damping of the j-line Jacobi smoother
#endif
#ifndef VFS_MG_WDEPTH
#define VFS_MG_WDEPTH 2     // levels l < this visit level l + 1 twice (W-cycle)
#endif
#ifndef VFS_MG_ALPHA
#define VFS_MG_ALPHA 1.4
#endif
#ifndef VFS_MG_ALPHA_SEMI
#define VFS_MG_ALPHA_SEMI 1.2
#endif

struct LV {
    int nx, ny, nz;
    long long n;
    int pi, pj, pk;
    int nslot;
    int sk;   // k cells of the finer level per cell of this level: 2, or 1 where k is not coarsened
};

static LV lv_of(const MGLevel &l) { return LV{l.nx, l.ny, l.nz, l.n, l.per_i, l.per_j, l.per_k, l.nslot, l.sk}; }

// neighbour offsets (di, dj, dk) of the 19 slots of poisson.c:951-969; a
// constexpr table so unrolled loops fold the offsets into immediates
struct Slot19 {
    static constexpr int t[19][3] = {
        {0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
        {1, 1, 0}, {1, -1, 0}, {-1, 1, 0}, {-1, -1, 0}, {0, 1, 1}, {0, 1, -1}, {0, -1, 1},
        {0, -1, -1}, {1, 0, 1}, {1, 0, -1}, {-1, 0, 1}, {-1, 0, -1}};
};

__device__ __forceinline__ void slot_off(int nslot, int s, int &di, int &dj, int &dk)
{
    if (nslot == 19) { di = Slot19::t[s][0]; dj = Slot19::t[s][1]; dk = Slot19::t[s][2]; }
    else { dk = s / 9 - 1; dj = (s / 3) % 3 - 1; di = s % 3 - 1; }
}

// wrap (periodic) or reject (bounded) a neighbour coordinate
__device__ __forceinline__ bool nb_ok(const LV &L, int &k, int &j, int &i)
{
    if (i < 0) { if (!L.pi) return false; i += L.nx; } else if (i >= L.nx) { if (!L.pi) return false; i -= L.nx; }
    if (j < 0) { if (!L.pj) return false; j += L.ny; } else if (j >= L.ny) { if (!L.pj) return false; j -= L.ny; }
    if (k < 0) { if (!L.pk) return false; k += L.nz; } else if (k >= L.nz) { if (!L.pk) return false; k -= L.nz; }
    return true;
}

__device__ __forceinline__ long long cix(const LV &L, int k, int j, int i) { return ((long long)k * L.ny + j) * L.nx + i; }

// ---------------------------------------------------------------------------
// set-up kernels
// ---------------------------------------------------------------------------
__global__ void k_gather_level0(Dims d, LV L, const double *__restrict__ coef, double *A64, float *A32)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L.n) return;
    int i = (int)(t % L.nx), j = (int)((t / L.nx) % L.ny), k = (int)(t / ((long long)L.nx * L.ny));
    long long g = IX(d, k + 1, j + 1, i + 1);
    for (int s = 0; s < 19; s++) {
        double v = coef[(long long)s * d.P + g];
        A64[(long long)s * L.n + t] = v;
        A32[(long long)s * L.n + t] = (float)v;
    }
}

// Galerkin product R A P for piecewise-constant 2x2x2 aggregation
__global__ void k_galerkin(LV F, LV C, const double *__restrict__ Af, double *Ac64, float *Ac32)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= C.n) return;
    int I = (int)(t % C.nx), J = (int)((t / C.nx) % C.ny), K = (int)(t / ((long long)C.nx * C.ny));
    double acc[27];
#pragma unroll
    for (int s = 0; s < 27; s++) acc[s] = 0;
    for (int ck = 0; ck < C.sk; ck++)
        for (int cj = 0; cj < 2; cj++)
            for (int ci = 0; ci < 2; ci++) {
                int fk = (C.sk == 1 ? K : 2 * K + ck), fj = (F.ny == 1 ? J : 2 * J + cj), fi = (F.nx == 1 ? I : 2 * I + ci);
                if ((F.ny == 1 && cj) || (F.nx == 1 && ci)) continue;
                if (fk >= F.nz || fj >= F.ny || fi >= F.nx) continue;
                long long f = cix(F, fk, fj, fi);
                for (int s = 0; s < F.nslot; s++) {
                    double a = Af[(long long)s * F.n + f];
                    if (a == 0.0) continue;
                    int di, dj, dk;
                    slot_off(F.nslot, s, di, dj, dk);
                    int nk = fk + dk, nj = fj + dj, ni = fi + di;
                    if (!nb_ok(F, nk, nj, ni)) continue;
                    int NK = C.sk == 1 ? nk : nk / 2, NJ = F.ny == 1 ? nj : nj / 2, NI = F.nx == 1 ? ni : ni / 2;
                    int oK = NK - K, oJ = NJ - J, oI = NI - I;
                    if (C.pk) { if (oK > 1) oK -= C.nz; if (oK < -1) oK += C.nz; }
                    if (C.pj) { if (oJ > 1) oJ -= C.ny; if (oJ < -1) oJ += C.ny; }
                    if (C.pi) { if (oI > 1) oI -= C.nx; if (oI < -1) oI += C.nx; }
                    acc[(oK + 1) * 9 + (oJ + 1) * 3 + (oI + 1)] += a;
                }
            }
    for (int s = 0; s < 27; s++) {
        Ac64[(long long)s * C.n + t] = acc[s];
        Ac32[(long long)s * C.n + t] = (float)acc[s];
    }
}

// Thomas factors of the j-line tridiagonal part, one thread per (k,i) line
__global__ void k_line_factors(LV L, const double *__restrict__ A, double *tl64, double *tu64, double *ti64,
                               float *tl32, float *tu32, float *ti32)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)L.nx * L.nz) return;
    int i = (int)(t % L.nx), k = (int)(t / L.nx);
    const int sl = L.nslot == 19 ? 4 : 10, sc = L.nslot == 19 ? 0 : 13, su = L.nslot == 19 ? 3 : 16;
    double up = 0;
    for (int j = 0; j < L.ny; j++) {
        long long id = cix(L, k, j, i);
        double a = j > 0 ? A[(long long)sl * L.n + id] : 0.0;
        double dgl = A[(long long)sc * L.n + id];
        double c = j < L.ny - 1 ? A[(long long)su * L.n + id] : 0.0;
        if (dgl == 0.0) { dgl = 1.0; a = 0.0; c = 0.0; }   // no unknown here: identity row
        double den = dgl - a * up;
        double inv = 1.0 / den;
        up = c * inv;
        // forward sweep coefficients of dp_j = r_j*inv_j - (a_j*inv_j)*dp_{j-1}
        tl64[id] = a * inv; tu64[id] = up; ti64[id] = inv;
        tl32[id] = (float)(a * inv); tu32[id] = (float)up; ti32[id] = (float)inv;
    }
}

// ---------------------------------------------------------------------------
// multigrid kernels (precision T, accumulation in T except reductions)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T apply_row(const LV &L, const T *__restrict__ A, const T *__restrict__ x, int k, int j, int i,
                                       long long id, const T *__restrict__ e, LV C, T alpha)
{
    T s = 0;
    for (int q = 0; q < L.nslot; q++) {
        T a = A[(long long)q * L.n + id];
        if (a == T(0)) continue;
        int di, dj, dk;
        slot_off(L.nslot, q, di, dj, dk);
        int nk = k + dk, nj = j + dj, ni = i + di;
        if (!nb_ok(L, nk, nj, ni)) continue;
        long long nb = cix(L, nk, nj, ni);
        T xv = x[nb];
        if (e) {
            int ck = C.sk == 1 ? nk : nk / 2, cj = L.ny == 1 ? nj : nj / 2, ci = L.nx == 1 ? ni : ni / 2;
            xv += alpha * e[cix(C, ck, cj, ci)];
        }
        s += a * xv;
    }
    return s;
}

// Row of A x with constant neighbour offsets: only for rows whose stencil
// stays inside the level (no wrap, no boundary), i.e. the bulk of the grid.
template <typename T, int NS>
__device__ __forceinline__ T apply_row_fast(const LV &L, const T *__restrict__ A, const T *__restrict__ x, long long id)
{
    const long long sx = 1, sy = L.nx, sz = (long long)L.nx * L.ny;
    T s = 0;
#pragma unroll
    for (int q = 0; q < NS; q++) {
        int di, dj, dk;
        slot_off(NS, q, di, dj, dk);
        s += A[(long long)q * L.n + id] * x[id + di * sx + dj * sy + dk * sz];
    }
    return s;
}

// Branch-free row of A x for any cell: periodic neighbours wrap, bounded
// neighbours are clamped (their coefficients are zero by construction: the
// drop rule removes couplings to non-unknowns on level 0 and Galerkin
// products inherit that), so every lane of a wavefront runs the same code.
template <typename T, int NS>
__device__ __forceinline__ T apply_row_any(const LV &L, const T *__restrict__ A, const T *__restrict__ x, long long id,
                                           int i, int j, int k)
{
    const int im = i > 0 ? i - 1 : (L.pi ? L.nx - 1 : 0), ip = i < L.nx - 1 ? i + 1 : (L.pi ? 0 : L.nx - 1);
    const int jm = j > 0 ? j - 1 : (L.pj ? L.ny - 1 : 0), jp = j < L.ny - 1 ? j + 1 : (L.pj ? 0 : L.ny - 1);
    const int km = k > 0 ? k - 1 : (L.pk ? L.nz - 1 : 0), kp = k < L.nz - 1 ? k + 1 : (L.pk ? 0 : L.nz - 1);
    const long long sy = L.nx, sz = (long long)L.nx * L.ny;
    const long long I[3] = {im, i, ip}, J[3] = {jm * sy, j * sy, jp * sy}, K[3] = {km * sz, k * sz, kp * sz};
    T s = 0;
#pragma unroll
    for (int q = 0; q < NS; q++) {
        int di, dj, dk;
        slot_off(NS, q, di, dj, dk);
        s += A[(long long)q * L.n + id] * x[K[dk + 1] + J[dj + 1] + I[di + 1]];
    }
    return s;
}

// Row kernels launch one thread per i and one block row per (j, k)
// (blockIdx.y = j, blockIdx.z = k), so all index arithmetic but i is scalar.
// (i, j, k) of this thread on a (ceil(nx/b), ny, nz) grid, with the
// workgroups of each XCD on a contiguous range of planes
#define ROW_IDX(L)                                                             \
    const unsigned _lb = xcd_block(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), \
                                   gridDim.x * gridDim.y * gridDim.z);         \
    const int i = (int)(_lb % gridDim.x) * blockDim.x + threadIdx.x;           \
    if (i >= (L).nx) return;                                                   \
    const int j = (int)((_lb / gridDim.x) % gridDim.y), k = (int)(_lb / (gridDim.x * gridDim.y)); \
    const long long t = ((long long)k * (L).ny + j) * (L).nx + i;

// y = A x
template <typename T, int NS>
__global__ void k_spmv(LV L, const T *__restrict__ A, const T *__restrict__ x, T *__restrict__ y)
{
    ROW_IDX(L)
    y[t] = apply_row_any<T, NS>(L, A, x, t, i, j, k);
}

// r = b - A x
template <typename T, int NS>
__global__ void k_resid(LV L, const T *__restrict__ A, const T *__restrict__ x, const T *__restrict__ b, T *__restrict__ r)
{
    ROW_IDX(L)
    r[t] = b[t] - apply_row_any<T, NS>(L, A, x, t, i, j, k);
}

// Restricted residual bc = P^T (b - A x) in one pass: a thread takes one fine
// i and the (up to) 2 x 2 fine cells of its aggregate in j and k, and the
// even lane adds its odd neighbour's sum (the aggregate's other i).  Grid
// (ceil(F.nx / bs), C.ny, C.nz) with an even block size, and no lane leaves
// before the exchange.
template <typename T, int NS>
__global__ void k_resid_restrict(LV F, LV C, const T *__restrict__ A, const T *__restrict__ x,
                                 const T *__restrict__ b, T *__restrict__ bc)
{
    const unsigned lb = xcd_block(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z),
                                  gridDim.x * gridDim.y * gridDim.z);
    const int i = (int)(lb % gridDim.x) * blockDim.x + threadIdx.x;
    const int J = (int)((lb / gridDim.x) % gridDim.y), K = (int)(lb / (gridDim.x * gridDim.y));
    const int nj = F.ny > 1 ? 2 : 1, nk = C.sk;
    T s = 0;
    if (i < F.nx) {
        for (int ck = 0; ck < nk; ck++) {
            const int k = nk == 2 ? 2 * K + ck : K;
            if (k >= F.nz) break;
            for (int cj = 0; cj < nj; cj++) {
                const int j = nj == 2 ? 2 * J + cj : J;
                if (j >= F.ny) break;
                const long long t = cix(F, k, j, i);
                s += b[t] - apply_row_any<T, NS>(F, A, x, t, i, j, k);
            }
        }
    }
    if (F.nx > 1) {
        s += __shfl_xor(s, 1, 64);
        if (i < F.nx && !(i & 1)) bc[cix(C, K, J, i >> 1)] = s;
    } else if (i == 0) {
        bc[cix(C, K, J, 0)] = s;
    }
}

// Row of A (x + alpha P e): x on level L, the coarse correction e on level C
template <typename T, int NS>
__device__ __forceinline__ T apply_row_pe(const LV &L, const LV &C, const T *__restrict__ A, const T *__restrict__ x,
                                          const T *__restrict__ e, T alpha, long long id, int i, int j, int k)
{
    const int im = i > 0 ? i - 1 : (L.pi ? L.nx - 1 : 0), ip = i < L.nx - 1 ? i + 1 : (L.pi ? 0 : L.nx - 1);
    const int jm = j > 0 ? j - 1 : (L.pj ? L.ny - 1 : 0), jp = j < L.ny - 1 ? j + 1 : (L.pj ? 0 : L.ny - 1);
    const int km = k > 0 ? k - 1 : (L.pk ? L.nz - 1 : 0), kp = k < L.nz - 1 ? k + 1 : (L.pk ? 0 : L.nz - 1);
    const long long sy = L.nx, sz = (long long)L.nx * L.ny, cy = C.nx, cz = (long long)C.nx * C.ny;
    const long long I[3] = {im, i, ip}, J[3] = {jm * sy, j * sy, jp * sy}, K[3] = {km * sz, k * sz, kp * sz};
    const long long CI[3] = {im >> 1, i >> 1, ip >> 1}, CJ[3] = {(jm >> 1) * cy, (j >> 1) * cy, (jp >> 1) * cy},
                    CK[3] = {(km >> (C.sk - 1)) * cz, (k >> (C.sk - 1)) * cz, (kp >> (C.sk - 1)) * cz};
    T s = 0;
#pragma unroll
    for (int q = 0; q < NS; q++) {
        int di, dj, dk;
        slot_off(NS, q, di, dj, dk);
        const T xv = x[K[dk + 1] + J[dj + 1] + I[di + 1]] + alpha * e[CK[dk + 1] + CJ[dj + 1] + CI[di + 1]];
        s += A[(long long)q * L.n + id] * xv;
    }
    return s;
}

// xe = x + alpha P e and r = b - A xe in one pass (x itself is not written,
// since the neighbours' old values are still being read)
template <typename T, int NS>
__global__ void k_resid_pe(LV L, LV C, const T *__restrict__ A, const T *__restrict__ x, const T *__restrict__ e,
                           T alpha, const T *__restrict__ b, T *__restrict__ r, T *__restrict__ xe)
{
    ROW_IDX(L)
    r[t] = b[t] - apply_row_pe<T, NS>(L, C, A, x, e, alpha, t, i, j, k);
    xe[t] = x[t] + alpha * e[cix(C, k >> (C.sk - 1), j >> 1, i >> 1)];
}

// x += alpha * P e  (piecewise-constant prolongation)
template <typename T>
__global__ void k_prolong(LV L, LV C, const T *__restrict__ e, T alpha, T *__restrict__ x)
{
    ROW_IDX(L)
    x[t] += alpha * e[cix(C, k >> (C.sk - 1), j >> 1, i >> 1)];
}

// Damped j-line Jacobi update xn = xo + omega * T^{-1} r along every (k,i)
// grid line (Thomas sweeps with the factors of k_line_factors); ZG: zero
// guess (xo unused).  xo may alias xn: each lane touches its own line only.
// One lane per line; the loads of a batch of B steps are issued before its
// dependent recurrence, so a wavefront keeps B x 3 loads in flight: the
// fine-level lines are few (nx * nz lanes) and long, so the sweeps are
// latency-bound.
template <typename T, int B, bool ZG>
__global__ void __launch_bounds__(64) k_thomas(LV L, const T *__restrict__ r, const T *__restrict__ ti,
                                               const T *__restrict__ tm, const T *__restrict__ tu, const T *xo,
                                               T omega, T *__restrict__ dscr, T *xn)
{
    // rows are addressed as (uniform row pointer)[lane], so the loads take a
    // scalar base and a 32-bit lane offset
    const unsigned i = blockIdx.x * 64u + threadIdx.x;
    if (i >= (unsigned)L.nx) return;
    const long long plane = (long long)blockIdx.y * L.ny * L.nx, sy = L.nx;
    T dp = 0;
    for (int j0 = 0; j0 < L.ny; j0 += B) {
        T rr[B], ii[B], mm[B];
#pragma unroll
        for (int q = 0; q < B; q++)
            if (j0 + q < L.ny) {
                const long long o = plane + (j0 + q) * sy;
                rr[q] = (r + o)[i]; ii[q] = (ti + o)[i]; mm[q] = (tm + o)[i];
            }
#pragma unroll
        for (int q = 0; q < B; q++)
            if (j0 + q < L.ny) {
                dp = rr[q] * ii[q] - mm[q] * dp;
                (dscr + plane + (j0 + q) * sy)[i] = dp;
            }
    }
    T xj = 0;
    for (int j1 = L.ny - 1; j1 >= 0; j1 -= B) {
        T dd[B], uu[B], oo[B];
#pragma unroll
        for (int q = 0; q < B; q++)
            if (j1 - q >= 0) {
                const long long o = plane + (j1 - q) * sy;
                dd[q] = (dscr + o)[i]; uu[q] = (tu + o)[i]; oo[q] = ZG ? T(0) : (xo + o)[i];
            }
#pragma unroll
        for (int q = 0; q < B; q++)
            if (j1 - q >= 0) {
                xj = dd[q] - uu[q] * xj;
                (xn + plane + (j1 - q) * sy)[i] = oo[q] + omega * xj;
            }
    }
}

// dense coarsest solve x = Cinv b over the listed unknowns; other cells 0
// (cinv column-major; nc <= VFS_MAX_DENSE; one workgroup)
#define VFS_MAX_DENSE 1024
template <typename T>
__device__ __forceinline__ void dense_solve(long long n, int nc, const long long *__restrict__ cidx,
                                            const T *__restrict__ cinv, const T *__restrict__ b, T *__restrict__ x,
                                            T *sb)
{
    for (int q = threadIdx.x; q < nc; q += blockDim.x) sb[q] = b[cidx[q]];
    __syncthreads();
    for (long long q = threadIdx.x; q < n; q += blockDim.x) x[q] = 0;
    __syncthreads();
    for (int r = threadIdx.x; r < nc; r += blockDim.x) {
        T s = 0;
        for (int c = 0; c < nc; c++) s += cinv[(long long)c * nc + r] * sb[c];
        x[cidx[r]] = s;
    }
}

// The same solve by wavefronts, x = F b with F the inverse scattered to all
// n cells of the coarsest level (zero rows and columns for the cells that
// are no unknowns), row-major: one wavefront per row, coalesced across the
// lanes, with no gather.  Grid ceil(n / 4) x 256.
template <typename T>
__global__ void __launch_bounds__(256) k_dense_mv(long long n, const T *__restrict__ F, const T *__restrict__ b,
                                                  T *__restrict__ x)
{
    const long long q = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (q >= n) return;   // uniform per wavefront
    const T *row = F + q * n;
    double s = 0;
    for (long long c = lane; c < n; c += 64) s += (double)row[c] * (double)b[c];
    s = wave_sum(s);
    if (lane == 0) x[q] = (T)s;
}

template <typename T>
__global__ void k_coarse_solve(long long n, int nc, const long long *__restrict__ cidx, const T *__restrict__ cinv,
                               const T *__restrict__ b, T *__restrict__ x)
{
    __shared__ T sb[VFS_MAX_DENSE];
    dense_solve<T>(n, nc, cidx, cinv, b, x, sb);
}

// ---------------------------------------------------------------------------
// The small levels in one workgroup.  Below ~16K unknowns every multigrid
// kernel is a few microseconds of launch latency around nanoseconds of work,
// so the whole V-cycle of levels first..last (the last one dense) runs as a
// single 1024-thread workgroup with barriers between its phases.  The phases
// are the ones of vcycle(): zero-guess line smoothing, residual, restriction,
// dense solve, prolongation, residual, line smoothing.
// ---------------------------------------------------------------------------
#define MG_MAXL 16
template <typename T>
struct SmallLevels {
    int first, last;
    LV lv[MG_MAXL];
    const T *A[MG_MAXL], *tl[MG_MAXL], *tu[MG_MAXL], *ti[MG_MAXL];
    T *b[MG_MAXL], *x[MG_MAXL], *y[MG_MAXL], *d[MG_MAXL];
    int nc;
    const long long *cidx;
    const T *cinv;
    T omega;
    T alpha[MG_MAXL];   // coarse-correction scale into level l from level l + 1
    int zero;    // the first level starts from x = 0 (else from the given x)
};

// xn = xo + omega T^{-1} r along every j-line (xo may be null: zero guess)
template <typename T>
__device__ __forceinline__ void wg_thomas(const LV &L, const T *r, const T *ti, const T *tm, const T *tu, const T *xo,
                                          T omega, T *dscr, T *xn)
{
    constexpr int MJ = 8;    // lines up to this length are held in registers
    const long long nl = (long long)L.nx * L.nz, sy = L.nx;
    for (long long ln = threadIdx.x; ln < nl; ln += blockDim.x) {
        const int i = (int)(ln % L.nx), k = (int)(ln / L.nx);
        const long long base = (long long)k * L.ny * L.nx + i;
        if (L.ny <= MJ) {
            // all loads of the line first, then the two recurrences
            T rr[MJ], ii[MJ], mm[MJ], uu[MJ], oo[MJ];
#pragma unroll
            for (int j = 0; j < MJ; j++)
                if (j < L.ny) {
                    const long long id = base + j * sy;
                    rr[j] = r[id]; ii[j] = ti[id]; mm[j] = tm[id]; uu[j] = tu[id]; oo[j] = xo ? xo[id] : T(0);
                }
            T dp = 0;
#pragma unroll
            for (int j = 0; j < MJ; j++)
                if (j < L.ny) { dp = rr[j] * ii[j] - mm[j] * dp; rr[j] = dp; }
            T xj = 0;
#pragma unroll
            for (int j = MJ - 1; j >= 0; j--)
                if (j < L.ny) { xj = rr[j] - uu[j] * xj; xn[base + j * sy] = oo[j] + omega * xj; }
            continue;
        }
        T dp = 0;
        for (int j = 0; j < L.ny; j++) {
            const long long id = base + j * sy;
            dp = r[id] * ti[id] - tm[id] * dp;
            dscr[id] = dp;
        }
        T xj = 0;
        for (int j = L.ny - 1; j >= 0; j--) {
            const long long id = base + j * sy;
            xj = dscr[id] - tu[id] * xj;
            xn[id] = (xo ? xo[id] : T(0)) + omega * xj;
        }
    }
}

template <typename T>
__device__ __forceinline__ void wg_resid(const LV &L, const T *A, const T *x, const T *b, T *r)
{
    for (long long t = threadIdx.x; t < L.n; t += blockDim.x) {
        const int i = (int)(t % L.nx), j = (int)((t / L.nx) % L.ny), k = (int)(t / ((long long)L.nx * L.ny));
        r[t] = b[t] - apply_row_any<T, 27>(L, A, x, t, i, j, k);
    }
}

template <typename T>
__global__ void __launch_bounds__(1024) k_small_cycle(SmallLevels<T> S)
{
    for (int l = S.first; l < S.last; l++) {
        const LV L = S.lv[l], C = S.lv[l + 1];
        if (l == S.first && !S.zero) {
            wg_resid<T>(L, S.A[l], S.x[l], S.b[l], S.y[l]);
            __syncthreads();
            wg_thomas<T>(L, S.y[l], S.ti[l], S.tl[l], S.tu[l], S.x[l], S.omega, S.d[l], S.x[l]);
        } else {
            wg_thomas<T>(L, S.b[l], S.ti[l], S.tl[l], S.tu[l], nullptr, S.omega, S.d[l], S.x[l]);
        }
        __syncthreads();
        wg_resid<T>(L, S.A[l], S.x[l], S.b[l], S.y[l]);
        __syncthreads();
        for (long long t = threadIdx.x; t < C.n; t += blockDim.x) {
            const int I = (int)(t % C.nx), J = (int)((t / C.nx) % C.ny), K = (int)(t / ((long long)C.nx * C.ny));
            T s = 0;
            for (int ck = 0; ck < C.sk; ck++)
                for (int cj = 0; cj < 2; cj++)
                    for (int ci = 0; ci < 2; ci++) {
                        if ((L.ny == 1 && cj) || (L.nx == 1 && ci)) continue;
                        const int fk = C.sk == 1 ? K : 2 * K + ck, fj = L.ny == 1 ? J : 2 * J + cj,
                                  fi = L.nx == 1 ? I : 2 * I + ci;
                        if (fk >= L.nz || fj >= L.ny || fi >= L.nx) continue;
                        s += S.y[l][cix(L, fk, fj, fi)];
                    }
            S.b[l + 1][t] = s;
        }
        __syncthreads();
    }
    {   // dense solve on the last level
        __shared__ T sb[VFS_MAX_DENSE];
        dense_solve<T>(S.lv[S.last].n, S.nc, S.cidx, S.cinv, S.b[S.last], S.x[S.last], sb);
        __syncthreads();
    }
    for (int l = S.last - 1; l >= S.first; l--) {
        const LV L = S.lv[l], C = S.lv[l + 1];
        T *x = S.x[l];
        const T *e = S.x[l + 1];
        for (long long t = threadIdx.x; t < L.n; t += blockDim.x) {
            const int i = (int)(t % L.nx), j = (int)((t / L.nx) % L.ny), k = (int)(t / ((long long)L.nx * L.ny));
            const int ck = C.sk == 1 ? k : k / 2, cj = L.ny == 1 ? j : j / 2, ci = L.nx == 1 ? i : i / 2;
            x[t] += S.alpha[l] * e[cix(C, ck, cj, ci)];
        }
        __syncthreads();
        wg_resid<T>(L, S.A[l], x, S.b[l], S.y[l]);
        __syncthreads();
        wg_thomas<T>(L, S.y[l], S.ti[l], S.tl[l], S.tu[l], x, S.omega, S.d[l], x);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// vector kernels for the Krylov loop (reductions accumulate in FP64)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_multidot(long long n, const T *__restrict__ V, long long ldv, int nv, const T *__restrict__ w, double *part)
{
    __shared__ double sh[4];
    for (int v = 0; v < nv; v++) {
        double acc = 0;
        const T *vv = V + (long long)v * ldv;
        for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
            acc += (double)vv[q] * (double)w[q];
        double r = block_sum<256>(acc, sh);
        if (threadIdx.x == 0) part[(long long)v * VFS_RED_BLOCKS + blockIdx.x] = r;
    }
}

// dots of w with NV basis vectors V_{v0..v0+NV-1} in one pass over w, plus
// |w|^2 when self != 0; partials per block at part[v*RB + block] (|w|^2 in
// the last slot).  The full trips of the grid-stride loop issue the loads of
// U element blocks with no predicate, so none waits for another; elements
// are addressed as (uniform pointer)[thread] (a scalar base and a 32-bit
// offset).  The partial blocks at the end go one at a time.  Block size 256.
template <typename T, int NV>
__global__ void __launch_bounds__(256) k_multidot8(long long n, const T *__restrict__ V, long long ldv, int v0, int self,
                                                   const T *__restrict__ w, double *part)
{
    constexpr int U = 4;
    __shared__ double sh[4];
    double acc[NV], accs = 0;
#pragma unroll
    for (int v = 0; v < NV; v++) acc[v] = 0;
    const unsigned t = threadIdx.x;
    const long long S = (long long)gridDim.x * 256;
    const T *Vb = V + (long long)v0 * ldv;
    long long b0 = (long long)blockIdx.x * 256;
    for (; b0 + (U - 1) * S + 256 <= n; b0 += U * S) {
        T wv[U], vv[U][NV];
#pragma unroll
        for (int u = 0; u < U; u++) {
            wv[u] = (w + b0 + u * S)[t];
#pragma unroll
            for (int v = 0; v < NV; v++) vv[u][v] = (Vb + v * ldv + b0 + u * S)[t];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
#pragma unroll
            for (int v = 0; v < NV; v++) acc[v] += (double)vv[u][v] * (double)wv[u];
            accs += (double)wv[u] * (double)wv[u];
        }
    }
    for (; b0 < n; b0 += S)
        if (b0 + t < n) {
            const double wq = (double)(w + b0)[t];
#pragma unroll
            for (int v = 0; v < NV; v++) acc[v] += (double)(Vb + v * ldv + b0)[t] * wq;
            accs += wq * wq;
        }
#pragma unroll
    for (int v = 0; v < NV; v++) {
        double r = block_sum<256>(acc[v], sh);
        if (threadIdx.x == 0) part[(long long)(v0 + v) * VFS_RED_BLOCKS + blockIdx.x] = r;
    }
    if (self) {
        double r = block_sum<256>(accs, sh);
        if (threadIdx.x == 0) part[(long long)(VFS_RED_SLOTS - 1) * VFS_RED_BLOCKS + blockIdx.x] = r;
    }
}

template <typename T>
static void launch_multidot8(hipStream_t st, int nbr, long long n, const T *V, long long ldv, int v0, int nv, int self,
                             const T *w, double *part)
{
    switch (nv) {
#define MD8(K) case K: hipLaunchKernelGGL((k_multidot8<T, K>), dim3(nbr), dim3(256), 0, st, n, V, ldv, v0, self, w, part); break;
        MD8(1) MD8(2) MD8(3) MD8(4) MD8(5) MD8(6) MD8(7) MD8(8)
#undef MD8
        default: break;
    }
}

// wout = win - sum_v (h[v] / nrm2[v]) V_v  (nrm2 null: 1), h and nrm2 on the
// device; win may alias wout.  With part, also the block partials of
// |wout|^2 at part[(VFS_RED_SLOTS-1)*RB + block].  nv <= 64, block size 256.
// At most two waves per SIMD: fewer basis-vector streams in flight at once
// keep the DRAM pages open longer (111 -> 87 us in FP64, 33 -> 24 us in FP32
// at 256x128x128 with 25 vectors).
template <typename T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 2))) k_multiaxpy(long long n, const T *__restrict__ V, long long ldv, int nv,
                                                   const double *__restrict__ h, const double *__restrict__ nrm2,
                                                   const T *win, T *wout, double *part)
{
    constexpr int U = 4;   // element blocks per thread and trip, loaded together
    __shared__ double sc[64];
    __shared__ double sh[4];
    if (threadIdx.x < nv) sc[threadIdx.x] = h[threadIdx.x] / (nrm2 ? nrm2[threadIdx.x] : 1.0);
    __syncthreads();
    const unsigned t = threadIdx.x;
    const long long S = (long long)gridDim.x * 256;
    double acc = 0;
    long long b0 = (long long)blockIdx.x * 256;
    for (; b0 + (U - 1) * S + 256 <= n; b0 += U * S) {
        T s[U];
#pragma unroll
        for (int u = 0; u < U; u++) s[u] = (win + b0 + u * S)[t];
        for (int v = 0; v < nv; v++) {
            const T cv = (T)sc[v];
            const T *vp = V + (long long)v * ldv + b0;
#pragma unroll
            for (int u = 0; u < U; u++) s[u] -= cv * (vp + u * S)[t];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            (wout + b0 + u * S)[t] = s[u];
            acc += (double)s[u] * (double)s[u];
        }
    }
    for (; b0 < n; b0 += S)
        if (b0 + t < n) {
            T s = (win + b0)[t];
            for (int v = 0; v < nv; v++) s -= (T)sc[v] * (V + (long long)v * ldv + b0)[t];
            (wout + b0)[t] = s;
            acc += (double)s * (double)s;
        }
    if (part) {
        double r = block_sum<256>(acc, sh);
        if (threadIdx.x == 0) part[(long long)(VFS_RED_SLOTS - 1) * VFS_RED_BLOCKS + blockIdx.x] = r;
    }
}

// u = sum_v y[v] V_v  (y on the device)
template <typename T>
__global__ void k_combine(long long n, const T *__restrict__ V, long long ldv, int nv, const double *__restrict__ y,
                          T *__restrict__ u)
{
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    T s = 0;
    for (int v = 0; v < nv; v++) s += (T)y[v] * V[(long long)v * ldv + q];
    u[q] = s;
}

template <typename T>
__global__ void k_cast_scale(long long n, const double *__restrict__ src, const double *__restrict__ nrm2, T *__restrict__ dst)
{
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    double s = nrm2 ? 1.0 / sqrt(nrm2[0]) : 1.0;
    dst[q] = (T)(src[q] * s);
}

// x64 += scale * e  (scale = sqrt(nrm2[0]))
template <typename T>
__global__ void k_add_scaled(long long n, const T *__restrict__ e, const double *__restrict__ nrm2, double *__restrict__ x)
{
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    double s = nrm2 ? sqrt(nrm2[0]) : 1.0;
    x[q] += s * (double)e[q];
}

template <typename T>
__global__ void k_axpy_dev(long long n, const T *__restrict__ z, T *__restrict__ x)
{
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    x[q] += z[q];
}

// DA <-> level-0 compact
__global__ void k_da_to_c(Dims d, LV L, const double *__restrict__ src, const unsigned char *__restrict__ rowflag, double *dst)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L.n) return;
    int i = (int)(t % L.nx), j = (int)((t / L.nx) % L.ny), k = (int)(t / ((long long)L.nx * L.ny));
    long long g = IX(d, k + 1, j + 1, i + 1);
    dst[t] = rowflag[g] ? src[g] : 0.0;
}

// phi on the DA layout: unknown cells from x, others 0 (Convert_Phi2_Phi,
// poisson.c:92-130); nullspace mean removed over nvert < 0.1 cells.
__global__ void k_c_to_da(Dims d, LV L, const double *__restrict__ x, const double *__restrict__ nvert,
                          const double *__restrict__ mean, double *phi)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.P) return;
    int i = (int)(t % d.mx), j = (int)((t / d.mx) % d.my), k = (int)(t / ((long long)d.mx * d.my));
    double v = 0;
    if (i >= 1 && j >= 1 && k >= 1 && i <= d.mx - 2 && j <= d.my - 2 && k <= d.mz - 2) {
        if (!((int)nvert[t] > 0.1)) {
            v = x[cix(L, k - 1, j - 1, i - 1)];
            if (nvert[t] < 0.1) v += mean[0];
        }
    }
    phi[t] = v;
}

// sum and count over nvert < 0.1 interior cells (Remove_Nullspace)
__global__ void k_nullsum(Dims d, LV L, const double *__restrict__ x, const double *__restrict__ nvert, double *part)
{
    __shared__ double sh[4];
    double s = 0, c = 0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < L.n; t += (long long)gridDim.x * blockDim.x) {
        int i = (int)(t % L.nx), j = (int)((t / L.nx) % L.ny), k = (int)(t / ((long long)L.nx * L.ny));
        if (nvert[IX(d, k + 1, j + 1, i + 1)] < 0.1) { s += x[t]; c += 1.0; }
    }
    double r0 = block_sum<256>(s, sh), r1 = block_sum<256>(c, sh);
    if (threadIdx.x == 0) { part[blockIdx.x] = r0; part[VFS_RED_BLOCKS + blockIdx.x] = r1; }
}

__global__ void k_neg_mean(double *v) { v[0] = -v[0] / v[1]; }

__global__ void k_set_scalar(double *p, double v) { p[0] = v; }

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <typename T> static T *vec_of(MGLevel &l, int which);
template <> double *vec_of<double>(MGLevel &l, int w)
{
    return w == 0 ? l.b64 : w == 1 ? l.x64 : w == 2 ? l.y64 : w == 3 ? l.d64 : l.e64;
}
template <> float *vec_of<float>(MGLevel &l, int w)
{
    return w == 0 ? l.b32 : w == 1 ? l.x32 : w == 2 ? l.y32 : w == 3 ? l.d32 : l.e32;
}
template <typename T> static const T *A_of(const MGLevel &l);
template <> const double *A_of<double>(const MGLevel &l) { return l.A64; }
template <> const float *A_of<float>(const MGLevel &l) { return l.A32; }
template <typename T> static void fac_of(const MGLevel &l, const T *&a, const T *&u, const T *&i);
template <> void fac_of<double>(const MGLevel &l, const double *&a, const double *&u, const double *&i) { a = l.tl64; u = l.tu64; i = l.ti64; }
template <> void fac_of<float>(const MGLevel &l, const float *&a, const float *&u, const float *&i) { a = l.tl32; u = l.tu32; i = l.ti32; }

static void free_level(MGLevel &l)
{
    hipFree(l.A64); hipFree(l.A32); hipFree(l.tl64); hipFree(l.tu64); hipFree(l.ti64);
    hipFree(l.tl32); hipFree(l.tu32); hipFree(l.ti32);
    hipFree(l.b64); hipFree(l.x64); hipFree(l.y64); hipFree(l.d64); hipFree(l.e64);
    hipFree(l.b32); hipFree(l.x32); hipFree(l.y32); hipFree(l.d32); hipFree(l.e32);
}

int vfs_poisson_destroy(vfs_ctx *c)
{
    Poisson *ps = c->pois;
    if (!ps) return 0;
    hipFree(ps->gid); hipFree(ps->blocksum); hipFree(ps->total); hipFree(ps->coef); hipFree(ps->coef32);
    hipFree(ps->rowflag);
    for (auto &l : ps->lv) free_level(l);
    hipFree(ps->cinv64); hipFree(ps->cinv32); hipFree(ps->cidx);
    hipFree(ps->cinvr64); hipFree(ps->cinvr32);
    hipFree(ps->V64); hipFree(ps->V32); hipFree(ps->w64); hipFree(ps->z64); hipFree(ps->r64); hipFree(ps->x64);
    hipFree(ps->bb64); hipFree(ps->e64); hipFree(ps->w32); hipFree(ps->z32); hipFree(ps->x32); hipFree(ps->bb32);
    hipFree(ps->dots);
    if (ps->hdots) hipHostFree(ps->hdots);
    delete ps;
    c->pois = nullptr;
    return 0;
}

#define DMALLOC(ptr, bytes)                                                         \
    do {                                                                            \
        if (hipMalloc((void **)&(ptr), (bytes) ? (bytes) : 8) != hipSuccess) {      \
            vfs_set_error("hipMalloc failed in the Poisson set-up");               \
            return -1;                                                              \
        }                                                                           \
        hipMemsetAsync((ptr), 0, (bytes) ? (bytes) : 8, c->stream);                 \
    } while (0)

int vfs_poisson_alloc(vfs_ctx *c)
{
    if (c->pois) vfs_poisson_destroy(c);
    Poisson *ps = new Poisson();
    c->pois = ps;
    const size_t P = (size_t)c->d.P;
    DMALLOC(ps->gid, P * sizeof(double));
    int nb = (int)grid_for((long long)P, 1024);
    DMALLOC(ps->blocksum, nb * sizeof(int));
    DMALLOC(ps->total, sizeof(int));
    DMALLOC(ps->coef, 19 * P * sizeof(double));
    DMALLOC(ps->coef32, 19 * P * sizeof(float));
    DMALLOC(ps->rowflag, P);
    // line-Jacobi damping: 0.9 took 25 GMRES iterations to 1e-8 on the
    // 256x128x128 channel against 26 for 0.8; 1.0 does not converge
    ps->omega = VFS_MG_OMEGA;
    ps->alpha = VFS_MG_ALPHA;
    ps->alpha_semi = VFS_MG_ALPHA_SEMI;
    ps->wdepth = VFS_MG_WDEPTH;
    ps->m = 40;
    return 0;
}

static int alloc_level(vfs_ctx *c, MGLevel &l)
{
    const size_t n = (size_t)l.n;
    DMALLOC(l.A64, (size_t)l.nslot * n * sizeof(double));
    DMALLOC(l.A32, (size_t)l.nslot * n * sizeof(float));
    DMALLOC(l.tl64, n * sizeof(double)); DMALLOC(l.tu64, n * sizeof(double)); DMALLOC(l.ti64, n * sizeof(double));
    DMALLOC(l.tl32, n * sizeof(float)); DMALLOC(l.tu32, n * sizeof(float)); DMALLOC(l.ti32, n * sizeof(float));
    DMALLOC(l.b64, n * sizeof(double)); DMALLOC(l.x64, n * sizeof(double));
    DMALLOC(l.y64, n * sizeof(double)); DMALLOC(l.d64, n * sizeof(double)); DMALLOC(l.e64, n * sizeof(double));
    DMALLOC(l.b32, n * sizeof(float)); DMALLOC(l.x32, n * sizeof(float));
    DMALLOC(l.y32, n * sizeof(float)); DMALLOC(l.d32, n * sizeof(float)); DMALLOC(l.e32, n * sizeof(float));
    return 0;
}

// Gauss-Jordan inverse with partial pivoting of the dense coarsest operator,
// on the augmented matrix [M | I] (row-major, n rows of 2n), one column at a
// time: k_gj_pivot then k_gj_update.  Columns left of the pivot column are
// zero in every row not yet used as a pivot, so both kernels start there.
//
// Pivot of column col (one workgroup): the largest |a[r][col]|, r >= col,
// the lowest such r on ties (what a sequential scan picks), is swapped into
// row col, which is then scaled by 1/pivot; f[r] = a[r][col] is kept for
// the update (f[col] = 0).  A zero or NaN pivot sets *bad, after which every
// later launch returns at once.
__global__ void __launch_bounds__(1024) k_gj_pivot(double *__restrict__ a, int n, int col, double *__restrict__ f,
                                                   int *__restrict__ bad)
{
    __shared__ double sv[16];
    __shared__ int si[16];
    __shared__ int s_piv, s_bad;
    __shared__ double s_sc;
    const long long w = 2LL * n;
    if (threadIdx.x == 0) s_bad = *bad;
    __syncthreads();
    if (s_bad) return;
    double best = -1.0;
    int bi = n;
    for (int r = col + (int)threadIdx.x; r < n; r += 1024) {
        const double v = fabs(a[(long long)r * w + col]);
        if (v > best) { best = v; bi = r; }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(best, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
    }
    if ((threadIdx.x & 63) == 0) { sv[threadIdx.x >> 6] = best; si[threadIdx.x >> 6] = bi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = sv[0];
        int p = si[0];
        for (int q = 1; q < 16; q++)
            if (sv[q] > b || (sv[q] == b && si[q] < p)) { b = sv[q]; p = si[q]; }
        s_piv = p;
        s_bad = !(b > 0.0);
        if (s_bad) *bad = 1;
        else s_sc = 1.0 / a[(long long)p * w + col];   // read before any thread swaps
    }
    __syncthreads();
    if (s_bad) return;
    const int p = s_piv;
    double *rc = a + (long long)col * w, *rp = a + (long long)p * w;
    const double sc = s_sc;
    for (long long q = col + threadIdx.x; q < w; q += 1024) {
        const double x = rp[q];
        if (p != col) rp[q] = rc[q];
        rc[q] = x * sc;
    }
    __syncthreads();
    for (int r = threadIdx.x; r < n; r += 1024) f[r] = r == col ? 0.0 : a[(long long)r * w + col];
}

// a[r][q] -= f[r] * a[col][q] for q >= col; rows with f[r] == 0 unchanged
// grid: (ceil((2n - col) / 256), n)
__global__ void __launch_bounds__(256) k_gj_update(double *__restrict__ a, int n, int col,
                                                   const double *__restrict__ f, const int *__restrict__ bad)
{
    if (*bad) return;
    const int r = blockIdx.y;
    const long long w = 2LL * n, q = col + (long long)blockIdx.x * 256 + threadIdx.x;
    const double fr = f[r];
    if (q >= w || fr == 0.0) return;
    a[(long long)r * w + q] -= fr * a[(long long)col * w + q];
}

// The inverse (right half of the augmented matrix) in the layouts of the
// coarse solves: column-major nc x nc (cinv, rows of a solve one per thread)
// and scattered to all ncell cells of the level (cinvr, row-major, zero
// elsewhere), FP64 and FP32.
__global__ void __launch_bounds__(256) k_gj_out(const double *__restrict__ a, int n, const long long *__restrict__ cidx,
                                                long long ncell, double *__restrict__ cinv64, float *__restrict__ cinv32,
                                                double *__restrict__ cinvr64, float *__restrict__ cinvr32)
{
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long long)n * n) return;
    const int r = (int)(t / n), q = (int)(t % n);
    const double v = a[(long long)r * 2 * n + n + q];
    cinv64[(long long)q * n + r] = v;
    cinv32[(long long)q * n + r] = (float)v;
    const long long o = cidx[r] * ncell + cidx[q];
    cinvr64[o] = v;
    cinvr32[o] = (float)v;
}

// sums of |coefficient| of the i and of the k neighbours of a level: the
// couplings that decide whether k is coarsened.  One partial per workgroup
// at part[b] (i) and part[VFS_RED_BLOCKS + b] (k); k_reduce_finish adds them
// in a fixed order, so the decision does not depend on the schedule.
__global__ void __launch_bounds__(256) k_axis_strength(LV L, const double *__restrict__ A, double *part)
{
    __shared__ double sh[4];
    const int i1 = L.nslot == 19 ? 1 : 14, i2 = L.nslot == 19 ? 2 : 12;
    const int k1 = L.nslot == 19 ? 5 : 22, k2 = L.nslot == 19 ? 6 : 4;
    double a = 0, b = 0;
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < L.n; t += (long long)gridDim.x * 256) {
        a += fabs(A[(long long)i1 * L.n + t]) + fabs(A[(long long)i2 * L.n + t]);
        b += fabs(A[(long long)k1 * L.n + t]) + fabs(A[(long long)k2 * L.n + t]);
    }
    a = block_sum<256>(a, sh);
    b = block_sum<256>(b, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = a;
        part[VFS_RED_BLOCKS + blockIdx.x] = b;
    }
}

// k is kept while its coupling is below this share of the i coupling.  On
// the 768x384x48 turbine slab (dz about 8 dx) with the over-corrected
// coarse correction, GMRES to 1e-8 took 113 iterations at 0 (k always
// coarsened), 53 at 1/16, 37 at 1/8, 41 at 1/4 and 43 at 1/2.
#ifndef VFS_MG_KEEPK
#define VFS_MG_KEEPK 0.125
#endif

int vfs_mg_setup(vfs_ctx *c)
{
    Poisson *ps = c->pois;
    const Dims d = c->d;
    ps->lv.clear();
    MGLevel l0;
    std::memset(&l0, 0, sizeof(l0));
    l0.nx = d.mx - 2; l0.ny = d.my - 2; l0.nz = d.mz - 2;
    l0.n = (long long)l0.nx * l0.ny * l0.nz;
    l0.per_i = d.ip; l0.per_j = d.jp; l0.per_k = d.kp;
    l0.nslot = 19;
    l0.sk = 1;
    if (alloc_level(c, l0)) return -1;
    hipLaunchKernelGGL(k_gather_level0, dim3(grid_for(l0.n, 256)), dim3(256), 0, c->stream, d, lv_of(l0), ps->coef, l0.A64, l0.A32);
    VFS_CHECK_LAUNCH();
    ps->lv.push_back(l0);
    // coarsen until the level fits the dense solve (the one-workgroup solve
    // of the small-level kernel holds at most VFS_MAX_DENSE unknowns)
    const long long NCOARSE = VFS_MAX_DENSE;
    while (ps->lv.back().n > NCOARSE) {
        MGLevel &f = ps->lv.back();
        MGLevel cl;
        std::memset(&cl, 0, sizeof(cl));
        // semi-coarsening: k is kept while its couplings are under an eighth
        // of the i couplings (cells about three times as long in k as in i
        // or more), where 2x2x2 aggregates would join cells the smoother
        // leaves rough
        int keep_k = f.nz <= 1;
        if (!keep_k && f.nx > 1) {
            const int nb = (int)std::min<long long>(VFS_RED_BLOCKS, grid_for(f.n, 256));
            hipLaunchKernelGGL(k_axis_strength, dim3(nb), dim3(256), 0, c->stream, lv_of(f), f.A64, c->red);
            VFS_CHECK_LAUNCH();
            double *res = c->red + (long long)VFS_RED_SLOTS * VFS_RED_BLOCKS;
            if (vfs_reduce_finish_slots(c, 0, nb, 2, 0, res)) return -1;
            double h[2];
            if (vfs_read_scalars(c, res, h, 2)) return -1;
            keep_k = h[1] < VFS_MG_KEEPK * h[0];
        }
        cl.sk = keep_k ? 1 : 2;
        cl.nx = f.nx > 1 ? (f.nx + 1) / 2 : 1;
        cl.ny = f.ny > 1 ? (f.ny + 1) / 2 : 1;
        cl.nz = keep_k ? f.nz : (f.nz + 1) / 2;
        cl.n = (long long)cl.nx * cl.ny * cl.nz;
        cl.per_i = f.per_i; cl.per_j = f.per_j; cl.per_k = f.per_k;
        cl.nslot = 27;
        if (cl.n == f.n) break;
        if (alloc_level(c, cl)) return -1;
        LV F = lv_of(f), C = lv_of(cl);
        hipLaunchKernelGGL(k_galerkin, dim3(grid_for(cl.n, 128)), dim3(128), 0, c->stream, F, C, f.A64, cl.A64, cl.A32);
        VFS_CHECK_LAUNCH();
        ps->lv.push_back(cl);
    }
    for (auto &l : ps->lv) {
        LV L = lv_of(l);
        hipLaunchKernelGGL(k_line_factors, dim3(grid_for((long long)l.nx * l.nz, 128)), dim3(128), 0, c->stream, L,
                           l.A64, l.tl64, l.tu64, l.ti64, l.tl32, l.tu32, l.ti32);
        VFS_CHECK_LAUNCH();
    }
    // dense coarsest operator: regularised by 1 1^T / n so the singular
    // Neumann/periodic system has the mean-free solution
    MGLevel &cl = ps->lv.back();
    std::vector<double> A((size_t)cl.nslot * cl.n);
    VFS_CHECK_HIP(hipMemcpyAsync(A.data(), cl.A64, A.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    VFS_CHECK_HIP(hipStreamSynchronize(c->stream));
    std::vector<long long> list;
    std::vector<int> pos(cl.n, -1);
    const int center = cl.nslot == 19 ? 0 : 13;
    for (long long q = 0; q < cl.n; q++)
        if (A[(size_t)center * cl.n + q] != 0.0) { pos[q] = (int)list.size(); list.push_back(q); }
    int nc = (int)list.size();
    VFS_REQUIRE(nc > 0, "no Poisson unknowns");
    std::vector<double> M((size_t)nc * nc, 1.0 / nc);
    LV L = lv_of(cl);
    for (int r = 0; r < nc; r++) {
        long long q = list[r];
        int i = (int)(q % L.nx), j = (int)((q / L.nx) % L.ny), k = (int)(q / ((long long)L.nx * L.ny));
        for (int s = 0; s < cl.nslot; s++) {
            double a = A[(size_t)s * cl.n + q];
            if (a == 0.0) continue;
            int di, dj, dk;
            if (cl.nslot == 19) {
                static const int t19[19][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
                                               {1, 1, 0}, {1, -1, 0}, {-1, 1, 0}, {-1, -1, 0}, {0, 1, 1}, {0, 1, -1}, {0, -1, 1},
                                               {0, -1, -1}, {1, 0, 1}, {1, 0, -1}, {-1, 0, 1}, {-1, 0, -1}};
                di = t19[s][0]; dj = t19[s][1]; dk = t19[s][2];
            } else { dk = s / 9 - 1; dj = (s / 3) % 3 - 1; di = s % 3 - 1; }
            int ni = i + di, nj = j + dj, nk = k + dk;
            auto wrap = [](int &v, int n, int per) -> bool {
                if (v < 0) { if (!per) return false; v += n; } else if (v >= n) { if (!per) return false; v -= n; }
                return true;
            };
            if (!wrap(ni, L.nx, L.pi) || !wrap(nj, L.ny, L.pj) || !wrap(nk, L.nz, L.pk)) continue;
            long long nbq = ((long long)nk * L.ny + nj) * L.nx + ni;
            if (pos[nbq] < 0) continue;
            M[(size_t)r * nc + pos[nbq]] += a;
        }
    }
    ps->nc = nc;
    // the first level (below the finest) small enough for the one-workgroup cycle
    ps->lsmall = (int)ps->lv.size();
    for (int q = 1; q < (int)ps->lv.size(); q++)
        if (ps->lv[q].n <= 2048 && (int)ps->lv.size() - q <= MG_MAXL) { ps->lsmall = q; break; }
    DMALLOC(ps->cinv64, (size_t)nc * nc * sizeof(double));
    DMALLOC(ps->cinv32, (size_t)nc * nc * sizeof(float));
    DMALLOC(ps->cidx, (size_t)nc * sizeof(long long));
    DMALLOC(ps->cinvr64, (size_t)cl.n * cl.n * sizeof(double));
    DMALLOC(ps->cinvr32, (size_t)cl.n * cl.n * sizeof(float));
    {
        // (M | I) inverted on the device (2 nc launches of a few microseconds)
        std::vector<double> aug((size_t)nc * 2 * nc, 0.0);
        for (int r = 0; r < nc; r++) {
            std::memcpy(&aug[(size_t)r * 2 * nc], &M[(size_t)r * nc], (size_t)nc * sizeof(double));
            aug[(size_t)r * 2 * nc + nc + r] = 1.0;
        }
        double *da = nullptr, *df = nullptr;
        int *dbad = nullptr;
        auto release = [&]() { hipFree(da); hipFree(df); hipFree(dbad); };
        if (hipMalloc((void **)&da, aug.size() * sizeof(double)) != hipSuccess ||
            hipMalloc((void **)&df, (size_t)nc * sizeof(double)) != hipSuccess ||
            hipMalloc((void **)&dbad, sizeof(int)) != hipSuccess) {
            release();
            vfs_set_error("hipMalloc failed in the Poisson set-up");
            return -1;
        }
        int bad = 0;
        if (hipMemcpyAsync(da, aug.data(), aug.size() * sizeof(double), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            hipMemsetAsync(dbad, 0, sizeof(int), c->stream) != hipSuccess ||
            hipMemcpyAsync(ps->cidx, list.data(), list.size() * sizeof(long long), hipMemcpyHostToDevice,
                           c->stream) != hipSuccess) {
            release();
            vfs_set_error("copy to the device failed in the Poisson set-up");
            return -1;
        }
        for (int col = 0; col < nc; col++) {
            hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(1024), 0, c->stream, da, nc, col, df, dbad);
            hipLaunchKernelGGL(k_gj_update, dim3(grid_for(2LL * nc - col, 256), nc), dim3(256), 0, c->stream, da, nc,
                               col, df, dbad);
        }
        hipLaunchKernelGGL(k_gj_out, dim3(grid_for((long long)nc * nc, 256)), dim3(256), 0, c->stream, da, nc,
                           ps->cidx, cl.n, ps->cinv64, ps->cinv32, ps->cinvr64, ps->cinvr32);
        const hipError_t le = hipGetLastError();
        const hipError_t ce = hipMemcpyAsync(&bad, dbad, sizeof(int), hipMemcpyDeviceToHost, c->stream);
        const hipError_t se = hipStreamSynchronize(c->stream);
        release();
        if (le != hipSuccess || ce != hipSuccess || se != hipSuccess) {
            vfs_set_error("the coarsest Poisson inverse failed on the device");
            return -1;
        }
        if (bad) { vfs_set_error("singular coarsest Poisson operator"); return -1; }
    }
    // Krylov space
    const size_t n0 = (size_t)ps->lv[0].n;
    DMALLOC(ps->V64, (size_t)(ps->m + 1) * n0 * sizeof(double));
    DMALLOC(ps->V32, (size_t)(ps->m + 1) * n0 * sizeof(float));
    DMALLOC(ps->w64, n0 * sizeof(double)); DMALLOC(ps->z64, n0 * sizeof(double)); DMALLOC(ps->r64, n0 * sizeof(double));
    DMALLOC(ps->x64, n0 * sizeof(double)); DMALLOC(ps->bb64, n0 * sizeof(double)); DMALLOC(ps->e64, n0 * sizeof(double));
    DMALLOC(ps->w32, n0 * sizeof(float)); DMALLOC(ps->z32, n0 * sizeof(float)); DMALLOC(ps->x32, n0 * sizeof(float));
    DMALLOC(ps->bb32, n0 * sizeof(float));
    DMALLOC(ps->dots, 512 * sizeof(double));
    VFS_CHECK_HIP(hipHostMalloc((void **)&ps->hdots, 512 * sizeof(double), hipHostMallocDefault));
    VFS_CHECK_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

// launch y = A x / r = b - A x with the level's slot count
static inline int row_block(int nx) { return nx >= 256 ? 256 : ((nx + 63) / 64) * 64; }
static inline dim3 row_grid(const MGLevel &L) { int b = row_block(L.nx); return dim3((L.nx + b - 1) / b, L.ny, L.nz); }

template <typename T>
static int launch_spmv(vfs_ctx *c, const MGLevel &L, const T *A, const T *x, T *y)
{
    LV Lv = lv_of(L);
    dim3 gb = row_grid(L), bb(row_block(L.nx));
    if (L.nslot == 19) hipLaunchKernelGGL((k_spmv<T, 19>), gb, bb, 0, c->stream, Lv, A, x, y);
    else hipLaunchKernelGGL((k_spmv<T, 27>), gb, bb, 0, c->stream, Lv, A, x, y);
    VFS_CHECK_LAUNCH();
    return 0;
}

template <typename T>
static int launch_resid(vfs_ctx *c, const MGLevel &L, const T *A, const T *x, const T *b, T *r)
{
    LV Lv = lv_of(L);
    dim3 gb = row_grid(L), bb(row_block(L.nx));
    if (L.nslot == 19) hipLaunchKernelGGL((k_resid<T, 19>), gb, bb, 0, c->stream, Lv, A, x, b, r);
    else hipLaunchKernelGGL((k_resid<T, 27>), gb, bb, 0, c->stream, Lv, A, x, b, r);
    VFS_CHECK_LAUNCH();
    return 0;
}

template <typename T, int B>
static void launch_thomas_b(vfs_ctx *c, dim3 gl, LV Lv, const T *r, const T *ti, const T *tm, const T *tu, const T *xo,
                            T *dd, T *xn)
{
    const T omega = (T)c->pois->omega;
    if (xo) hipLaunchKernelGGL((k_thomas<T, B, false>), gl, dim3(64), 0, c->stream, Lv, r, ti, tm, tu, xo, omega, dd, xn);
    else hipLaunchKernelGGL((k_thomas<T, B, true>), gl, dim3(64), 0, c->stream, Lv, r, ti, tm, tu, xo, omega, dd, xn);
}

// xn = xo + omega T^{-1} r on level L (xo null: zero guess)
template <typename T>
static int launch_thomas(vfs_ctx *c, const MGLevel &L, const T *r, const T *xo, T *xn)
{
    const T *tm, *tu, *ti;
    fac_of<T>(L, tm, tu, ti);
    T *dd = vec_of<T>(const_cast<MGLevel &>(L), 3);
    dim3 gl((L.nx + 63) / 64, L.nz);
    launch_thomas_b<T, 8>(c, gl, lv_of(L), r, ti, tm, tu, xo, dd, xn);
    VFS_CHECK_LAUNCH();
    return 0;
}

// One multigrid cycle on level l: x = M_l^{-1} b from x = 0 (zero_guess), or
// one cycle started from the given x.  Per level: line smoothing, the
// restricted residual, the coarse cycle (twice on levels l < wdepth: a
// W-cycle there, since a V-cycle loses about one digit per added level on
// the piecewise-constant aggregation hierarchy), the residual of the
// corrected iterate, and line smoothing with the prolongation folded in.
// Scratch: the level's y and d vectors and the coarser levels' b and x.
// the coarse-correction scale into level l, by the shape of the aggregates of level l + 1
static double mg_alpha(const Poisson *ps, int l)
{
    return l + 1 < (int)ps->lv.size() && ps->lv[l + 1].sk == 1 ? ps->alpha_semi : ps->alpha;
}

template <typename T>
static int vcycle(vfs_ctx *c, int l, const T *b, T *x, bool zero_guess)
{
    Poisson *ps = c->pois;
    MGLevel &L = ps->lv[l];
    LV Lv = lv_of(L);
    const T alpha = (T)mg_alpha(ps, l);
    T *y = vec_of<T>(L, 2);
    const T *A = A_of<T>(L);
    if (l == (int)ps->lv.size() - 1) {
        // dense solve: exact, whatever the starting guess
        const T *cinvr = std::is_same<T, double>::value ? (const T *)ps->cinvr64 : (const T *)ps->cinvr32;
        hipLaunchKernelGGL(k_dense_mv<T>, dim3((unsigned)((L.n + 3) / 4)), dim3(256), 0, c->stream, L.n, cinvr, b, x);
        VFS_CHECK_LAUNCH();
        return 0;
    }
    if (l == ps->lsmall) {
        SmallLevels<T> S;
        S.first = l;
        S.last = (int)ps->lv.size() - 1;
        for (int q = l; q <= S.last; q++) {
            MGLevel &Q = ps->lv[q];
            S.lv[q] = lv_of(Q);
            S.A[q] = A_of<T>(Q);
            fac_of<T>(Q, S.tl[q], S.tu[q], S.ti[q]);
            S.b[q] = vec_of<T>(Q, 0); S.x[q] = vec_of<T>(Q, 1); S.y[q] = vec_of<T>(Q, 2); S.d[q] = vec_of<T>(Q, 3);
        }
        S.b[l] = const_cast<T *>(b);
        S.x[l] = x;
        S.nc = ps->nc;
        S.cidx = ps->cidx;
        S.cinv = std::is_same<T, double>::value ? (const T *)ps->cinv64 : (const T *)ps->cinv32;
        S.omega = (T)ps->omega;
        for (int q = l; q < S.last; q++) S.alpha[q] = (T)mg_alpha(ps, q);
        S.zero = zero_guess ? 1 : 0;
        hipLaunchKernelGGL(k_small_cycle<T>, dim3(1), dim3(1024), 0, c->stream, S);
        VFS_CHECK_LAUNCH();
        return 0;
    }
    MGLevel &C = ps->lv[l + 1];
    LV Cv = lv_of(C);
    T *cb = vec_of<T>(C, 0), *cx = vec_of<T>(C, 1);
    // pre-smoothing
    if (zero_guess) {
        if (launch_thomas<T>(c, L, b, nullptr, x)) return -1;
    } else {
        if (launch_resid<T>(c, L, A, x, b, y)) return -1;
        if (launch_thomas<T>(c, L, y, x, x)) return -1;
    }
    {
        const int bs = row_block(L.nx);
        dim3 g((L.nx + bs - 1) / bs, C.ny, C.nz);
        if (L.nslot == 19)
            hipLaunchKernelGGL((k_resid_restrict<T, 19>), g, dim3(bs), 0, c->stream, Lv, Cv, A, (const T *)x, b, cb);
        else
            hipLaunchKernelGGL((k_resid_restrict<T, 27>), g, dim3(bs), 0, c->stream, Lv, Cv, A, (const T *)x, b, cb);
        VFS_CHECK_LAUNCH();
    }
    if (vcycle<T>(c, l + 1, cb, cx, true)) return -1;
    if (l < ps->wdepth && l + 2 < (int)ps->lv.size())
        if (vcycle<T>(c, l + 1, cb, cx, false)) return -1;
    // post-smoothing of x + alpha P e.  On the 19-point fine level a
    // prolongation pass and the plain residual (the coarse reads at every
    // stencil point cost more there than the pass); on the 27-point levels,
    // where launches dominate, one kernel forms both x + alpha P e and its
    // residual.
    if (L.nslot == 19) {
        hipLaunchKernelGGL(k_prolong<T>, row_grid(L), dim3(row_block(L.nx)), 0, c->stream, Lv, Cv, (const T *)cx, alpha, x);
        VFS_CHECK_LAUNCH();
        if (launch_resid<T>(c, L, A, x, b, y)) return -1;
        return launch_thomas<T>(c, L, y, x, x);
    }
    T *xe = vec_of<T>(L, 4);
    hipLaunchKernelGGL((k_resid_pe<T, 27>), row_grid(L), dim3(row_block(L.nx)), 0, c->stream, Lv, Cv, A, (const T *)x,
                       (const T *)cx, alpha, b, y, xe);
    VFS_CHECK_LAUNCH();
    return launch_thomas<T>(c, L, y, xe, x);
}

template <typename T> static T *kv(Poisson *ps, int w);
template <> double *kv<double>(Poisson *ps, int w) { return w == 0 ? ps->V64 : w == 1 ? ps->w64 : w == 2 ? ps->z64 : ps->x64; }
template <> float *kv<float>(Poisson *ps, int w) { return w == 0 ? ps->V32 : w == 1 ? ps->w32 : w == 2 ? ps->z32 : ps->x32; }

static int multidot_finish(vfs_ctx *c, int nv, int nb, double *dst)
{
    return vfs_reduce_finish(c, nb, nv, 0, dst);
}

// The launches of Arnoldi step kk up to the first orthogonalisation pass:
// xl = M^{-1} U_kk, w = A xl, dots of w with U_0..U_kk and |w|^2,
// U_{kk+1} = w - sum_q U_q d_q / |U_q|^2 and its norm (all on the device).
template <typename T>
static int arnoldi_enqueue(vfs_ctx *c, int kk, const T *uk, T *un, T *xl, T *w)
{
    Poisson *ps = c->pois;
    MGLevel &L0 = ps->lv[0];
    const long long n = L0.n;
    const T *V = kv<T>(ps, 0);
    const int nbr = (int)std::min<long long>(VFS_RED_BLOCKS, grid_for(n, 256));
    double *nrm = ps->dots + 256;
    if (vcycle<T>(c, 0, uk, xl, true)) return -1;
    if (launch_spmv<T>(c, L0, A_of<T>(L0), (const T *)xl, w)) return -1;
    for (int v0 = 0; v0 <= kk; v0 += 8) {
        const int nv = std::min(8, kk + 1 - v0);
        launch_multidot8<T>(c->stream, nbr, n, V, n, v0, nv, (int)(v0 == 0), (const T *)w, c->red);
        VFS_CHECK_LAUNCH();
    }
    if (multidot_finish(c, kk + 1, nbr, ps->dots)) return -1;
    if (vfs_reduce_finish_slots(c, VFS_RED_SLOTS - 1, nbr, 1, 0, ps->dots + 130)) return -1;
    hipLaunchKernelGGL(k_multiaxpy<T>, dim3(nbr), dim3(256), 0, c->stream, n, V, n, kk + 1, (const double *)ps->dots,
                       (const double *)nrm, (const T *)w, un, c->red);
    VFS_CHECK_LAUNCH();
    return vfs_reduce_finish_slots(c, VFS_RED_SLOTS - 1, nbr, 1, 0, nrm + kk + 1);
}

// Right-preconditioned restarted GMRES in precision T on A_T e = rhs from
// e = 0 (rhs: level-0 compact, precision T), with classical Gram-Schmidt
// (a second pass only on heavy cancellation, see below).
// The basis vectors U_q are kept unnormalised, |U_q|^2 in nrm (device) and
// hnrm (host), and the Hessenberg entries are scaled on the host:
//   H[q][k] = (d1_q + d2_q) / (|U_q| |U_k|),  H[k+1][k] = |U_{k+1}| / |U_k|,
// where d1, d2 are the passes' dots of U_q with w = A M^{-1} U_k.  The
// orthogonalisation writes U_{k+1} and returns its norm in the same pass, so
// a new basis vector costs no normalisation pass.  One host synchronisation
// per pass.  Stops when the estimated residual drops below tol * |rhs| or
// after maxit iterations; the result is in kv(..., 3).
template <typename T>
static int gmres(vfs_ctx *c, const T *rhs, double tol, int maxit, int *iters, double *est)
{
    Poisson *ps = c->pois;
    MGLevel &L0 = ps->lv[0];
    const long long n = L0.n;
    const int m = ps->m;
    VFS_REQUIRE(m + 1 <= 64, "GMRES restart length above 63");
    T *V = kv<T>(ps, 0), *w = kv<T>(ps, 1), *z = kv<T>(ps, 2), *x = kv<T>(ps, 3);
    T *xl = vec_of<T>(L0, 1);   // preconditioner output
    const T *A = A_of<T>(L0);
    unsigned gb = grid_for(n, 256);
    const int nbr = (int)std::min<long long>(VFS_RED_BLOCKS, grid_for(n, 256));
    double *one = c->red + VFS_RED_SLOTS * VFS_RED_BLOCKS + 24;   // device constant 1.0
    double *nrm = ps->dots + 256, *hnrm = ps->hdots + 256;
    VFS_CHECK_HIP(hipMemsetAsync(x, 0, n * sizeof(T), c->stream));
    // U_0 = rhs (x = 0)
    VFS_CHECK_HIP(hipMemcpyAsync(V, rhs, n * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
    hipLaunchKernelGGL(k_multidot<T>, dim3(nbr), dim3(256), 0, c->stream, n, (const T *)V, n, 1, (const T *)V, c->red);
    VFS_CHECK_LAUNCH();
    if (multidot_finish(c, 1, nbr, nrm)) return -1;
    VFS_CHECK_HIP(hipMemcpyAsync(hnrm, nrm, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    VFS_CHECK_HIP(hipStreamSynchronize(c->stream));
    const double bnorm = sqrt(hnrm[0]);
    int it = 0;
    if (bnorm == 0.0) { *iters = 0; *est = 0; return 0; }
    hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, c->stream, one, 1.0);
    VFS_CHECK_LAUNCH();
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), s(m + 1), y(m);
    double res = 1.0;
    bool first = true;
    while (it < maxit) {
        s.assign(m + 1, 0.0);
        if (first) {
            s[0] = bnorm;
        } else {
            // U_0 = A x - rhs = -r, and |r|
            if (launch_spmv<T>(c, L0, A, (const T *)x, V)) return -1;
            hipLaunchKernelGGL(k_multiaxpy<T>, dim3(nbr), dim3(256), 0, c->stream, n, rhs, n, 1, (const double *)one,
                               (const double *)nullptr, (const T *)V, V, c->red);
            VFS_CHECK_LAUNCH();
            if (vfs_reduce_finish_slots(c, VFS_RED_SLOTS - 1, nbr, 1, 0, nrm)) return -1;
            VFS_CHECK_HIP(hipMemcpyAsync(hnrm, nrm, sizeof(double), hipMemcpyDeviceToHost, c->stream));
            VFS_CHECK_HIP(hipStreamSynchronize(c->stream));
            const double beta = sqrt(hnrm[0]);
            res = beta / bnorm;
            if (res <= tol) break;
            s[0] = -beta;
        }
        first = false;
        int kk = 0;
        for (; kk < m && it < maxit; kk++, it++) {
            // w = A M^{-1} U_kk
            T *uk = V + (size_t)kk * n, *un = V + (size_t)(kk + 1) * n;
            if (arnoldi_enqueue<T>(c, kk, uk, un, xl, w)) return -1;
            // one classical Gram-Schmidt pass; a second one only when the first
            // cancelled so much of w that its orthogonality error, about
            // u |w| / |w_perp|, would exceed u / sqrt(tau) (tau 1e-12 in FP64,
            // 1e-4 in FP32: errors below 1e-10 and 1e-5, far under the
            // tolerances the two precisions are asked for)
            const double tau = std::is_same<T, double>::value ? 1e-12 : 1e-4;
            double h1[64], h2[64];
            VFS_CHECK_HIP(hipMemcpyAsync(ps->hdots, ps->dots, (256 + kk + 2) * sizeof(double), hipMemcpyDeviceToHost,
                                         c->stream));
            VFS_CHECK_HIP(hipStreamSynchronize(c->stream));
            for (int q = 0; q <= kk; q++) { h1[q] = ps->hdots[q]; h2[q] = 0; }
            if (hnrm[kk + 1] < tau * ps->hdots[130]) {
                // second pass, in place on U_{kk+1}
                for (int v0 = 0; v0 <= kk; v0 += 8) {
                    const int nv = std::min(8, kk + 1 - v0);
                    launch_multidot8<T>(c->stream, nbr, n, (const T *)V, n, v0, nv, 0, (const T *)un, c->red);
                    VFS_CHECK_LAUNCH();
                }
                if (multidot_finish(c, kk + 1, nbr, ps->dots + 64)) return -1;
                hipLaunchKernelGGL(k_multiaxpy<T>, dim3(nbr), dim3(256), 0, c->stream, n, (const T *)V, n, kk + 1,
                                   (const double *)(ps->dots + 64), (const double *)nrm, (const T *)un, un, c->red);
                VFS_CHECK_LAUNCH();
                if (vfs_reduce_finish_slots(c, VFS_RED_SLOTS - 1, nbr, 1, 0, nrm + kk + 1)) return -1;
                VFS_CHECK_HIP(hipMemcpyAsync(ps->hdots, ps->dots, (256 + kk + 2) * sizeof(double),
                                             hipMemcpyDeviceToHost, c->stream));
                VFS_CHECK_HIP(hipStreamSynchronize(c->stream));
                for (int q = 0; q <= kk; q++) h2[q] = ps->hdots[64 + q];
            }
            const double nk = sqrt(hnrm[kk]);
            for (int q = 0; q <= kk; q++) H[(size_t)q * m + kk] = (h1[q] + h2[q]) / (sqrt(hnrm[q]) * nk);
            H[(size_t)(kk + 1) * m + kk] = sqrt(hnrm[kk + 1]) / nk;
            for (int q = 0; q < kk; q++) {
                double t0 = cs[q] * H[(size_t)q * m + kk] + sn[q] * H[(size_t)(q + 1) * m + kk];
                H[(size_t)(q + 1) * m + kk] = -sn[q] * H[(size_t)q * m + kk] + cs[q] * H[(size_t)(q + 1) * m + kk];
                H[(size_t)q * m + kk] = t0;
            }
            double a = H[(size_t)kk * m + kk], bb = H[(size_t)(kk + 1) * m + kk];
            double den = sqrt(a * a + bb * bb);
            cs[kk] = den > 0 ? a / den : 1.0;
            sn[kk] = den > 0 ? bb / den : 0.0;
            H[(size_t)kk * m + kk] = den;
            H[(size_t)(kk + 1) * m + kk] = 0;
            s[kk + 1] = -sn[kk] * s[kk];
            s[kk] = cs[kk] * s[kk];
            res = fabs(s[kk + 1]) / bnorm;
            if (res <= tol || hnrm[kk + 1] == 0.0) { kk++; it++; break; }
        }
        // x += M^{-1} sum_q U_q y_q / |U_q|
        for (int q = kk - 1; q >= 0; q--) {
            double t0 = s[q];
            for (int p = q + 1; p < kk; p++) t0 -= H[(size_t)q * m + p] * y[p];
            y[q] = t0 / H[(size_t)q * m + q];
        }
        for (int q = 0; q < kk; q++) ps->hdots[192 + q] = y[q] / sqrt(hnrm[q]);   // pinned staging
        VFS_CHECK_HIP(hipMemcpyAsync(ps->dots + 192, ps->hdots + 192, kk * sizeof(double), hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_combine<T>, dim3(gb), dim3(256), 0, c->stream, n, (const T *)V, n, kk,
                           (const double *)(ps->dots + 192), z);
        VFS_CHECK_LAUNCH();
        if (vcycle<T>(c, 0, (const T *)z, xl, true)) return -1;
        hipLaunchKernelGGL(k_axpy_dev<T>, dim3(gb), dim3(256), 0, c->stream, n, (const T *)xl, x);
        VFS_CHECK_LAUNCH();
        if (res <= tol) break;
    }
    *iters = it;
    *est = res;
    return 0;
}

static int norm2_64(vfs_ctx *c, const double *v, long long n, double *dev_out, double *host_out)
{
    int nbr = (int)std::min<long long>(VFS_RED_BLOCKS, grid_for(n, 256));
    hipLaunchKernelGGL(k_multidot<double>, dim3(nbr), dim3(256), 0, c->stream, n, v, n, 1, v, c->red);
    VFS_CHECK_LAUNCH();
    if (vfs_reduce_finish(c, nbr, 1, 0, dev_out)) return -1;
    if (host_out) return vfs_read_scalars(c, dev_out, host_out, 1);
    return 0;
}

extern "C" int vfs_poisson_solve(vfs_ctx *c, const double *b, double *phi, double rtol, int precision, int maxit,
                                 int *iters, double *relres)
{
    VFS_REQUIRE(c->pois && c->pois->ready, "vfs_poisson_setup has not been called");
    VFS_REQUIRE(precision == 0 || precision == 1, "precision must be 0 (FP64) or 1 (FP32 + FP64 refinement)");
    Poisson *ps = c->pois;
    const Dims d = c->d;
    MGLevel &L0 = ps->lv[0];
    LV Lv = lv_of(L0);
    const long long n = L0.n;
    unsigned gb = grid_for(n, 256);
    double *dres = c->red + VFS_RED_SLOTS * VFS_RED_BLOCKS + 32;
    // b and the initial guess on the compact layout
    hipLaunchKernelGGL(k_da_to_c, dim3(gb), dim3(256), 0, c->stream, d, Lv, b, ps->rowflag, ps->bb64);
    VFS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_da_to_c, dim3(gb), dim3(256), 0, c->stream, d, Lv, phi, ps->rowflag, ps->e64);
    VFS_CHECK_LAUNCH();
    double bn2;
    if (norm2_64(c, ps->bb64, n, dres, &bn2)) return -1;
    if (!std::isfinite(bn2)) {
        vfs_set_error("vfs_poisson_solve: the right-hand side is not finite (NaN or Inf)");
        return -6;
    }
    double bnorm = sqrt(bn2);
    int total = 0;
    double rel = 0;
    bool finite = true;
    if (bnorm == 0.0) bnorm = 1.0;
    for (int outer = 0; outer < 50; outer++) {
        // FP64 defect r = b - A x (FP64 operator)
        if (launch_resid<double>(c, L0, L0.A64, ps->e64, ps->bb64, ps->r64)) return -1;
        double rn2;
        if (norm2_64(c, ps->r64, n, dres, &rn2)) return -1;
        rel = sqrt(rn2) / bnorm;
        if (!std::isfinite(rn2)) { finite = false; break; }
        if (rel <= rtol || total >= maxit || rn2 == 0.0) break;
        int it = 0;
        double est = 0;
        if (precision == 0) {
            // FP64 GMRES on the defect to the remaining reduction
            double inner = std::max(rtol / rel, 1e-12) * 0.5;
            if (gmres<double>(c, ps->r64, inner, maxit - total, &it, &est)) return -1;
            hipLaunchKernelGGL(k_add_scaled<double>, dim3(gb), dim3(256), 0, c->stream, n, (const double *)ps->x64,
                               (const double *)nullptr, ps->e64);
            VFS_CHECK_LAUNCH();
        } else {
            // FP32 inner solve on the normalised defect, FP64 update
            hipLaunchKernelGGL(k_cast_scale<float>, dim3(gb), dim3(256), 0, c->stream, n, (const double *)ps->r64,
                               (const double *)dres, ps->bb32);
            VFS_CHECK_LAUNCH();
            double inner = std::max(rtol / rel * 0.5, 2e-5);
            if (gmres<float>(c, ps->bb32, inner, maxit - total, &it, &est)) return -1;
            hipLaunchKernelGGL(k_add_scaled<float>, dim3(gb), dim3(256), 0, c->stream, n, (const float *)ps->x32,
                               (const double *)dres, ps->e64);
            VFS_CHECK_LAUNCH();
        }
        total += it;
        if (it == 0) break;
    }
    // Remove_Nullspace over nvert < 0.1 cells, then Convert_Phi2_Phi
    int nbr = (int)std::min<long long>(VFS_RED_BLOCKS, grid_for(n, 256));
    hipLaunchKernelGGL(k_nullsum, dim3(nbr), dim3(256), 0, c->stream, d, Lv, ps->e64, c->nvert, c->red);
    VFS_CHECK_LAUNCH();
    if (vfs_reduce_finish(c, nbr, 2, 0, dres + 4)) return -1;
    hipLaunchKernelGGL(k_neg_mean, dim3(1), dim3(1), 0, c->stream, dres + 4);
    VFS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_c_to_da, dim3(grid_for(d.P, 256)), dim3(256), 0, c->stream, d, Lv, ps->e64, c->nvert, dres + 4, phi);
    VFS_CHECK_LAUNCH();
    if (vfs_periodic_sca(c, phi)) return -1;
    if (iters) *iters = total;
    if (relres) *relres = rel;
    if (!finite) {
        vfs_set_error("vfs_poisson_solve: the residual became non-finite after " + std::to_string(total) +
                      " iterations");
        return -6;
    }
    if (!(rel <= rtol)) {
        vfs_set_error("vfs_poisson_solve: not converged, relres " + std::to_string(rel) + " > rtol " +
                      std::to_string(rtol) + " after " + std::to_string(total) + " iterations");
        return 1;
    }
    return 0;
}
