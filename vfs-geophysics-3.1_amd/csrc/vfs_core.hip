// Context life cycle, device memory helpers, periodic ghost copies and the
// second pass of the deterministic reductions.
#include <cstring>
#include <string>
#include <vector>

#include "vfs_internal.h"
#include "vfs_reduce.h"

static thread_local std::string g_last_error;

void vfs_set_error(const std::string &msg) { g_last_error = msg; }

extern "C" const char *vfs_last_error(void) { return g_last_error.c_str(); }

// ---------------------------------------------------------------------------
// periodic ghost copies over the ghost shell only
// ---------------------------------------------------------------------------
__global__ void k_periodic_vec(Dims d, double *v)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int k, j, i;
    if (t >= shell_count(d) || !shell_point(d, t, k, j, i)) return;
    int a, b, c;
    if (!per_src(d, i, j, k, a, b, c)) return;
    long long s = IX(d, c, b, a), o = IX(d, k, j, i);
    v[3 * o] = v[3 * s]; v[3 * o + 1] = v[3 * s + 1]; v[3 * o + 2] = v[3 * s + 2];
}

__global__ void k_periodic_sca(Dims d, double *v)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int k, j, i;
    if (t >= shell_count(d) || !shell_point(d, t, k, j, i)) return;
    int a, b, c;
    if (!per_src(d, i, j, k, a, b, c)) return;
    v[IX(d, k, j, i)] = v[IX(d, c, b, a)];
}

int vfs_periodic_vec(vfs_ctx *c, double *v)
{
    if (!(c->d.ip || c->d.jp || c->d.kp)) return 0;
    long long n = shell_count(c->d);
    hipLaunchKernelGGL(k_periodic_vec, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, c->d, v);
    VFS_CHECK_LAUNCH();
    return 0;
}

int vfs_periodic_sca(vfs_ctx *c, double *v)
{
    if (!(c->d.ip || c->d.jp || c->d.kp)) return 0;
    long long n = shell_count(c->d);
    hipLaunchKernelGGL(k_periodic_sca, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, c->d, v);
    VFS_CHECK_LAUNCH();
    return 0;
}

// ---------------------------------------------------------------------------
// reduction finish: partials at red[v*VFS_RED_BLOCKS + b], b < nblocks
// ---------------------------------------------------------------------------
__global__ void k_reduce_finish(const double *part, int nblocks, int op, double *out)
{
    __shared__ double sh[4];
    const int v = blockIdx.x;
    double acc = op == 0 ? 0.0 : -1.0e308;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
        double x = part[(long long)v * VFS_RED_BLOCKS + b];
        acc = op == 0 ? acc + x : fmax(acc, x);
    }
    double r = op == 0 ? block_sum<256>(acc, sh) : block_max<256>(acc, sh);
    if (threadIdx.x == 0) out[v] = r;
}

int vfs_reduce_finish(vfs_ctx *c, int nblocks, int nvals, int op, double *dev_out)
{
    return vfs_reduce_finish_slots(c, 0, nblocks, nvals, op, dev_out);
}

int vfs_reduce_finish_slots(vfs_ctx *c, int first_slot, int nblocks, int nvals, int op, double *dev_out)
{
    hipLaunchKernelGGL(k_reduce_finish, dim3(nvals), dim3(256), 0, c->stream,
                       c->red + (long long)first_slot * VFS_RED_BLOCKS, nblocks, op, dev_out);
    VFS_CHECK_LAUNCH();
    return 0;
}

int vfs_read_scalars(vfs_ctx *c, const double *dev, double *host, int n)
{
    VFS_CHECK_HIP(hipMemcpyAsync(host, dev, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    VFS_CHECK_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
static int check_config(const vfs_config *cfg)
{
    if (cfg->mx < 7 || cfg->my < 7 || cfg->mz < 7) {
        vfs_set_error("grid too small: every direction needs at least 5 cells");
        return -2;
    }
    const int per[3] = {cfg->i_periodic, cfg->j_periodic, cfg->k_periodic};
    for (int f = 0; f < 6; f++) {
        int bt = cfg->bctype[f];
        if (per[f / 2]) {
            if (bt != 7) { vfs_set_error("periodic faces must carry bctype 7"); return -3; }
        } else if (bt != 1 && !(bt == -1 && f < 4) && !(bt == 5 && f == 4) && !(bt == 4 && f == 5)) {
            vfs_set_error("unsupported bctype (supported: 1 no-slip, 7 periodic, -1 wall function on i/j walls, "
                          "5 inlet at k = 0, 4 outlet at k = mz-1)");
            return -3;
        }
    }
    if (cfg->immersed != 0 && cfg->immersed != 1) { vfs_set_error("immersed must be 0 or 1"); return -4; }
    if (cfg->weno != 0 && cfg->weno != 1) { vfs_set_error("weno must be 0 or 1"); return -4; }
    if (cfg->inviscid != 0 && cfg->inviscid != 1) { vfs_set_error("inviscid must be 0 or 1"); return -4; }
    if (cfg->skew != 0 && cfg->skew != 1) { vfs_set_error("skew must be 0 or 1"); return -4; }
    if (cfg->clark != 0 && cfg->clark != 1) { vfs_set_error("clark must be 0 or 1"); return -4; }
    return 0;
}

extern "C" int vfs_ctx_create(const vfs_config *cfg, vfs_ctx **out)
{
    if (!cfg || !out) { vfs_set_error("null argument"); return -2; }
    int rc = check_config(cfg);
    if (rc) return rc;
    vfs_ctx *c = new vfs_ctx();
    std::memset((void *)c, 0, sizeof(vfs_ctx));
    c->cfg = *cfg;
    Dims &d = c->d;
    d.mx = cfg->mx; d.my = cfg->my; d.mz = cfg->mz;
    d.ip = cfg->i_periodic; d.jp = cfg->j_periodic; d.kp = cfg->k_periodic;
    for (int f = 0; f < 6; f++) d.bc[f] = cfg->bctype[f];
    d.immersed = cfg->immersed; d.les = cfg->les; d.second_order = cfg->second_order;
    d.dpdz_set = cfg->dpdz_set;
    d.weno = cfg->weno;
    d.inviscid = cfg->inviscid;
    d.skew = cfg->skew;
    d.clark = cfg->clark;
    d.ren = cfg->ren; d.dt = cfg->dt; d.st = cfg->st; d.mpg = cfg->mean_pressure_gradient;
    d.P = (long long)d.mx * d.my * d.mz;
    if (hipGetDevice(&c->device) != hipSuccess) { delete c; vfs_set_error("no HIP device"); return -1; }
    if (vfs_ib_scratch_setup(c->device)) { delete c; return -1; }
    c->stream = 0;
    const size_t P = (size_t)d.P;
    // metrics: 13 vector arrays + 4 scalar arrays
    size_t ng = (13 * 3 + 4) * P;
    if (hipMalloc(&c->geom_block, ng * sizeof(double)) != hipSuccess) { delete c; vfs_set_error("hipMalloc geometry"); return -1; }
    hipMemset(c->geom_block, 0, ng * sizeof(double));
    double *p = c->geom_block;
    double **vecs[13] = {&c->g.csi, &c->g.eta, &c->g.zet, &c->g.icsi, &c->g.ieta, &c->g.izet,
                         &c->g.jcsi, &c->g.jeta, &c->g.jzet, &c->g.kcsi, &c->g.keta, &c->g.kzet, &c->g.cent};
    for (int n = 0; n < 13; n++) { *vecs[n] = p; p += 3 * P; }
    double **scas[4] = {&c->g.aj, &c->g.iaj, &c->g.jaj, &c->g.kaj};
    for (int n = 0; n < 4; n++) { *scas[n] = p; p += P; }
    bool ok = hipMalloc(&c->nvert, P * sizeof(double)) == hipSuccess &&
              hipMalloc(&c->face, VFS_FACE_SLOTS(d) * P * sizeof(double)) == hipSuccess &&
              hipMalloc(&c->fp, 3 * P * sizeof(double)) == hipSuccess &&
              hipMalloc(&c->snap, 3 * P * sizeof(double)) == hipSuccess &&
              hipMalloc(&c->tmp3, 3 * P * sizeof(double)) == hipSuccess &&
              hipMalloc(&c->red, (VFS_RED_SLOTS * VFS_RED_BLOCKS + 256) * sizeof(double)) == hipSuccess &&
              hipMalloc(&c->kflux, 4 * sizeof(double)) == hipSuccess;
    if (!ok) { vfs_ctx_destroy(c); vfs_set_error("hipMalloc scratch"); return -1; }
    hipMemset(c->nvert, 0, P * sizeof(double));
    hipMemset(c->face, 0, VFS_FACE_SLOTS(d) * P * sizeof(double));
    hipMemset(c->fp, 0, 3 * P * sizeof(double));
    hipMemset(c->snap, 0, 3 * P * sizeof(double));
    hipMemset(c->tmp3, 0, 3 * P * sizeof(double));
    hipMemset(c->kflux, 0, 4 * sizeof(double));
    c->inlet_flux = -1.0;
    c->red_blocks = VFS_RED_BLOCKS;
    c->nvert_f32 = 1;
    if (hipDeviceSynchronize() != hipSuccess) { vfs_ctx_destroy(c); vfs_set_error("device sync"); return -1; }
    if (vfs_wall_setup(c) || vfs_wall_mask(c) || hipDeviceSynchronize() != hipSuccess) {
        vfs_ctx_destroy(c);
        return -1;
    }
    *out = c;
    return 0;
}

extern "C" int vfs_ctx_destroy(vfs_ctx *c)
{
    if (!c) return 0;
    hipDeviceSynchronize();
    vfs_poisson_destroy(c);
    vfs_implicit_destroy(c);
    hipFree(c->geom_block); hipFree(c->nvert); hipFree(c->face); hipFree(c->fp);
    hipFree(c->snap); hipFree(c->tmp3); hipFree(c->red); hipFree(c->kflux);
    if (c->ad) hipFree(c->ad);
    if (c->vjp_ws) hipFree(c->vjp_ws);
    if (c->wf_table) hipFree(c->wf_table);
    if (c->ustar) hipFree(c->ustar);
    if (c->dsm) hipFree(c->dsm);
    if (c->ff2_flags) hipFree(c->ff2_flags);
    delete c;
    return 0;
}

extern "C" int vfs_set_inlet_flux(vfs_ctx *c, double inlet_flux)
{
    c->inlet_flux = inlet_flux;
    return 0;
}

extern "C" int vfs_set_stream(vfs_ctx *c, void *s)
{
    c->stream = (hipStream_t)s;
    return 0;
}

extern "C" int vfs_synchronize(vfs_ctx *c)
{
    VFS_CHECK_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int vfs_device_malloc(size_t bytes, void **ptr)
{
    VFS_CHECK_HIP(hipMalloc(ptr, bytes ? bytes : 8));
    VFS_CHECK_HIP(hipMemset(*ptr, 0, bytes ? bytes : 8));
    return 0;
}

extern "C" int vfs_device_free(void *ptr)
{
    VFS_CHECK_HIP(hipFree(ptr));
    return 0;
}

extern "C" int vfs_memcpy_htod(void *dst, const void *src, size_t bytes)
{
    VFS_CHECK_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return 0;
}

extern "C" int vfs_memcpy_dtoh(void *dst, const void *src, size_t bytes)
{
    VFS_CHECK_HIP(hipDeviceSynchronize());
    VFS_CHECK_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int vfs_memcpy_dtod(void *dst, const void *src, size_t bytes)
{
    VFS_CHECK_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return 0;
}

// ---------------------------------------------------------------------------
// restart files (host I/O in vfs_io.c)
// ---------------------------------------------------------------------------
extern "C" void vfs_io_set_error(const char *msg) { vfs_set_error(msg); }
extern "C" int vfs_vec_write_petsc(const char *path, const double *v, long long n);
extern "C" int vfs_vec_read_petsc(const char *path, double *v, long long n);
extern "C" int vfs_restart_path(char *out, size_t cap, const char *dir, const char *stem, int ti, int block);

// Ucont_P_Binary_Output (main.c:799-847) without the optional averaging and
// Q-criterion fields: vfield (ucont), ufield (ucat), pfield (p) and nvfield
// (nvert) of one block at step ti, from device arrays.
extern "C" int vfs_restart_write(vfs_ctx *c, const char *dir, int ti, int block, const double *ucont,
                                 const double *ucat, const double *p, const double *nvert)
{
    VFS_REQUIRE(c && dir && ucont && ucat && p && nvert, "null argument");
    const long long P = c->d.P;
    std::vector<double> h((size_t)(3 * P));
    char path[1024];
    const struct { const char *stem; const double *src; int dof; } f[4] = {
        {"vfield", ucont, 3}, {"ufield", ucat, 3}, {"pfield", p, 1}, {"nvfield", nvert, 1}};
    VFS_CHECK_HIP(hipStreamSynchronize(c->stream));
    for (int q = 0; q < 4; q++) {
        const long long n = f[q].dof * P;
        VFS_CHECK_HIP(hipMemcpy(h.data(), f[q].src, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
        if (vfs_restart_path(path, sizeof path, dir, f[q].stem, ti, block)) return -2;
        if (vfs_vec_write_petsc(path, h.data(), n)) return -3;
    }
    return 0;
}

// Ucont_P_Binary_Input (main.c:394-420): vfield -> ucont, pfield -> p,
// nvfield -> nvert_o (the reference loads it into Nvert_o), ufield -> ucat,
// into device arrays.
extern "C" int vfs_restart_read(vfs_ctx *c, const char *dir, int ti, int block, double *ucont, double *ucat,
                                double *p, double *nvert_o)
{
    VFS_REQUIRE(c && dir && ucont && ucat && p && nvert_o, "null argument");
    const long long P = c->d.P;
    std::vector<double> h((size_t)(3 * P));
    char path[1024];
    const struct { const char *stem; double *dst; int dof; } f[4] = {
        {"vfield", ucont, 3}, {"pfield", p, 1}, {"nvfield", nvert_o, 1}, {"ufield", ucat, 3}};
    VFS_CHECK_HIP(hipStreamSynchronize(c->stream));
    for (int q = 0; q < 4; q++) {
        const long long n = f[q].dof * P;
        if (vfs_restart_path(path, sizeof path, dir, f[q].stem, ti, block)) return -2;
        if (vfs_vec_read_petsc(path, h.data(), n)) return -3;
        VFS_CHECK_HIP(hipMemcpy(f[q].dst, h.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    }
    return 0;
}
