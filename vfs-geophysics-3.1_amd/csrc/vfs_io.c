/*
 * The reference's on-disk formats on either side of the hot path, host code:
 *  - the grid file reader of MG_Initial (init.c:316-436): text grid.dat,
 *    binary grid.dat (-binary 1) and the tensor-product xyz.dat (-xyz 1);
 *  - PETSc binary vector files, the restart set written by
 *    Ucont_P_Binary_Output and read by Ucont_P_Binary_Input
 *    (main.c:394-420, 799-847): vfield (Ucont), ufield (Ucat), pfield (P),
 *    nvfield (Nvert) "%06d_%1d.dat" files.
 *
 * A PETSc binary vector file is the class id 1211214 and the length as
 * big-endian int32, then the values as big-endian float64; a DA vector is
 * written in the DA's natural order, which is this library's padded
 * [k][j][i] (x dof) layout.  The .info side file the writer removes
 * (main.c:812) is not produced.
 */
#include <errno.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define VFS_PETSC_VEC_ID 1211214

void vfs_io_set_error(const char *msg);   /* vfs_core.hip */

static void fail(const char *what, const char *path)
{
    char b[512];
    snprintf(b, sizeof b, "%s: %s", what, path);
    vfs_io_set_error(b);
}

/* fscanf "%le" as the reference reads (init.c:365-367, 392-430) */
static int read_le(FILE *f, double *v) { return fscanf(f, "%le", v) == 1 ? 0 : -1; }

/* format 0 text grid.dat, 1 binary grid.dat, 2 xyz.dat.  Coordinates are
 * value / cl * L_dim (init.c:406-433), stored like the DA coordinate vector:
 * (IM+1)*(JM+1)*(KM+1) Cmpnts, node (k,j,i) at ((k*(JM+1) + j)*(IM+1) + i),
 * the points with i = IM, j = JM or k = KM left as they are.  coor == NULL:
 * dimensions only. */
int vfs_read_grid(const char *path, int format, double cl, double L_dim, int dims[3], double *coor)
{
    if (format < 0 || format > 2) { vfs_io_set_error("grid format must be 0 (text), 1 (binary) or 2 (xyz)"); return -2; }
    FILE *f = fopen(path, format == 1 ? "rb" : "r");
    if (!f) { fail("cannot open the grid file", path); return -3; }
    int nb = 1, IM = 0, JM = 0, KM = 0, rc = 0;
    if (format == 0) rc = fscanf(f, "%i\n", &nb) == 1 ? 0 : -1;
    else if (format == 1) rc = fread(&nb, sizeof(int), 1, f) == 1 ? 0 : -1;
    if (rc || nb != 1) {
        fclose(f);
        if (rc) fail("cannot read the block count of", path);
        else vfs_io_set_error("multi-block grid files are not supported (block count != 1)");
        return -4;
    }
    if (format == 1) {
        rc = (fread(&IM, sizeof(int), 1, f) == 1 && fread(&JM, sizeof(int), 1, f) == 1 &&
              fread(&KM, sizeof(int), 1, f) == 1) ? 0 : -1;
    } else {
        rc = fscanf(f, "%i %i %i\n", &IM, &JM, &KM) == 3 ? 0 : -1;
    }
    if (rc || IM < 2 || JM < 2 || KM < 2) { fclose(f); fail("bad grid dimensions in", path); return -5; }
    dims[0] = IM; dims[1] = JM; dims[2] = KM;
    if (!coor) { fclose(f); return 0; }
    const size_t n = (size_t)IM * JM * KM;
#define NODE(k, j, i) (coor + 3 * (((size_t)(k) * (JM + 1) + (j)) * (IM + 1) + (i)))
    if (format == 2) {
        double *X = malloc(sizeof(double) * (IM + JM + KM)), *Y = X + IM, *Z = Y + JM, t;
        if (!X) { fclose(f); vfs_io_set_error("out of host memory"); return -6; }
        for (int i = 0; i < IM && !rc; i++) rc = fscanf(f, "%le %le %le\n", &X[i], &t, &t) == 3 ? 0 : -1;
        for (int j = 0; j < JM && !rc; j++) rc = fscanf(f, "%le %le %le\n", &t, &Y[j], &t) == 3 ? 0 : -1;
        for (int k = 0; k < KM && !rc; k++) rc = fscanf(f, "%le %le %le\n", &t, &t, &Z[k]) == 3 ? 0 : -1;
        if (!rc)
            for (int k = 0; k < KM; k++)
                for (int j = 0; j < JM; j++)
                    for (int i = 0; i < IM; i++) {
                        double *c = NODE(k, j, i);
                        c[0] = X[i] / cl * L_dim;
                        c[1] = Y[j] / cl * L_dim;
                        c[2] = Z[k] / cl * L_dim;
                    }
        free(X);
    } else {
        /* every x, then every y, then every z, each over k, j, i (i fastest) */
        for (int q = 0; q < 3 && !rc; q++)
            for (size_t p = 0; p < n && !rc; p++) {
                double v;
                if (format == 1) rc = fread(&v, sizeof(double), 1, f) == 1 ? 0 : -1;
                else rc = read_le(f, &v);
                const int i = (int)(p % IM), j = (int)((p / IM) % JM), k = (int)(p / ((size_t)IM * JM));
                NODE(k, j, i)[q] = v / cl * L_dim;
            }
    }
#undef NODE
    fclose(f);
    if (rc) { fail("truncated grid file", path); return -7; }
    return 0;
}

static uint32_t bswap32(uint32_t v)
{
    return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24);
}
static uint64_t bswap64(uint64_t v) { return ((uint64_t)bswap32((uint32_t)v) << 32) | bswap32((uint32_t)(v >> 32)); }
static int little_endian(void)
{
    const uint16_t one = 1;
    return *(const uint8_t *)&one == 1;
}

/* n values from host memory to a PETSc binary vector file */
int vfs_vec_write_petsc(const char *path, const double *v, long long n)
{
    if (n < 0 || n > 0x7fffffffLL) { vfs_io_set_error("vector length outside PETSc's 32-bit range"); return -2; }
    FILE *f = fopen(path, "wb");
    if (!f) { fail("cannot create", path); return -3; }
    const int le = little_endian();
    uint32_t h[2] = {(uint32_t)VFS_PETSC_VEC_ID, (uint32_t)n};
    if (le) { h[0] = bswap32(h[0]); h[1] = bswap32(h[1]); }
    int rc = fwrite(h, sizeof(uint32_t), 2, f) == 2 ? 0 : -1;
    uint64_t buf[4096];
    for (long long p = 0; p < n && !rc; p += 4096) {
        const size_t m = (size_t)(n - p < 4096 ? n - p : 4096);
        memcpy(buf, v + p, m * sizeof(double));
        if (le)
            for (size_t q = 0; q < m; q++) buf[q] = bswap64(buf[q]);
        rc = fwrite(buf, sizeof(uint64_t), m, f) == m ? 0 : -1;
    }
    if (fclose(f) != 0) rc = -1;
    if (rc) { fail("write failed", path); return -4; }
    return 0;
}

/* a PETSc binary vector file of exactly n values into host memory */
int vfs_vec_read_petsc(const char *path, double *v, long long n)
{
    FILE *f = fopen(path, "rb");
    if (!f) { fail("cannot open", path); return -3; }
    const int le = little_endian();
    uint32_t h[2];
    int rc = fread(h, sizeof(uint32_t), 2, f) == 2 ? 0 : -1;
    if (!rc && le) { h[0] = bswap32(h[0]); h[1] = bswap32(h[1]); }
    if (rc || h[0] != (uint32_t)VFS_PETSC_VEC_ID) { fclose(f); fail("not a PETSc binary vector file", path); return -4; }
    if ((long long)h[1] != n) { fclose(f); fail("vector length in the file differs from the field size", path); return -5; }
    uint64_t buf[4096];
    for (long long p = 0; p < n && !rc; p += 4096) {
        const size_t m = (size_t)(n - p < 4096 ? n - p : 4096);
        rc = fread(buf, sizeof(uint64_t), m, f) == m ? 0 : -1;
        if (le)
            for (size_t q = 0; q < m; q++) buf[q] = bswap64(buf[q]);
        memcpy(v + p, buf, m * sizeof(double));
    }
    fclose(f);
    if (rc) { fail("truncated vector file", path); return -6; }
    return 0;
}

/* "<dir>/<stem>%06d_%1d.dat" (main.c:404, 804) */
int vfs_restart_path(char *out, size_t cap, const char *dir, const char *stem, int ti, int block)
{
    const int w = snprintf(out, cap, "%s/%s%06d_%1d.dat", dir, stem, ti, block);
    if (w < 0 || (size_t)w >= cap) { vfs_io_set_error("restart file path too long"); return -2; }
    return 0;
}
