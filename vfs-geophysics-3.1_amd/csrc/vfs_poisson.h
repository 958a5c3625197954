// Pressure-solver state owned by a context.
#pragma once
#include <vector>

#include "vfs_internal.h"

// One level of the aggregation multigrid.  Vectors use a compact layout over
// the level's cells, index (k*ny + j)*nx + i, i fastest.  Level 0 is the
// interior of the DA grid with the reference's 19-point operator; coarser
// levels carry 27-point Galerkin operators of 2x2x2 aggregates.
struct MGLevel {
    int nx, ny, nz;
    long long n;
    int per_i, per_j, per_k;
    int nslot;                         // 19 (level 0) or 27
    int sk;                            // k cells of the finer level per cell of this one (2, or 1 if k is kept)
    double *A64;  float *A32;          // slot-major coefficients, nslot*n
    // j-line Thomas factors: lower coupling, modified upper coupling, inverse pivot
    double *tl64, *tu64, *ti64;
    float *tl32, *tu32, *ti32;
    // work vectors (both precisions)
    double *b64, *x64, *y64, *d64, *e64;
    float *b32, *x32, *y32, *d32, *e32;
};

struct Poisson {
    int ready;
    int n;                              // unknowns
    double *gid;                        // DA layout
    int *blocksum, *total;
    double *coef;  float *coef32;       // DA layout, 19*P
    unsigned char *rowflag;             // DA layout
    std::vector<MGLevel> lv;
    int nc;                             // unknowns of the dense coarsest level
    double *cinv64; float *cinv32;      // (A_c + 1 1^T / nc)^{-1}, nc*nc, column-major
    double *cinvr64; float *cinvr32;    // the same scattered to all n cells of the level, row-major (n*n)
    long long *cidx;                    // coarsest compact index of each dense unknown
    // Krylov space (level-0 compact vectors)
    int m;
    double *V64; float *V32;            // (m+1) vectors
    double *w64, *z64, *r64, *x64, *bb64, *e64;
    float *w32, *z32, *x32, *bb32;
    double *dots;                       // device scalars for multi-dot results
    double *hdots;                      // pinned host mirror
    double omega;                       // line-smoother damping
    double alpha, alpha_semi;           // scale of the coarse correction from 2x2x2 / 2x2x1 (k kept) aggregates
    int wdepth;                         // levels l < wdepth visit level l+1 twice (W-cycle)
    int lsmall;                         // first level cycled by the one-workgroup kernel
};
