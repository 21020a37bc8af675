/*
 * CPU oracle body, included twice by baorec_oracle.c with REAL = float and
 * REAL = double (suffix SFX = _f32 / _f64).  TEST INFRASTRUCTURE ONLY: this is
 * the checker the HIP path is compared against, never part of the product.
 *
 * Every function restates one reference routine (BAOrec.jl, Julia) loop for
 * loop, translated to 0-based indexing on x-fastest grids
 * (linear index i + nx*(j + ny*k), Julia's column-major layout).
 * The float instantiation keeps the reference's FP32 operation order
 * (no contraction: build with -ffp-contract=off); the double instantiation is
 * the same algorithm in FP64, used to state errors against FP64.
 */

/* ORACLE_PLANT selects a deliberately wrong variant, built only as the
 * separate test libraries _build/liboracle_plant<k>.so, so that the tests can
 * show that their checks catch the mistake (0, the default, is the oracle):
 *   1  the sign of the operator's first-derivative term sG is flipped;
 *   2  the cell-centre position p = x/cell drops box_min;
 *   3  prolongation feeds fine cell 2c (not 2c+1) from coarse cell c;
 *   4  fmg starts each level from v = 0 instead of the prolonged coarse solve. */
#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif
#if ORACLE_PLANT == 1
#define PLANT_GSIGN -
#else
#define PLANT_GSIGN
#endif
#if ORACLE_PLANT == 2
#define PLANT_XC(a, i) (L->xv[a][i] - (REAL)L->bmin[a])
#else
#define PLANT_XC(a, i) (L->xv[a][i])
#endif

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

/* Cell centres, src/utils.jl:21-25.  The reference builds the range in
 * Float64 from a Float32 cell size fl32(L/n) and rounds each entry to Float32;
 * the double instantiation uses the exact cell size L/n. */
static void FN(xvec_axis)(REAL *out, int n, float L, float b) {
    if (sizeof(REAL) == sizeof(float)) {
        float h = L / (float)n;
        for (int i = 0; i < n; i++)
            out[i] = (REAL)(float)((double)b + ((double)i + 0.5) * (double)h);
    } else {
        double h = (double)L / (double)n;
        for (int i = 0; i < n; i++)
            out[i] = (REAL)((double)b + ((double)i + 0.5) * h);
    }
}

/* CIC scatter, src/mas.jl:1-52 (CPU, serial) and :53-99 (GPU, same
 * arithmetic).  wrap=1 mutates x, y, z exactly like the reference (one-sided,
 * using box_min[1]/box_size[1] for all three axes, src/mas.jl:7-11).
 * Returns 0, or -5 when a corner index falls outside 1..n (the reference throws
 * a BoundsError there). */
int FN(oracle_cic)(REAL *rho, int nx, int ny, int nz, float *x, float *y,
                   float *z, const float *w, int64_t n, const float *box_size,
                   const float *box_min, int wrap) {
    const int nb[3] = {nx, ny, nz};
    for (int64_t i = 0; i < n; i++) {
        if (wrap) {
            x[i] = (x[i] - box_min[0]) > box_size[0] ? x[i] - box_size[0] : x[i];
            y[i] = (y[i] - box_min[0]) > box_size[0] ? y[i] - box_size[0] : y[i];
            z[i] = (z[i] - box_min[0]) > box_size[0] ? z[i] - box_size[0] : z[i];
        }
        float p[3] = {x[i], y[i], z[i]};
        int64_t i0[3], i1[3];
        REAL w0[3], w1[3];
        for (int a = 0; a < 3; a++) {
            REAL u;
            if (sizeof(REAL) == sizeof(float))
                u = (REAL)((p[a] - box_min[a]) * (float)nb[a] / box_size[a] + 1.0f);
            else
                u = ((double)p[a] - (double)box_min[a]) * (double)nb[a] / (double)box_size[a] + 1.0;
            if (!(u == u)) return -5; /* NaN: Int(floor(NaN)) throws */
            if (u < -1e18 || u > 1e18) return -5;
            int64_t f = (int64_t)floor((double)u);
            w1[a] = u - (REAL)f;
            w0[a] = (REAL)1 - w1[a];
            if (f == nb[a] + 1) f = 1;
            int64_t g = (f == nb[a] && wrap) ? 1 : f + 1;
            if (f < 1 || f > nb[a] || g < 1 || g > nb[a]) return -5;
            i0[a] = f - 1;
            i1[a] = g - 1;
        }
        REAL wt = (REAL)w[i];
        w0[0] *= wt;
        w1[0] *= wt;
#define IDX(a, b, c) ((a) + (int64_t)nx * ((b) + (int64_t)ny * (c)))
        rho[IDX(i0[0], i0[1], i0[2])] += w0[0] * w0[1] * w0[2];
        rho[IDX(i1[0], i0[1], i0[2])] += w1[0] * w0[1] * w0[2];
        rho[IDX(i0[0], i1[1], i0[2])] += w0[0] * w1[1] * w0[2];
        rho[IDX(i0[0], i0[1], i1[2])] += w0[0] * w0[1] * w1[2];
        rho[IDX(i1[0], i1[1], i0[2])] += w1[0] * w1[1] * w0[2];
        rho[IDX(i1[0], i0[1], i1[2])] += w1[0] * w0[1] * w1[2];
        rho[IDX(i0[0], i1[1], i1[2])] += w0[0] * w1[1] * w1[2];
        rho[IDX(i1[0], i1[1], i1[2])] += w1[0] * w1[1] * w1[2];
    }
    return 0;
}

/* Trilinear read-back, src/mas.jl:218-269 (CPU: u = (p-b)/cell) and
 * :271-318 (GPU: u = (p-b)*n/L); gpu_formula selects which.  wrap=true as
 * called from compute_displacements. Returns -5 on an out-of-range index. */
int FN(oracle_read_cic)(REAL *out, const REAL *field, int nx, int ny, int nz,
                        const float *x, const float *y, const float *z,
                        int64_t n, const float *box_size, const float *box_min,
                        int gpu_formula) {
    const int nb[3] = {nx, ny, nz};
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t i = 0; i < n; i++) {
        float p[3] = {x[i], y[i], z[i]};
        int64_t d[3], up[3];
        REAL uu[3], dd[3];
        for (int a = 0; a < 3; a++) {
            REAL dist;
            if (sizeof(REAL) == sizeof(float)) {
                if (gpu_formula)
                    dist = (REAL)((p[a] - box_min[a]) * (float)nb[a] / box_size[a]);
                else
                    dist = (REAL)((p[a] - box_min[a]) / (box_size[a] / (float)nb[a]));
            } else {
                dist = ((double)p[a] - (double)box_min[a]) * (double)nb[a] / (double)box_size[a];
            }
            if (!(dist == dist) || dist < -1e18 || dist > 1e18) { bad = 1; dist = 0; }
            int64_t f = (int64_t)floor((double)dist);
            uu[a] = dist - (REAL)f;
            dd[a] = (REAL)1 - uu[a];
            int64_t id = f + 1;
            if (id > nb[a]) id -= nb[a];
            int64_t iu = id + 1;
            if (iu > nb[a]) iu -= nb[a];
            if (id < 1 || id > nb[a] || iu < 1 || iu > nb[a]) { bad = 1; id = iu = 1; }
            d[a] = id - 1;
            up[a] = iu - 1;
        }
        out[i] = field[IDX(d[0], d[1], d[2])] * dd[0] * dd[1] * dd[2] +
                 field[IDX(d[0], d[1], up[2])] * dd[0] * dd[1] * uu[2] +
                 field[IDX(d[0], up[1], d[2])] * dd[0] * uu[1] * dd[2] +
                 field[IDX(d[0], up[1], up[2])] * dd[0] * uu[1] * uu[2] +
                 field[IDX(up[0], d[1], d[2])] * uu[0] * dd[1] * dd[2] +
                 field[IDX(up[0], d[1], up[2])] * uu[0] * dd[1] * uu[2] +
                 field[IDX(up[0], up[1], d[2])] * uu[0] * uu[1] * dd[2] +
                 field[IDX(up[0], up[1], up[2])] * uu[0] * uu[1] * uu[2];
    }
    return bad ? -5 : 0;
}

/* ---------------------------------------------------------------- multigrid */

/* Per-level constants of src/multigrid.jl:16-29 (cell, cell^2, 1/cell^2,
 * los/cell) and the level's cell centres. */
typedef struct {
    int n[3];
    REAL cell[3], cell2[3], icell2[3], losn[3];
    REAL *xv[3];
    float bmin[3];
} FN(mglevel);

static void FN(mg_setup)(FN(mglevel) *L, int nx, int ny, int nz,
                         const float *box_size, const float *box_min,
                         int has_los, const float *los) {
    L->n[0] = nx; L->n[1] = ny; L->n[2] = nz;
    for (int a = 0; a < 3; a++) L->bmin[a] = box_min[a];
    for (int a = 0; a < 3; a++) {
        if (sizeof(REAL) == sizeof(float)) {
            float c = box_size[a] / (float)L->n[a];
            float c2 = c * c;
            L->cell[a] = c;
            L->cell2[a] = c2;
            L->icell2[a] = 1.0f / c2;
            L->losn[a] = has_los ? los[a] / c : 0;
        } else {
            double c = (double)box_size[a] / (double)L->n[a];
            L->cell[a] = c;
            L->cell2[a] = c * c;
            L->icell2[a] = 1.0 / (c * c);
            L->losn[a] = has_los ? (double)los[a] / c : 0;
        }
        L->xv[a] = (REAL *)malloc(sizeof(REAL) * L->n[a]);
        FN(xvec_axis)(L->xv[a], L->n[a], box_size[a], box_min[a]);
    }
}

static void FN(mg_free)(FN(mglevel) *L) {
    for (int a = 0; a < 3; a++) free(L->xv[a]);
}

#define WRAPP(i, n) ((i) + 1 >= (n) ? (i) + 1 - (n) : (i) + 1)
#define WRAPM(i, n) ((i) == 0 ? (n) - 1 : (i) - 1)

/* Anisotropic operator pieces shared by jacobi! and residual!
 * (src/multigrid.jl:37-82 and :235-276). */
#define MG_STENCIL_BODY                                                         \
    int64_t nx = L->n[0], ny = L->n[1], nz = L->n[2];                           \
    int64_t ixp = WRAPP(ix, nx), ixm = WRAPM(ix, nx);                           \
    int64_t iyp = WRAPP(iy, ny), iym = WRAPM(iy, ny);                           \
    int64_t izp = WRAPP(iz, nz), izm = WRAPM(iz, nz);                           \
    REAL px = has_los ? L->losn[0] : PLANT_XC(0, ix) / L->cell[0];              \
    REAL py = has_los ? L->losn[1] : PLANT_XC(1, iy) / L->cell[1];              \
    REAL pz = has_los ? L->losn[2] : PLANT_XC(2, iz) / L->cell[2];              \
    REAL g = beta / (L->cell2[0] * (px * px) + L->cell2[1] * (py * py) +         \
                     L->cell2[2] * (pz * pz));                                  \
    REAL gpx2 = L->icell2[0] + g * (px * px);                                   \
    REAL gpy2 = L->icell2[1] + g * (py * py);                                   \
    REAL gpz2 = L->icell2[2] + g * (pz * pz);                                   \
    REAL sA = gpx2 * (v[IDX(ixp, iy, iz)] + v[IDX(ixm, iy, iz)]);              \
    REAL sB = gpy2 * (v[IDX(ix, iyp, iz)] + v[IDX(ix, iym, iz)]);              \
    REAL sC = gpz2 * (v[IDX(ix, iy, izp)] + v[IDX(ix, iy, izm)]);              \
    REAL sX = (g / (REAL)2) *                                                   \
              ((px * py) * (v[IDX(ixp, iyp, iz)] + v[IDX(ixm, iym, iz)] -      \
                            v[IDX(ixm, iyp, iz)] - v[IDX(ixp, iym, iz)]) +     \
               (px * pz) * (v[IDX(ixp, iy, izp)] + v[IDX(ixm, iy, izm)] -      \
                            v[IDX(ixm, iy, izp)] - v[IDX(ixp, iy, izm)]) +     \
               (py * pz) * (v[IDX(ix, iyp, izp)] + v[IDX(ix, iym, izm)] -      \
                            v[IDX(ix, iym, izp)] - v[IDX(ix, iyp, izm)]));     \
    REAL sG = PLANT_GSIGN g * (px * (v[IDX(ixp, iy, iz)] - v[IDX(ixm, iy, iz)]) +           \
                   py * (v[IDX(ix, iyp, iz)] - v[IDX(ix, iym, iz)]) +           \
                   pz * (v[IDX(ix, iy, izp)] - v[IDX(ix, iy, izm)]));

/* One damped-Jacobi sweep set, src/multigrid.jl:14-135. */
static void FN(mg_jacobi)(REAL *v, const REAL *f, const FN(mglevel) *L, REAL beta,
                          REAL omega, int niter, int has_los) {
    int64_t nx0 = L->n[0], ny0 = L->n[1], nz0 = L->n[2];
    int64_t N = nx0 * ny0 * nz0;
    REAL *jac = (REAL *)malloc(sizeof(REAL) * N);
    for (int it = 0; it < niter; it++) {
#pragma omp parallel for collapse(2) schedule(static)
        for (int64_t iz = 0; iz < nz0; iz++)
            for (int64_t iy = 0; iy < ny0; iy++)
                for (int64_t ix = 0; ix < nx0; ix++) {
                    MG_STENCIL_BODY
                    int64_t I = IDX(ix, iy, iz);
                    REAL j = f[I] + sA + sB + sC + sX;
                    if (!has_los) j += sG;
                    j /= (REAL)2 * (gpx2 + gpy2 + gpz2);
                    jac[I] = j;
                }
        REAL a = (REAL)1 - omega;
#pragma omp parallel for schedule(static)
        for (int64_t I = 0; I < N; I++) v[I] = a * v[I] + omega * jac[I];
    }
    free(jac);
}

/* r = f - A v, src/multigrid.jl:213-326. */
static void FN(mg_residual)(REAL *r, const REAL *v, const REAL *f,
                            const FN(mglevel) *L, REAL beta, int has_los) {
    int64_t nx0 = L->n[0], ny0 = L->n[1], nz0 = L->n[2];
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t iz = 0; iz < nz0; iz++)
        for (int64_t iy = 0; iy < ny0; iy++)
            for (int64_t ix = 0; ix < nx0; ix++) {
                MG_STENCIL_BODY
                int64_t I = IDX(ix, iy, iz);
                REAL t = (REAL)2 * (gpx2 + gpy2 + gpz2) * v[I] - (sA + sB + sC + sX);
                if (!has_los) t -= sG;
                r[I] = f[I] - t;
            }
}

/* Trilinear prolongation, src/multigrid.jl:398-453: coarse cell c feeds fine
 * cells 2c+1 and 2c+2 (mod n_f) per axis (the 1-based 2c, 2c+1). Overwrites. */
static void FN(mg_prolong)(REAL *v1h, const REAL *v2h, int cx, int cy, int cz) {
    int64_t fx = 2 * (int64_t)cx, fy = 2 * (int64_t)cy;
    int64_t fz = 2 * (int64_t)cz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t iz = 0; iz < cz; iz++)
        for (int64_t iy = 0; iy < cy; iy++)
            for (int64_t ix = 0; ix < cx; ix++) {
                int64_t ixp = WRAPP(ix, cx), iyp = WRAPP(iy, cy), izp = WRAPP(iz, cz);
#define CI(a, b, c) ((a) + (int64_t)cx * ((b) + (int64_t)cy * (c)))
#define FI(a, b, c) ((a) + fx * ((b) + fy * (c)))
                int64_t x0 = 2 * ix + 1, y0 = 2 * iy + 1, z0 = 2 * iz + 1;
#if ORACLE_PLANT == 3
                x0 = 2 * ix; y0 = 2 * iy; z0 = 2 * iz;
#endif
                int64_t xp = x0 + 1 >= fx ? x0 + 1 - fx : x0 + 1;
                int64_t yp = y0 + 1 >= fy ? y0 + 1 - fy : y0 + 1;
                int64_t zp = z0 + 1 >= fz ? z0 + 1 - fz : z0 + 1;
                REAL c000 = v2h[CI(ix, iy, iz)];
                REAL c100 = v2h[CI(ixp, iy, iz)], c010 = v2h[CI(ix, iyp, iz)];
                REAL c001 = v2h[CI(ix, iy, izp)], c110 = v2h[CI(ixp, iyp, iz)];
                REAL c011 = v2h[CI(ix, iyp, izp)], c101 = v2h[CI(ixp, iy, izp)];
                REAL c111 = v2h[CI(ixp, iyp, izp)];
                v1h[FI(x0, y0, z0)] = c000;
                v1h[FI(xp, y0, z0)] = (c000 + c100) / (REAL)2;
                v1h[FI(x0, yp, z0)] = (c000 + c010) / (REAL)2;
                v1h[FI(x0, y0, zp)] = (c000 + c001) / (REAL)2;
                v1h[FI(xp, yp, z0)] = (c000 + c100 + c010 + c110) / (REAL)4;
                v1h[FI(x0, yp, zp)] = (c000 + c010 + c001 + c011) / (REAL)4;
                v1h[FI(xp, y0, zp)] = (c000 + c100 + c001 + c101) / (REAL)4;
                v1h[FI(xp, yp, zp)] = (c000 + c100 + c010 + c001 + c110 + c101 +
                                       c011 + c111) / (REAL)8;
            }
}

/* 27-point full weighting, src/multigrid.jl:520-584, centred on fine cell
 * 2c+1 (the 1-based 2c). */
static void FN(mg_reduce)(REAL *v2h, const REAL *v1h, int fxn, int fyn, int fzn) {
    int64_t cx = fxn / 2, cy = fyn / 2, cz = fzn / 2;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t kz = 0; kz < cz; kz++)
        for (int64_t ky = 0; ky < cy; ky++)
            for (int64_t kx = 0; kx < cx; kx++) {
                int64_t ix = 2 * kx + 1, iy = 2 * ky + 1, iz = 2 * kz + 1;
                int64_t ixp = WRAPP(ix, fxn), ixm = WRAPM(ix, fxn);
                int64_t iyp = WRAPP(iy, fyn), iym = WRAPM(iy, fyn);
                int64_t izp = WRAPP(iz, fzn), izm = WRAPM(iz, fzn);
#define FJ(a, b, c) ((a) + (int64_t)fxn * ((b) + (int64_t)fyn * (c)))
                const REAL *w = v1h;
                REAL s = (REAL)8 * w[FJ(ix, iy, iz)] +
                         (REAL)4 * (w[FJ(ixp, iy, iz)] + w[FJ(ixm, iy, iz)] +
                                    w[FJ(ix, iyp, iz)] + w[FJ(ix, iym, iz)] +
                                    w[FJ(ix, iy, izp)] + w[FJ(ix, iy, izm)]) +
                         (REAL)2 * (w[FJ(ixp, iyp, iz)] + w[FJ(ixm, iyp, iz)] +
                                    w[FJ(ixp, iym, iz)] + w[FJ(ixm, iym, iz)] +
                                    w[FJ(ixp, iy, izp)] + w[FJ(ixm, iy, izp)] +
                                    w[FJ(ixp, iy, izm)] + w[FJ(ixm, iy, izm)] +
                                    w[FJ(ix, iyp, izp)] + w[FJ(ix, iym, izp)] +
                                    w[FJ(ix, iyp, izm)] + w[FJ(ix, iym, izm)]) +
                         w[FJ(ixp, iyp, izp)] + w[FJ(ixm, iyp, izp)] +
                         w[FJ(ixp, iym, izp)] + w[FJ(ixm, iym, izp)] +
                         w[FJ(ixp, iyp, izm)] + w[FJ(ixm, iyp, izm)] +
                         w[FJ(ixp, iym, izm)] + w[FJ(ixm, iym, izm)];
                v2h[kx + cx * (ky + cy * kz)] = (REAL)((double)s / 64.0);
            }
}

static int FN(mg_recurse)(const int *n) {
    int r = 1;
    for (int a = 0; a < 3; a++) r &= (n[a] > 4 && n[a] % 2 == 0);
    return r;
}

/* src/multigrid.jl:654-687. */
static void FN(mg_vcycle)(REAL *v, const REAL *f, int nx, int ny, int nz,
                          const float *box_size, const float *box_min, REAL beta,
                          REAL omega, int niter, int has_los, const float *los) {
    FN(mglevel) L;
    FN(mg_setup)(&L, nx, ny, nz, box_size, box_min, has_los, los);
    int64_t N = (int64_t)nx * ny * nz;
    FN(mg_jacobi)(v, f, &L, beta, omega, niter, has_los);
    if (FN(mg_recurse)(L.n)) {
        int cx = nx / 2, cy = ny / 2, cz = nz / 2;
        int64_t Nc = (int64_t)cx * cy * cz;
        REAL *r = (REAL *)malloc(sizeof(REAL) * N);
        FN(mg_residual)(r, v, f, &L, beta, has_los);
        REAL *f2h = (REAL *)calloc(Nc, sizeof(REAL));
        FN(mg_reduce)(f2h, r, nx, ny, nz);
        free(r);
        REAL *v2h = (REAL *)calloc(Nc, sizeof(REAL));
        FN(mg_vcycle)(v2h, f2h, cx, cy, cz, box_size, box_min, beta, omega, niter,
                      has_los, los);
        free(f2h);
        REAL *v1h = (REAL *)calloc(N, sizeof(REAL));
        FN(mg_prolong)(v1h, v2h, cx, cy, cz);
        free(v2h);
        for (int64_t I = 0; I < N; I++) v[I] += v1h[I];
        free(v1h);
    }
    FN(mg_jacobi)(v, f, &L, beta, omega, niter, has_los);
    FN(mg_free)(&L);
}

/* Full multigrid, src/multigrid.jl:722-752; v is overwritten. */
static void FN(mg_fmg)(REAL *v, const REAL *f, int nx, int ny, int nz,
                       const float *box_size, const float *box_min, REAL beta,
                       REAL omega, int niter, int ncycles, int has_los,
                       const float *los) {
    int n[3] = {nx, ny, nz};
    int64_t N = (int64_t)nx * ny * nz;
    if (FN(mg_recurse)(n)) {
        int cx = nx / 2, cy = ny / 2, cz = nz / 2;
        int64_t Nc = (int64_t)cx * cy * cz;
        REAL *f2h = (REAL *)calloc(Nc, sizeof(REAL));
        FN(mg_reduce)(f2h, f, nx, ny, nz);
        REAL *v2h = (REAL *)calloc(Nc, sizeof(REAL));
        FN(mg_fmg)(v2h, f2h, cx, cy, cz, box_size, box_min, beta, omega, niter,
                   ncycles, has_los, los);
        free(f2h);
        FN(mg_prolong)(v, v2h, cx, cy, cz);
        free(v2h);
#if ORACLE_PLANT == 4
        for (int64_t I = 0; I < N; I++) v[I] = 0;
#endif
    } else {
        for (int64_t I = 0; I < N; I++) v[I] = 0;
    }
    for (int c = 0; c < ncycles; c++)
        FN(mg_vcycle)(v, f, nx, ny, nz, box_size, box_min, beta, omega, niter,
                      has_los, los);
}

/* ------------------------------------------------ exported oracle entries */

void FN(oracle_xvec)(REAL *out, int n, float L, float b) { FN(xvec_axis)(out, n, L, b); }

void FN(oracle_mg_jacobi)(REAL *v, const REAL *f, int nx, int ny, int nz,
                          const float *box_size, const float *box_min, REAL beta,
                          REAL omega, int niter, int has_los, const float *los) {
    FN(mglevel) L;
    FN(mg_setup)(&L, nx, ny, nz, box_size, box_min, has_los, los);
    FN(mg_jacobi)(v, f, &L, beta, omega, niter, has_los);
    FN(mg_free)(&L);
}

void FN(oracle_mg_residual)(REAL *r, const REAL *v, const REAL *f, int nx, int ny,
                            int nz, const float *box_size, const float *box_min,
                            REAL beta, int has_los, const float *los) {
    FN(mglevel) L;
    FN(mg_setup)(&L, nx, ny, nz, box_size, box_min, has_los, los);
    FN(mg_residual)(r, v, f, &L, beta, has_los);
    FN(mg_free)(&L);
}

void FN(oracle_mg_prolong)(REAL *v1h, const REAL *v2h, int cx, int cy, int cz) {
    FN(mg_prolong)(v1h, v2h, cx, cy, cz);
}

void FN(oracle_mg_reduce)(REAL *v2h, const REAL *v1h, int fx, int fy, int fz) {
    FN(mg_reduce)(v2h, v1h, fx, fy, fz);
}

void FN(oracle_mg_vcycle)(REAL *v, const REAL *f, int nx, int ny, int nz,
                          const float *box_size, const float *box_min, REAL beta,
                          REAL omega, int niter, int has_los, const float *los) {
    FN(mg_vcycle)(v, f, nx, ny, nz, box_size, box_min, beta, omega, niter, has_los, los);
}

void FN(oracle_mg_fmg)(REAL *v, const REAL *f, int nx, int ny, int nz,
                       const float *box_size, const float *box_min, REAL beta,
                       REAL omega, int niter, int ncycles, int has_los,
                       const float *los) {
    FN(mg_fmg)(v, f, nx, ny, nz, box_size, box_min, beta, omega, niter, ncycles,
               has_los, los);
}

#undef IDX
#undef CI
#undef FI
#undef FJ
#undef MG_STENCIL_BODY
#undef WRAPP
#undef WRAPM
#undef PLANT_GSIGN
#undef PLANT_XC
