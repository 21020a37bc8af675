/*
 * baorec.h — C ABI of the MI355X (gfx950) reconstruction library
 * libbaorec_hip.so.
 *
 * It replaces the reconstruction entry points of BAOrec.jl (a Julia package
 * whose GPU path is CUDA.jl + KernelAbstractions + CUFFT).  Each declaration
 * names the Julia method it stands in for (file:line in BAOrec.jl).  A Julia
 * host calls these through `ccall`; the Python package baorec_hip mirrors that
 * binding through ctypes.
 *
 * Conventions
 *  - Grids are FP32, x fastest: element (i, j, k) of an (nx, ny, nz) grid is at
 *    i + nx*(j + ny*k), which is Julia's column-major Array{Float32,3}.  Each
 *    side is a power of two from 16 to 1024 (other sizes return
 *    BAOREC_ERR_UNSUPPORTED); the reference's drivers and tests use 64 and 512.
 *  - Particles are FP32 structure-of-arrays (x, y, z, w).  One catalog of a
 *    mass assignment holds at most 536,870,911 particles with CIC and
 *    159,072,862 with TSC (32-bit placement indices); larger catalogs return
 *    BAOREC_ERR_UNSUPPORTED.
 *  - `mem` says where every pointer argument of the call lives:
 *    BAOREC_MEM_HOST (Julia Array) or BAOREC_MEM_DEVICE (a device pointer on
 *    the context's GPU, e.g. a ROCArray).  Host inputs are copied in, host
 *    outputs copied out; device pointers are used in place.
 *  - Every call is synchronous on return and returns 0 or a negative
 *    baorec_status; it never aborts.  baorec_last_error() gives the message.
 *  - One context per host thread; calls on one context are serialised.
 */
#ifndef BAOREC_H
#define BAOREC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct baorec_ctx baorec_ctx;

typedef enum {
    BAOREC_OK = 0,
    BAOREC_ERR_INVALID_ARGUMENT = -1, /* null pointer, bad size, bad enum */
    BAOREC_ERR_NO_DEVICE = -2,        /* no usable gfx950 GPU */
    BAOREC_ERR_HIP = -3,              /* a HIP runtime call failed */
    BAOREC_ERR_OUT_OF_MEMORY = -4,
    BAOREC_ERR_OUT_OF_BOX = -5,       /* a particle indexes outside the grid:
                                         the reference throws BoundsError */
    BAOREC_ERR_NO_RESULT = -6,        /* no cached result for this call */
    BAOREC_ERR_UNSUPPORTED = -7       /* e.g. a grid side that is not 2^k */
} baorec_status;

typedef enum { BAOREC_MEM_HOST = 0, BAOREC_MEM_DEVICE = 1 } baorec_mem;

typedef enum { BAOREC_ITERATIVE = 0, BAOREC_MULTIGRID = 1 } baorec_recon_kind;

/* read_shifts(...; field = :disp | :rsd | :sum), src/recon.jl:264-306 */
typedef enum {
    BAOREC_FIELD_DISP = 0,
    BAOREC_FIELD_RSD = 1,
    BAOREC_FIELD_SUM = 2
} baorec_field;

/* Mass-assignment scheme.  CIC is the reference's (src/mas.jl:1-99); TSC is
 * an extension with no reference counterpart. */
typedef enum { BAOREC_MAS_CIC = 0, BAOREC_MAS_TSC = 1 } baorec_mas;

/* Multigrid schedule.  REFERENCE is fmg with damped Jacobi, a fixed number of
 * V-cycles per level (src/multigrid.jl:654-752).  RBGS is an extension:
 * Gauss-Seidel V-cycles (rb_sweeps pre- and post-smoothing sweeps) in the
 * four-colour ordering that generalises red-black to the operator's 19-point
 * stencil (its cross-derivative terms couple diagonal neighbours), with levels
 * of side <= fp64_coarse_max in FP64.  The V-cycles refine
 * an FP64 solution: its fine residual is computed in FP64 and each V-cycle
 * solves for the correction (FP32 fine levels), so the residual can fall
 * below the FP32 rounding of A v (about 1e-4 of ||f|| at 512^3).  It stops
 * when the relative residual
 * (FP64 norm) is <= `tolerance`, when one V-cycle changes the residual by
 * <= tolerance * ||f||, or when two V-cycles in a row reduce it by less than
 * 10 % (a safeguard).  The solution is returned with zero mean (A
 * annihilates constants).
 * mg_warm_start = 1 starts RBGS from a spectral seed instead of v = 0: the
 * exact inverse of the discrete operator with the LOS fixed (the given LOS,
 * or for a local LOS the direction of the box centre seen from the origin,
 * without the operator's first-derivative term), one forward and inverse
 * FFT.  With a fixed LOS that is the solution itself up to FP32 rounding; the
 * V-cycles then continue to the same stopping rule.  mg_warm_start = 3 then
 * takes mg_seed_steps preconditioned Richardson steps v += A_c^-1 (f - A v)
 * (A_c^-1: that spectral inverse, A: the local-LOS operator).
 * mg_warm_start = 2 adds the learned correction of the context's network
 * (baorec_warmstart_*) to the spectral seed and then takes the same steps,
 * so 2 against 3 isolates what the network contributes.
 * With a local LOS the operator is not symmetric and only f - alpha, alpha
 * = (w . f) / (w . 1) for its left null vector w, is in its range: the solve
 * moves the residual's mean into alpha after every V-cycle and reports it in
 * rhs_shift (with a fixed LOS, w = 1 and alpha is the mean of f). */
typedef enum { BAOREC_MG_REFERENCE = 0, BAOREC_MG_RBGS = 1 } baorec_mg_schedule;

/*
 * Fields of IterativeRecon (src/recon.jl:2-16) and MultigridRecon
 * (src/recon.jl:18-33), plus extension fields.  Fill with
 * baorec_params_default() and override.
 */
typedef struct baorec_params {
    float bias;
    float f;
    float smoothing_radius;
    float box_size[3];
    float box_min[3];
    int32_t has_los;          /* 0: los = nothing (local radial LOS) */
    float los[3];
    int32_t n_iter;           /* IterativeRecon.n_iter, default 3 */
    float beta;               /* f / bias */
    float jacobi_damping_factor; /* 0.4 */
    int32_t jacobi_niterations;  /* 5 */
    int32_t vcycle_niterations;  /* 6 */
    /* ---- extensions (defaults reproduce the reference) ---- */
    int32_t mas;              /* baorec_mas, default CIC */
    int32_t mg_schedule;      /* baorec_mg_schedule, default REFERENCE */
    double tolerance;         /* RBGS: relative residual target, 1e-6 */
    int32_t max_vcycles;      /* RBGS: cap on V-cycles, 200 */
    int32_t rb_sweeps;        /* RBGS: pre- and post-smoothing sweeps, 2 */
    int32_t fp64_coarse_max;  /* RBGS: levels of side <= this run in FP64, 8 */
    float box_pad;            /* run with randoms: setup_box pad, 500 */
    float ran_min;            /* randoms threshold factor, 0.01 */
    int32_t mutate_inputs;    /* write wrapped positions back (wrap=true), 1 */
    int32_t mg_warm_start;    /* RBGS: 0 from v = 0; 1 the spectral seed; 2 the learned seed;
                                 3 the spectral seed refined by spectral Richardson steps */
    int32_t mg_seed_steps;    /* RBGS, warm starts 2 and 3: spectral Richardson steps, 3 */
    /* ---- outputs of the last run ---- */
    int32_t vcycles_done;     /* RBGS: V-cycles used on the finest level */
    double final_residual;    /* RBGS: last relative residual ||f - A v|| / ||f|| */
    double final_change;      /* RBGS: last V-cycle's residual change / ||f|| */
    double seed_residual;     /* RBGS: relative residual of the starting v */
    double rhs_shift;         /* RBGS: the constant alpha subtracted from f */
} baorec_params;

/* Defaults of the Julia constructors (src/recon.jl:2-33); beta = f/bias in
 * FP32, box and los unset. */
void baorec_params_default(baorec_params *p, float bias, float f,
                           float smoothing_radius);

/* ------------------------------------------------------------- context */

/* Opens a context on HIP device `device` (one stream, cached work buffers).
 * Returns BAOREC_ERR_NO_DEVICE when there is no GPU. */
int baorec_create(baorec_ctx **ctx, int device);
int baorec_destroy(baorec_ctx *ctx);
/* Message of the last failed call on ctx (or of the last failed
 * baorec_create on this thread when ctx is NULL). */
const char *baorec_last_error(const baorec_ctx *ctx);
/* Library version string. */
const char *baorec_version(void);

/* ------------------------------------------------------------- run! */

/*
 * run!(::IterativeRecon, grid_size, x,y,z,w[, rx,ry,rz,rw])
 *   src/recon.jl:134-153 (periodic box) and :155-180 (with randoms).
 * Without randoms (rx == NULL, nr == 0) the box in `p` is used and CIC wraps:
 * x, y, z are mutated like the reference when p->mutate_inputs.  With randoms,
 * setup_box(randoms, p->box_pad) overwrites p->box_size / p->box_min.
 * The result (the reconstructed real-space overdensity) is kept in the
 * context as recon.result_cache.
 */
int baorec_run_iterative(baorec_ctx *ctx, baorec_params *p, const int32_t grid[3],
                         int mem, float *x, float *y, float *z, const float *w,
                         int64_t n, const float *rx, const float *ry,
                         const float *rz, const float *rw, int64_t nr);

/* run!(::MultigridRecon, ...), src/recon.jl:215-234 and :236-261; caches the
 * potential phi. */
int baorec_run_multigrid(baorec_ctx *ctx, baorec_params *p, const int32_t grid[3],
                         int mem, float *x, float *y, float *z, const float *w,
                         int64_t n, const float *rx, const float *ry,
                         const float *rz, const float *rw, int64_t nr);

/* recon.result_cache (src/recon.jl:153, 180, 234, 261).  The context keeps
 * the result of its last run!; each run! gives it a new generation number.
 * A binding that hands out several recons on one context records the
 * generation with each recon and passes it back: baorec_result_copy returns
 * BAOREC_ERR_NO_RESULT ("stale result") when a later run! has replaced that
 * result, instead of copying another recon's grid.  `generation` may be NULL
 * in baorec_result_info. */
int baorec_result_info(const baorec_ctx *ctx, int32_t grid[3], int32_t *kind,
                       uint64_t *generation);
int baorec_result_copy(baorec_ctx *ctx, int mem, float *dst, uint64_t generation);

/* ------------------------------------------------ grid-level entry points */

/* setup_overdensity!(delta, recon, x,y,z,w[, rx..rw]), src/recon.jl:42-57 and
 * :60-91; writes delta (nx*ny*nz floats). Uses the box in p. */
int baorec_setup_overdensity(baorec_ctx *ctx, const baorec_params *p,
                             const int32_t grid[3], int mem, float *delta,
                             float *x, float *y, float *z, const float *w,
                             int64_t n, const float *rx, const float *ry,
                             const float *rz, const float *rw, int64_t nr);

/* reconstructed_overdensity!(delta, recon::IterativeRecon, ...),
 * src/recon.jl:93-132. */
int baorec_reconstructed_overdensity(baorec_ctx *ctx, const baorec_params *p,
                                     const int32_t grid[3], int mem, float *delta,
                                     float *x, float *y, float *z, const float *w,
                                     int64_t n, const float *rx, const float *ry,
                                     const float *rz, const float *rw, int64_t nr);

/* reconstructed_potential!(phi, recon::MultigridRecon, ...),
 * src/recon.jl:184-212. */
int baorec_reconstructed_potential(baorec_ctx *ctx, baorec_params *p,
                                   const int32_t grid[3], int mem, float *phi,
                                   float *x, float *y, float *z, const float *w,
                                   int64_t n, const float *rx, const float *ry,
                                   const float *rz, const float *rw, int64_t nr);

/* ------------------------------------------------------ shifts / positions */

/*
 * read_shifts(recon, x,y,z[, field3d]; field), src/recon.jl:264-306 (CPU) and
 * :333-364 (GPU).  `field3d` == NULL uses the context's cached result (whose
 * displacement grids are kept, so repeated calls only gather).  Otherwise
 * field3d (grid[0]*grid[1]*grid[2] floats in `mem`) is the field and `kind`
 * says how to read it: BAOREC_ITERATIVE (delta, Psi_k = i k delta_k / k^2,
 * src/iterative.jl:229-275) or BAOREC_MULTIGRID (phi, Psi_k = i k phi_k,
 * src/multigrid.jl:755-799).  Writes n shifts per component.
 */
int baorec_read_shifts(baorec_ctx *ctx, const baorec_params *p, int mem,
                       const float *x, const float *y, const float *z, int64_t n,
                       const float *field3d, const int32_t grid[3], int kind,
                       int field, float *sx, float *sy, float *sz);

/* reconstructed_positions(recon, x,y,z[, field3d]; field),
 * src/recon.jl:366-388: out = pos - shift. */
int baorec_reconstructed_positions(baorec_ctx *ctx, const baorec_params *p, int mem,
                                   const float *x, const float *y, const float *z,
                                   int64_t n, const float *field3d,
                                   const int32_t grid[3], int kind, int field,
                                   float *ox, float *oy, float *oz);

/* ------------------------------------------------------------- primitives */

/* run!(::IterativeRecon, grid, x, y, z, w) followed by
 * reconstructed_positions(recon, x, y, z; field = :sum) (src/recon.jl:134-180,
 * 366-388), replayed as a HIP graph (an extension; the reference's drivers
 * call the two per catalog, examples/simulation.jl:22-56).  The first call with
 * a given set of arguments (device pointers, n, grid, the input fields of p,
 * the autotuned choices) runs eagerly and captures the launch sequence;
 * later calls with the same arguments replay it, without the per-launch
 * host cost.  A replay is discarded when a library buffer it uses has been
 * reallocated since.  Periodic box only (no randoms: the box comes from p),
 * device memory only; results and errors as the two calls. */
int baorec_run_positions_graph(baorec_ctx *ctx, baorec_params *p, const int32_t grid[3],
                               float *x, float *y, float *z, const float *w, int64_t n,
                               float *ox, float *oy, float *oz);

/* setup_box(x, y, z, pad), src/utils.jl:100-109. */
int baorec_setup_box(baorec_ctx *ctx, int mem, const float *x, const float *y,
                     const float *z, int64_t n, float pad, float box_size[3],
                     float box_min[3]);

/* cic!(rho, x,y,z,w, box_size, box_min; wrap), src/mas.jl:1-107: ADDS the
 * mass of n particles to rho (p->mas selects CIC or TSC).  wrap mutates x,y,z
 * when p->mutate_inputs. */
int baorec_cic(baorec_ctx *ctx, const baorec_params *p, const int32_t grid[3],
               int mem, float *rho, float *x, float *y, float *z, const float *w,
               int64_t n, int wrap);

/* read_cic!(out, field, x,y,z, box_size, box_min), src/mas.jl:271-327
 * (wrap = true). */
int baorec_read_cic(baorec_ctx *ctx, const baorec_params *p, const int32_t grid[3],
                    int mem, float *out, const float *field, const float *x,
                    const float *y, const float *z, int64_t n);

/* smooth!(field, R, box_size, plan), src/utils.jl:85-96 (in place).  The
 * Gaussian exp(-R^2 k^2 / 2) is applied one axis at a time, exp(-R^2 k_a^2 / 2)
 * per axis with the exponent in FP64: the same function as the reference's
 * single 3D factor, within FP32 rounding. */
int baorec_smooth(baorec_ctx *ctx, const baorec_params *p, const int32_t grid[3],
                  int mem, float *field);

/* fft_plan * field and ldiv!(field, fft_plan, field_k) (setup_fft!,
 * src/recon.jl:35-37): forward R2C into (nx/2+1)*ny*nz interleaved complex
 * (re, im) floats, and normalised inverse C2R (the input is not modified). */
int baorec_rfft3d(baorec_ctx *ctx, const int32_t grid[3], int mem, const float *in,
                  float *out_complex);
int baorec_irfft3d(baorec_ctx *ctx, const int32_t grid[3], int mem,
                   const float *in_complex, float *out);

/* k_vec / x_vec (src/utils.jl:3-10, 21-25) as used by the kernels:
 * kx has nx/2+1 entries, ky ny, kz nz; xv* hold the cell centres. Host only. */
int baorec_k_vec(const int32_t grid[3], const float box_size[3], float *kx,
                 float *ky, float *kz);
int baorec_x_vec(const int32_t grid[3], const float box_size[3],
                 const float box_min[3], float *xx, float *xy, float *xz);

/* Multigrid pieces on one level (all in place / out of place as named):
 *   jacobi!   src/multigrid.jl:193-210    v <- niter damped sweeps
 *   residual! src/multigrid.jl:378-395    r = f - A v
 *   reduce!   src/multigrid.jl:641-651    coarse (grid/2) <- fine
 *   prolong!  src/multigrid.jl:511-518    fine <- coarse (grid = coarse side),
 *                                          add != 0 adds instead of overwriting
 *   vcycle!   src/multigrid.jl:689-719
 *   fmg       src/multigrid.jl:722-752    v overwritten
 * The box, beta, los, damping and sweep counts come from p. */
int baorec_mg_jacobi(baorec_ctx *ctx, const baorec_params *p, const int32_t grid[3],
                     int mem, float *v, const float *f);
/* The RBGS schedule's smoother (an extension: the reference smooths with
 * jacobi!, src/multigrid.jl:193-210): `sweeps` multicolour Gauss-Seidel
 * sweeps in place on v, the colours ((ix+iy)&1) + 2((iy+iz)&1) in order 0..3,
 * each cell v = (f + neighbour sums of src/multigrid.jl:170-185) / diagonal.
 * fused = 0: one launch per colour; fused = 1: one launch per sweep (sides
 * multiples of 64, 16 and 8, else BAOREC_ERR_UNSUPPORTED), bit for bit the
 * same result.  Sides must be even. */
int baorec_mg_gauss_seidel(baorec_ctx *ctx, const baorec_params *p, const int32_t grid[3],
                           int mem, float *v, const float *f, int sweeps, int fused);
int baorec_mg_residual(baorec_ctx *ctx, const baorec_params *p,
                       const int32_t grid[3], int mem, float *r, const float *v,
                       const float *f);
int baorec_mg_reduce(baorec_ctx *ctx, const int32_t fine_grid[3], int mem,
                     float *coarse, const float *fine);
/* The V-cycle's step down (src/multigrid.jl:700-704): coarse = reduce!(
 * residual!(v, f)) on a grid of even sides, coarse of grid/2.  fused = 0: the
 * two kernels of baorec_mg_residual and baorec_mg_reduce through a scratch
 * grid; fused = 1: one pass that does not store the residual (sides multiples
 * of 64, 16 and 2 and fewer than 2^31 cells, else BAOREC_ERR_UNSUPPORTED),
 * bit for bit the same result. */
int baorec_mg_residual_restrict(baorec_ctx *ctx, const baorec_params *p, const int32_t grid[3],
                                int mem, float *coarse, const float *v, const float *f, int fused);
int baorec_mg_prolong(baorec_ctx *ctx, const int32_t coarse_grid[3], int mem,
                      float *fine, const float *coarse, int add);
int baorec_mg_vcycle(baorec_ctx *ctx, const baorec_params *p, const int32_t grid[3],
                     int mem, float *v, const float *f);
int baorec_mg_fmg(baorec_ctx *ctx, baorec_params *p, const int32_t grid[3], int mem,
                  float *v, const float *f);

/* ------------------------------------------------ sky <-> Cartesian */

/*
 * Parameters of BAOrec.Cosmology (src/cosmo.jl:22-72): the Float32 fields of
 * the Julia struct; the library derives the densities exactly as the Julia
 * constructor does (photons from T_cmb, neutrinos from Neff, Omega_Lambda
 * closing the budget).  The comoving-distance table of
 * comoving_distance_interp / redshift_interp (src/cosmo.jl:86-103) has
 * z_tab_num knots spanning [z_tab_min, z_tab_max] and is built on the GPU
 * (cached per context and parameter set).
 */
typedef struct baorec_cosmology {
    float omega_b, omega_c, h, Neff, Omega_k0, T_cmb, w0, wa;
    double z_tab_min, z_tab_max;
    int32_t z_tab_num;
} baorec_cosmology;

/* Cosmology() and DESICosmology() defaults (src/cosmo.jl:22-62), z_tab_max 3. */
void baorec_cosmology_default(baorec_cosmology *c);
void baorec_cosmology_desi(baorec_cosmology *c);

/* comoving_distance_interp(c)(z) (src/cosmo.jl:86-93): linear interpolation
 * in the table, Mpc, FP64 in and out.  A z outside the table returns
 * BAOREC_ERR_OUT_OF_BOX (the reference's BoundsError). */
int baorec_comoving_distance(baorec_ctx *ctx, const baorec_cosmology *c, int mem,
                             const double *z, int64_t n, double *r);

/* sky_to_cartesian(ra, dec, z, cosmo) of the reference's drivers
 * (examples/lightcone.jl:30-48): degrees and redshift in, Mpc/h out. */
int baorec_sky_to_cartesian(baorec_ctx *ctx, const baorec_cosmology *c, int mem,
                            const float *ra, const float *dec, const float *red,
                            int64_t n, float *x, float *y, float *z);

/*
 * Allocation-free sky_to_cartesian.  No counterpart in the reference: for
 * callers that cannot allocate device memory during the call (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * conversion and results as baorec_sky_to_cartesian with mem =
 * BAOREC_MEM_DEVICE: ra, dec, red, x, y, z are device pointers on the
 * context's GPU, and the call makes no hipMalloc, hipFree or other allocating
 * call.  Its distance table (zk, rk, interval integrals, guide) and error flag
 * live in a static device array of 3.67 MB that each GPU the library runs on
 * holds until the process ends; z_tab_num is therefore limited to 131072
 * knots (more returns BAOREC_ERR_UNSUPPORTED).  The table is kept between
 * calls and rebuilt when the cosmology changes.  All calls in a process share
 * that array per GPU: the library serialises them with a lock, and two
 * processes each have their own copy.
 */
int baorec_sky_to_cartesian_noalloc(baorec_ctx *ctx, const baorec_cosmology *c,
                                    const float *ra, const float *dec, const float *red,
                                    int64_t n, float *x, float *y, float *z);

/* cartesian_to_sky(x, y, z, cosmo) (examples/lightcone.jl:50-80): RA is
 * rem(lon - 360, 360) as the driver computes it (values in (-360, 0]). */
int baorec_cartesian_to_sky(baorec_ctx *ctx, const baorec_cosmology *c, int mem,
                            const float *x, const float *y, const float *z,
                            int64_t n, float *ra, float *dec, float *red);

/* ------------------------------------------------------- power spectrum */

/*
 * Multipoles P_0, P_2, P_4 of a periodic box catalog, the estimator of the
 * reference's validation runs (test_helpers/powspec_auto.conf: cubic box,
 * TSC, grid interlacing, multipoles 0 2 4, linear bins of width 0.005 from
 * k = 0).  The weights are assigned to the grid (and, with interlacing, to a
 * second grid holding the particles shifted by half a cell), both are
 * Fourier transformed, combined with the half-cell phase, divided by the
 * assignment window prod_a sinc(k_a H_a / 2)^p (p = 2 CIC, 3 TSC), and
 * |delta_k|^2 L_l(mu) is averaged per |k| bin over the modes of the full
 * k-space; mu = k.los / |k|.  P_0 has the shot noise V sum(w^2) / sum(w)^2
 * subtracted.  This is an extension: its reference is a third-party code
 * (pypowspec) that is not part of BAOrec.jl.
 */
typedef struct baorec_pk_params {
    float box_size[3];   /* periodic box, positions in [box_min, box_min + box_size) */
    float box_min[3];
    float los[3];        /* line of sight (normalised by the call), default z */
    int32_t mas;         /* BAOREC_MAS_CIC or BAOREC_MAS_TSC (default) */
    int32_t interlace;   /* 1: interlaced grids (default) */
    double kmin, dk;     /* bin b covers [kmin + b dk, kmin + (b+1) dk) */
    int32_t nbins;       /* 1 .. 1600 */
} baorec_pk_params;

/* box 1000^3 at the origin, los z, TSC, interlacing, kmin 0, dk 0.005, 300 bins */
void baorec_pk_params_default(baorec_pk_params *p);

/* x, y, z, w (w may be NULL: unit weights) in `mem`; the bin outputs (nbins
 * doubles each: mean |k| of the modes, number of modes, P_0, P_2, P_4 in
 * (Mpc/h)^3) are host arrays.  Empty bins give 0 modes and NaN. */
int baorec_power_multipoles(baorec_ctx *ctx, const baorec_pk_params *pk, const int32_t grid[3],
                            int mem, const float *x, const float *y, const float *z,
                            const float *w, int64_t n, double *k_mean, double *nmodes,
                            double *p0, double *p2, double *p4);

/* ---------------------------------------------------------- catalog I/O */
/* Host-only calls (no context, no GPU); a failure's message is in
 * baorec_io_last_error() (per thread). */

/*
 * ASCII catalogs as the reference's drivers read them with CSV.File(path;
 * delim = " ", types = Float32) (test/benchmarks_cpu.jl:22-23,
 * examples/lightcone.jl:20-21): one galaxy per line, whitespace-separated
 * numeric columns, every value parsed to the nearest Float32.  Blank lines
 * and lines starting with '#' are skipped.
 * baorec_catalog_rows counts the data lines (and the columns of the first).
 * baorec_catalog_read parses columns cols[0..nsel-1] (0-based) into the host
 * arrays out[0..nsel-1] (capacity rows each); a short or malformed line is
 * BAOREC_ERR_INVALID_ARGUMENT naming the line.
 */
int baorec_catalog_rows(const char *path, int64_t *nrows, int32_t *ncols);
int baorec_catalog_read(const char *path, int32_t nsel, const int32_t *cols,
                        float *const *out, int64_t capacity, int64_t *nrows);

/* npzwrite(path, hcat(cols...)) of the drivers (examples/lightcone.jl:
 * 125-127): ncols float32 columns of n rows as an .npy array of shape
 * (n, ncols), Fortran order (columns contiguous, as NPZ.jl writes a Julia
 * Matrix{Float32}). */
int baorec_write_npy(const char *path, int32_t ncols, const float *const *cols, int64_t n);

/* Message of the last failed catalog call on this thread. */
const char *baorec_io_last_error(void);

/* --------------------------------------------------------------- batching */

/*
 * Extension (no reference counterpart; the reference's drivers loop over
 * catalogs): for each of `ncat` periodic-box catalogs, the equivalent of
 *   run!(IterativeRecon, grid, x, y, z, w)                     src/recon.jl:134
 *   reconstructed_positions(recon, x, y, z; field = :sum)      src/recon.jl:366
 * with every launch covering a group of catalogs.  Catalog c occupies
 * particles [offsets[c], offsets[c+1]) of x, y, z, w (offsets is a host
 * array of ncat + 1 entries, offsets[0] = 0, non-decreasing); all catalogs
 * share the grid and the params (box_size / box_min must be set; no randoms).
 * Outputs: the reconstructed positions of every particle in ox, oy, oz.
 * Inputs are wrapped in place when mutate_inputs, as run! does.  Results are
 * bitwise those of the single-catalog calls.  The sort and mass assignment
 * run over scatter groups of up to 16 catalogs (density grids within 8 GiB,
 * at most 2^29 placements); the rest runs over FFT subgroups of a scatter
 * group whose grid and spectrum buffers stay within 2 GiB (0.80 GB per
 * catalog at 256^3, so pairs of catalogs; at least one catalog per group),
 * all on the context's stream.  A particle outside the grid gives
 * BAOREC_ERR_OUT_OF_BOX.
 */
int baorec_run_batch(baorec_ctx *ctx, baorec_params *p, const int32_t grid[3],
                     int mem, int ncat, const int64_t *offsets, float *x, float *y,
                     float *z, const float *w, float *ox, float *oy, float *oz);

/* ------------------------------------------------------------- autotuning */

/*
 * Extension, no reference counterpart (the reference has no launch shapes to
 * tune).  Times the library's runtime choices on this device with the given
 * catalog and keeps the fastest in the context, where later calls use them:
 *  - the scatter tile shape (every tile of 16..64 x 8..64 x 8..64 cells that
 *    fits the LDS tile and the sort's histogram) for this grid, mass
 *    assignment and catalog size class (floor(log2 n));
 *  - per FFT line length >= 256 and axis, the kernel of single-output strided
 *    passes (one or two columns per thread).
 * The transforms are identical for every choice; the density grids agree
 * to the fixed-point quantum of the scatter (2^-61 of a tile's total weight,
 * in practice bitwise).  Needs box_size and box_min in params; the inputs
 * are not modified.
 */
int baorec_autotune(baorec_ctx *ctx, const baorec_params *p, const int32_t grid[3], int mem,
                    const float *x, const float *y, const float *z, const float *w, int64_t n);
/* The tuned choices as JSON into buf; returns the length needed including
 * the terminating NUL. */
int64_t baorec_tune_report(baorec_ctx *ctx, char *buf, int64_t len);
/* Writes the tuned choices to a JSON file / adds those of a file written by
 * baorec_tune_save (entries the library cannot use are rejected). */
int baorec_tune_save(baorec_ctx *ctx, const char *path);
int baorec_tune_load(baorec_ctx *ctx, const char *path);

/* ---------------------------------------------------- learned warm start */

/*
 * Extension, no reference counterpart (BAOrec.jl's fmg starts every level
 * from v = 0, src/multigrid.jl:722-752).  A 3D U-Net held in the context
 * predicts, from the right-hand side f of the multigrid solve, the
 * correction that turns the spectral seed (mg_warm_start = 1) into the
 * local-LOS potential; mg_warm_start = 2 starts RBGS from spectral seed +
 * correction, and the V-cycles then run to the same tolerance as a cold
 * start.  Network: levels of 16, 32, 64, 128 channels, two 3^3 convolutions
 * (+ bias, ReLU) per level down and up, 2x2x2 average pooling, nearest
 * upsampling concatenated after the skip input, a 1x1 head; BF16 weights and
 * activations on the matrix cores, FP32 accumulation, periodic padding.
 * Input features per cell (8 channels): f / s, phi_s / (s h^2) (phi_s the
 * spectral seed), r_s / s (its residual under the local-LOS operator),
 * psi / (s h^2) (psi the spectral inverse of r_s), rhat - nhat (local LOS minus
 * the seed's fixed LOS, 3), 100 h / |x|; s is the rms of f, h the cell size.
 * Output: the correction / (s h^2).  In the solve the network runs on the pyramid level
 * whose cell is closest to h_train (f restricted there) and its correction is
 * prolonged back.
 */
/* Number of parameters (FP32): per layer W[27][cin][cout] then b[cout], and
 * the head's 16 weights and bias. */
int64_t baorec_warmstart_nparams(void);
/* A fresh network with He-normal weights from `seed`; h_train is the cell
 * size (Mpc/h) of the grids it will be trained on. */
int baorec_warmstart_init(baorec_ctx *ctx, uint64_t seed, float h_train);
/* Network file: "BRUNET01", nparams, Adam step, h_train, parameters (FP32). */
int baorec_warmstart_save(baorec_ctx *ctx, const char *path);
int baorec_warmstart_load(baorec_ctx *ctx, const char *path);
/* which = 0: copy the parameters to host[n]; 1: set them from host[n]
 * (resets Adam); 2: copy the gradients of the last step / eval. */
int baorec_warmstart_params(baorec_ctx *ctx, int which, float *host, int64_t n);
/* One training step on nbatch pairs (f, phi) of grid-sized fields (phi the
 * converged local-LOS potential of f): features, target
 * (phi - phi_spectral) / (s h^2), forward, loss = mean square of the
 * finite-difference gradient of (output - target), backward, and an Adam step
 * of rate lr (lr = 0: gradients only).  Box, beta and LOS come from p. */
int baorec_warmstart_step(baorec_ctx *ctx, const baorec_params *p, const int32_t grid[3], int mem,
                          const float *f, const float *phi, int32_t nbatch, float lr, double *loss);
/* The network on given features (nbatch x cells x 8 BF16 bit patterns,
 * channels last) against a given target: output (FP32), loss, and the
 * parameter gradients (read with baorec_warmstart_params(ctx, 2, ...)). */
int baorec_warmstart_eval(baorec_ctx *ctx, const int32_t grid[3], int mem, const uint16_t *features,
                          const float *target, int32_t nbatch, float *out, double *loss);
/* A tensor of the last eval / step (BF16 bit patterns, channels last), for
 * checking each layer on its own inputs: kind 0 the output of layer `index`
 * (0..13), 1 its gradient after the ReLU mask, 2 pooled tensor `index`
 * (0..2), 3 its gradient, 4 the input features.  n must be its length. */
int baorec_warmstart_tensor(baorec_ctx *ctx, int kind, int index, uint16_t *host, int64_t n);
/* The learned seed of a multigrid solve on this grid (spectral seed + the
 * network's correction), as mg_warm_start = 2 starts from it. */
int baorec_warmstart_predict(baorec_ctx *ctx, const baorec_params *p, const int32_t grid[3], int mem,
                             const float *f, float *seed);

/* ------------------------------------------------------------- profiling */

/* Extension, no reference counterpart (the reference times whole calls with
 * BenchmarkTools): when enabled, every kernel launch on ctx is bracketed with
 * HIP events on the context's stream and aggregated per kernel family with
 * its algorithmic HBM bytes.  Enabling or disabling clears the tallies. */
int baorec_profile(baorec_ctx *ctx, int enable);
/* JSON {"name": {"count": c, "ms": total, "bytes": total}, ...} into buf;
 * returns the length needed including the terminating NUL. */
int64_t baorec_profile_report(baorec_ctx *ctx, char *buf, int64_t len);

#ifdef __cplusplus
}
#endif

#endif /* BAOREC_H */
