/*
 * mimadv.h -- C ABI of the MI355X advection library (libmimadv.so).
 *
 * The library replaces the advection operator of the MIMICA LES model:
 *   module advection (src/TOOLS/advection.f90:79, public :: advect,
 *     advection_x, advection_y, advection_z, advection_ac, ...) and its
 *     per-line kernels src/TOOLS/advq_{weno,ppm,muscl}.f90;
 *   module windadv   (src/INTEG/windadv.f90:41, public :: qnadv_uvw).
 * The Fortran shim fdd3260_amd/fortran/mimadv_advection.f90 keeps those
 * routine names and argument meanings and forwards here through
 * iso_c_binding.  Every entry point returns an int status: MIMADV_OK (0) or
 * a negative MIMADV_ERR_* code; mimadv_last_error() gives the message.
 *
 * Data layout (all entry points): Fortran column-major 3D arrays with x
 * fastest, extents (nx, ny, nz) INCLUDING the reference's lateral ghosts:
 * nx = MAXX+5 (3 west + 2 east ghosts), ny = MAXY+5 or 1 for a 2D run, no
 * vertical ghosts (src/TYPES/typedef_state.f90:85-87, INCLUDE/wgrid.csh:17-23).
 * Interior columns are i = 4..nx-2, j = 4..ny-2 (1-based).  C-grid
 * staggering: u(i) on the west face of cell i, v(j) on the south face, w(k)
 * on the bottom face.  Pointers are device pointers unless the call's flags
 * contain MIMADV_HOST_PTRS, in which case the library stages the arrays
 * through device memory and synchronises before returning.
 *
 * Threading: one context per host thread; a context owns one HIP stream and
 * its device buffers.  Calls on one context are not reentrant.
 */
#ifndef MIMADV_H
#define MIMADV_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define MIMADV_OK 0
#define MIMADV_ERR_ARG (-1)          /* invalid argument / shape */
#define MIMADV_ERR_HIP (-2)          /* HIP runtime failure */
#define MIMADV_ERR_UNSUPPORTED (-3)  /* option combination not implemented */
#define MIMADV_ERR_NOMEM (-4)        /* device allocation failed */

/* ---- boundary codes: bcl(6) of namelist cm_bc (shared_data.f90:325) ---- */
#define MIMADV_BC_PER 0 /* 'per' */
#define MIMADV_BC_NNF 1 /* 'nnf' */
#define MIMADV_BC_OPE 2 /* 'ope' */

/* ---- scalar schemes: scal_adv (shared_data.f90:320) ---- */
#define MIMADV_WENO 0  /* 'weno': the reference's 3rd-order switch WENO */
#define MIMADV_PPM 1   /* 'ppm' */
#define MIMADV_MUSCL 2 /* 'muscl' */
#define MIMADV_WENO5 3 /* WENO5-JS (no reference counterpart); in z on faces whose
                          5-cell upwind stencil lies in the column, the 'weno'
                          face elsewhere */
#define MIMADV_QUICK 4 /* 'quick' (advq_quick.f90) */
#define MIMADV_LW 5    /* Lax-Wendroff with the MC limiter (advq_lw.f90; also any other scal_adv) */

/* ---- MUSCL limiters: muscllimiter (advq_muscl.f90:49) ---- */
#define MIMADV_LIM_MINMOD 0
#define MIMADV_LIM_MC 1
#define MIMADV_LIM_VL 2
#define MIMADV_LIM_FROMM 3

/* advsz_weno reads x(0) on its first loop pass (advq_weno.f90:245-247). */
#define MIMADV_WENO_BOTTOM_X0_IS_X1 0 /* x(0) := x(1) (default) */
#define MIMADV_WENO_BOTTOM_CLOSURE 1  /* keep the k=1 closure flux */

/* ---- time integrators ---- */
#define MIMADV_INT_EULER 1  /* sorder = 1 (integ.h) */
#define MIMADV_INT_HEUN 2   /* sorder = 2: (a_lm,b_lm) = (1,0),(0.5,1) (integ.h:1-12) */
#define MIMADV_INT_SSPRK3 3 /* Shu-Osher SSP-RK3 (no reference counterpart) */

/* ---- call flags ---- */
#define MIMADV_DIR_X 1u
#define MIMADV_DIR_Y 2u
#define MIMADV_DIR_Z 4u
#define MIMADV_DIR_ALL 7u
#define MIMADV_OVERWRITE 8u    /* xs = -div q (xs is not read) */
#define MIMADV_HOST_PTRS 16u   /* array arguments are host memory */
/* limit = .false. of advection_{x,y,z}: unlimited QUICK for MIMADV_QUICK,
 * unlimited Lax-Wendroff for every other scheme (advection.f90:580-592,
 * advq_quick_nolim.f90, advq_lw_nolim.f90) */
#define MIMADV_UNLIMITED 128u
/* sorder = 1: limited Lax-Wendroff whatever the scheme (advection.f90:598-601) */
#define MIMADV_SORDER1 256u
/* Positivity-preserving stages for mimadv_advance (no reference
 * counterpart): every Euler substep q + dt*L(q) of the Shu-Osher form
 * scales each cell's outgoing face fluxes by one factor theta = 1, or
 * (1 - 2^-40) max(q,0) / (dt * outgoing flux per width) where that flux
 * exceeds the content, applied to the
 * fluxes the cell donates (left/lower cell for a positive flux, right/upper
 * for a negative one).  Mass is conserved, non-negative fields stay
 * non-negative, and where no cell needs it (theta = 1 everywhere) the step
 * is the reference step bit for bit.  Scheme 'weno' in flux form, periodic
 * sides, 'nnf' floor and lid, integrators MIMADV_INT_SSPRK3 and
 * MIMADV_INT_HEUN (Heun in its Shu-Osher form 1/2 q + 1/2 (q* + dt L(q*))). */
#define MIMADV_POSITIVE 2048u

/* Grid: extents, spacings and vertical metrics (define_grid.f90:179-187).
 * Metric arrays are host arrays of length nz; they are copied at create. */
typedef struct mimadv_grid {
    int nx, ny, nz;
    double dx, dy, dz;
    const double *fdz0;  /* fdz0(k) */
    const double *dzw;   /* dz*fdz0(k) */
    const double *dz1w;  /* 1/dzw(k) */
    const double *dz1;   /* 1/(dz*(fdz0(k)+fdz0(k+1))/2), dz1(nz)=dz1(nz-1) */
    int bcl[6];          /* MIMADV_BC_*: west, east, south, north, bottom, top */
} mimadv_grid;

/* Options: the namelist values the reference reads as module globals
 * (shared_data.f90:251-349, advq_muscl.f90:49, advq_ppm.f90:49). */
typedef struct mimadv_opts {
    int scheme;         /* MIMADV_WENO / PPM / MUSCL / WENO5 / QUICK / LW */
    int muscl_limiter;  /* MIMADV_LIM_* (reference default minmod) */
    int ppm_limiter;    /* 1 = 'PPM' selective limiter (default), 0 = none */
    double lim_tol;     /* reference default 1e-9 */
    int weno_bottom;    /* MIMADV_WENO_BOTTOM_* */
    int mom_ord;        /* 2 or 4 (default 4) */
    int lavisc;         /* momentum artificial viscosity (default 1) */
} mimadv_opts;

typedef struct mimadv_ctx mimadv_ctx;

/* Create a context on the current HIP device.  Replaces advection_init
 * (advection.f90:1045-1067) plus the module globals the reference reads. */
int mimadv_create(const mimadv_grid *grid, const mimadv_opts *opts, mimadv_ctx **out);
void mimadv_destroy(mimadv_ctx *ctx);
const char *mimadv_last_error(const mimadv_ctx *ctx);
/* Run subsequent work on a caller stream (hipStream_t).  NULL is HIP's null
 * (default) stream, the one a framework's default stream maps to;
 * MIMADV_OWN_STREAM returns to the context's own non-blocking stream (the
 * initial setting), which is not ordered with any other stream. */
#define MIMADV_OWN_STREAM ((void *)-1)
int mimadv_set_stream(mimadv_ctx *ctx, void *stream);
int mimadv_synchronize(mimadv_ctx *ctx);
/* Change the scheme-related options of an existing context. */
int mimadv_set_opts(mimadv_ctx *ctx, const mimadv_opts *opts);

/* Scalar advective tendency for nf fields sharing one velocity (u, v, w):
 *   xs_f = ((xs_f - X_f) - Y_f) - Z_f
 * where X/Y/Z are the per-direction flux divergences the reference computes
 * in advection_x/y/z (advection.f90:556-798; `xs = xs - xx` at :625, :706,
 * :788), restricted to the directions in flags & MIMADV_DIR_ALL.  adv_form[f]
 * (NULL = all 0) selects the advective form (`adv` argument of advs*_*).
 * dt0 is the full step; PPM and Lax-Wendroff use it (advq_ppm.f90:81-82,
 * advq_lw.f90:88).  flags may add MIMADV_UNLIMITED (limit = .false.) and
 * MIMADV_SORDER1 (sorder = 1), which select the line kernel as
 * advection_{x,y,z} does.  For ny == 1 the y direction is skipped as in a
 * 2D build.  Ghost values are read as given by the caller; ghost tendencies
 * are left untouched.
 * Replaces: advection_x/y/z (advection.f90:556,637,719), advection_ac
 * (advection.f90:801, xs += xsx+xsy+xsz), and the loops of advect /
 * qnadv_scal_{x,y,z} (advection.f90:84-553). */
int mimadv_scalar_tendency(mimadv_ctx *ctx, double dt0, const double *u, const double *v,
                           const double *w, int nf, const double *const *q, double *const *xs,
                           const int *adv_form, unsigned flags);

/* Momentum flux divergence (qnadv_uvw, windadv.f90:46-79): ua, va, wa are
 * OVERWRITTEN with +div(rho0 u u) of the momenta mu, mv, mw advected by the
 * velocities ut, vt, wt (the reference's wind_adv); ntime selects the
 * direction order (windadv.f90:103,158).  For ny == 1, mv/va are unused. */
int mimadv_momentum_tendency(mimadv_ctx *ctx, int ntime, const double *mu, const double *mv,
                             const double *mw, const double *ut, const double *vt, const double *wt,
                             double *ua, double *va, double *wa, unsigned flags);

/* Ghost fill of nf scalar fields (qnbc_p4, bc.f90:378-503, serial branch). */
int mimadv_fill_halo(mimadv_ctx *ctx, int nf, double *const *x, unsigned flags);

/* Stage combine of integ_scal (integpack.f90:174-227) for nf fields:
 * sorder > 1: tend = a + b_lm*tend; x = xold + a_lm*dt*tend with (a_lm,b_lm)
 * of integ.h for stage im; sorder == 1: x = xold + dt*a.  Whole arrays. */
int mimadv_integ_scal(mimadv_ctx *ctx, int sorder, int im, double dt, int nf,
                      const double *const *xold, const double *const *a, double *const *tend,
                      double *const *x, unsigned flags);

/* Ghost fill of the winds u, v, w (lbcu_p4, bc.f90:81-344, serial branch
 * of a single tile): periodic images in x, y and z per the context's bcl,
 * then on 'nnf' / 'ope' sides the normal component (u in x, v in y) is
 * mirrored with a sign change / copied into the ghosts (u(3:1) =
 * +-u(4:6), u(nx-1:nx) = +-u(nx-2), +-u(nx-3)); the tangential components
 * keep their ghost values there, as in the reference's single-tile run.
 * v is unused when ny == 1. */
int mimadv_fill_winds(mimadv_ctx *ctx, double *u, double *v, double *w, unsigned flags);

/* Momentum stage combine integ_mom (integpack.f90:54-135) with integ.h, on
 * the component triples old/a/tend/x = (u, v, w) (whole arrays; v unused in
 * 2D; x may alias old): sorder > 1: tend = a + b_lm*tend, x = old +
 * a_lm*dt*tend; sorder == 1: x = old + dt*a; w is then multiplied by the
 * implicit sponge factor when sponge == 2: tauw = 0.25*clamp(cos(pi/2*(zmax -
 * z0(k))/(zmax - zdamp))^2) of the LAST level k with z0(k) >= zdamp (the
 * reference's level loop assigns the whole tauw array), factor 1/(1 +
 * a_lm*dt*tauw) for every level (1/(1 + a_lm*dt) when no level reaches
 * zdamp), pi = 3.14159265.  z0 is a host array of nz level heights. */
int mimadv_integ_mom(mimadv_ctx *ctx, int sorder, int im, double dt, const double *z0, double zmax, double zdamp,
                     int sponge, const double *const *old, const double *const *a, double *const *tend,
                     double *const *x, unsigned flags);

/* totalav (averages.f90:507-551): unweighted interior mean of x; also the
 * dzw-weighted interior sum (the quantity flux form conserves). */
int mimadv_totalav(mimadv_ctx *ctx, const double *x, double *mean, double *dzw_sum, unsigned flags);

/* Advance nf scalar fields by nsteps steps of the advection-only dynamics
 * loop (time_step.f90:238-284 with integrate_all :337-444 and the ghost
 * refresh statebc :405-413): per stage, tendency of every field, stage
 * combine and periodic ghost refresh, all device-resident.  integrator is
 * MIMADV_INT_*: MIMADV_INT_EULER (sorder = 1, which runs the limited
 * Lax-Wendroff scheme, or the unlimited one with MIMADV_UNLIMITED),
 * MIMADV_INT_HEUN or MIMADV_INT_SSPRK3.  Lateral boundaries must be
 * periodic.  q holds the fields (in place); dt is also the dt0 of PPM and
 * Lax-Wendroff.  flags may contain MIMADV_UNLIMITED or MIMADV_POSITIVE. */
int mimadv_advance(mimadv_ctx *ctx, int integrator, int nsteps, double dt, const double *u,
                   const double *v, const double *w, int nf, double *const *q, const int *adv_form,
                   unsigned flags);

/* ---- density weighting of the anelastic build (funcpack.f90:65-542) ----
 * The reference density profile den0(k) and its face average avden0(k) are
 * the module globals of shared_data; they are handed over once per context.
 * avden0 may be NULL: it is then computed as initial.f90:543-546 does,
 * avden0(1) = den0(1), avden0(k) = (fdz0(k) den0(k) + fdz0(k-1) den0(k-1)) /
 * (fdz0(k) + fdz0(k-1)).  Host arrays of nz values. */
int mimadv_set_reference_density(mimadv_ctx *ctx, const double *den0, const double *avden0);

/* get_momentum (inverse = 0, funcpack.f90:99-178): u, v *= den0(k),
 * w *= avden0(k); get_velocities (inverse = 1, :236-316): the same
 * divisions.  In place, whole arrays; v unused when ny == 1. */
int mimadv_get_momentum(mimadv_ctx *ctx, double *u, double *v, double *w, int inverse, unsigned flags);

/* get_advective (funcpack.f90:65-92): ua = u/den0(k) - u0shift,
 * va = v/den0(k) - v0shift, wa = w/avden0(k) from the momenta u, v, w. */
int mimadv_get_advective(mimadv_ctx *ctx, double u0shift, double v0shift, const double *u, const double *v,
                         const double *w, double *ua, double *va, double *wa, unsigned flags);

/* get_conserved (inverse = 0, funcpack.f90:397-454): q_f *= dens;
 * get_scalars (inverse = 1, :457-514): q_f /= dens; dens is the 3D density
 * array (pressurel%dens).  In place on nf fields. */
int mimadv_get_conserved(mimadv_ctx *ctx, const double *dens, int nf, double *const *q, int inverse,
                         unsigned flags);

/* ---- large-scale vertical advection (advection.f90:918-1042) ----
 * advection_ac_1d: for nf fields, xs_f -= dref*w0*grad_1d(q_f/dref) on the
 * interior columns, with grad_1d (gradients.f90:400-473) the derivative of
 * the quadratic through levels k-1..k+1 (the floor reads x(1) at distance
 * dz*fdz0(1); level nz gets no tendency).  w0 is a host array of nz values
 * (the subsidence profile w0).  dref is the context's den0 profile with
 * MIMADV_DREF_DEN0 (CONSERVATIVE + ANELASTIC build; needs
 * mimadv_set_reference_density), else the 3D array dens when not NULL
 * (pressure%dens of the CONSERVATIVE compressible build), else 1.
 * advection_1d (:918-990) is the same call on es (and qt with with_lssub). */
#define MIMADV_DREF_DEN0 512u
int mimadv_advection_1d(mimadv_ctx *ctx, const double *w0, const double *dens, int nf, const double *const *q,
                        double *const *xs, unsigned flags);

/* ---- batched mixed-precision path (no reference counterpart: the
 * reference advects field by field in FP64) ---- */
#define MIMADV_F64 0
#define MIMADV_F32 1
#define MIMADV_GRAPH 32u /* capture one step in a hipGraph and replay it */
/* q[0], q[1], q[2] (MIMADV_F32) are the momenta rho0*u, rho0*v, rho0*w,
 * advanced with the momentum operator of qnadv_uvw (windadv.f90:46-1127;
 * the context's mom_ord and lavisc, directions in the even-ntime order) by
 * the batch's velocities instead of with the scalar scheme; FP32
 * arithmetic with the level metrics as precomputed factors, the stage
 * update in FP64 rounded once; 'nnf' floor and lid.  In 2D q[1] is not
 * touched. */
#define MIMADV_MOMENTUM 64u
/* Record an event pair around every launch of the call (eager launches
 * only; not with MIMADV_GRAPH); mimadv_batch_kernel_ms reads them. */
#define MIMADV_PROFILE 1024u

/* Tuning of the batched launches; 0 selects the default of each field. */
typedef struct mimadv_tuning {
    int group32;        /* FP32 fields per launch (default and maximum 18) */
    int kchunk;         /* levels per work item (default and maximum 8; 6 for FP32 WENO/MUSCL
                           launches with 256-thread tiles) */
    int blocks_per_cu;  /* persistent grid size per CU (default: occupancy) */
    int tile_y;         /* rows of the column tile in 3D: 2, 4 or 8 */
    int tile_x;         /* columns of the tile: 64, 32 with tile_y 8 (or 0), or 16 with
                           tile_y 0 (16x16).  Both 0: 32x8 tiles in 3D; tile_x 0 with tile_y
                           set: 64 columns; tile_x 64 with tile_y 0: 64x4 */
} mimadv_tuning;
int mimadv_set_tuning(mimadv_ctx *ctx, const mimadv_tuning *t);

/* Tendency xs_f = -div(q_f) of nf fields sharing one FP32 velocity, with
 * the context's scheme (MIMADV_WENO, MIMADV_WENO5 or MIMADV_MUSCL) in the batched kernel:
 * FP32 reconstruction for MIMADV_F32 fields, FP64 for MIMADV_F64 ones (at
 * most 2), flux divergence accumulated in FP64.  Device arrays; xs_f has the
 * precision of q_f.  Lateral boundaries periodic, vertical no-flux.  flags
 * may restrict the directions (MIMADV_DIR_*; none given = all three).
 * Replaces, for a whole batch, the per-field loops of advect /
 * qnadv_scal_{x,y,z} (advection.f90:84-553) with flux-form fields. */
int mimadv_batch_tendency(mimadv_ctx *ctx, const float *u, const float *v, const float *w, int nf,
                          const int *precision, const void *const *q, void *const *xs, unsigned flags);

/* Advance nf fields by nsteps Shu-Osher steps (MIMADV_INT_SSPRK3, or
 * MIMADV_INT_HEUN as its two-stage SSP form q1 = q + dt L(q),
 * q+ = q/2 + (q1 + dt L(q1))/2), every stage one batched launch per field
 * group with the periodic ghost refresh fused in.  flags may contain
 * MIMADV_GRAPH. */
int mimadv_batch_advance(mimadv_ctx *ctx, int integrator, int nsteps, double dt, const float *u, const float *v,
                         const float *w, int nf, const int *precision, void *const *q, unsigned flags);

/* Kernel time of the launches of the last batched call made with
 * MIMADV_PROFILE, summed per kind: [0] momenta, [1] FP64 fields, [2] FP32
 * fields (ms, from the events on the context stream), and the number of
 * launches of each kind.  Synchronises with the last launch. */
int mimadv_batch_kernel_ms(mimadv_ctx *ctx, double ms[3], int count[3]);

/* Domain budgets: dzw-weighted interior sums of nf fields, accumulated in
 * FP64 (out[f]). */
int mimadv_batch_budget(mimadv_ctx *ctx, int nf, const int *precision, const void *const *q, double *out);

/* mimadv_batch_budget for callers that cannot allocate device memory (code
 * under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc); it has no counterpart in the reference.  It allocates
 * nothing: the per-block partials go to a static device array of the
 * library (262 KB per GPU, held until the process ends), so nf is at most
 * 32 (MIMADV_ERR_ARG beyond), and the results go to the caller's DEVICE
 * array d_out[nf]; the call is ordered on the context stream and does not
 * wait.  Every call on one GPU shares that array: calls that overlap on
 * one GPU (two streams, or two host threads) race.  Same values as
 * mimadv_batch_budget, bit for bit. */
int mimadv_batch_budget_device(mimadv_ctx *ctx, int nf, const int *precision, const void *const *q,
                               double *d_out);

/* ---- many independent subdomains in one launch (no reference counterpart:
 * the reference runs one domain per MPI job) ----
 * Subdomain d is described by a context ctx[d] (its grid, periodic lateral
 * and 'nnf' vertical boundaries, scheme MIMADV_WENO, MIMADV_WENO5 or MIMADV_MUSCL, 3D),
 * its FP32 velocity (u[d], v[d], w[d]) and nf[d] FP32 fields q[d][0..nf-1]
 * (device arrays, at most 18 per subdomain).  Every stage of every
 * subdomain is one work list of one persistent launch; the tile shape
 * (tile_x * tile_y = 256 with tile_x in 16, 32, 64) and the level
 * chunk are chosen per subdomain.  The plan keeps two stage arrays per
 * field and runs on the stream of ctx[0]. */
typedef struct mimadv_multi mimadv_multi;
int mimadv_multi_create(int ndom, mimadv_ctx *const *ctx, const float *const *u, const float *const *v,
                        const float *const *w, const int *nf, float *const *const *q, mimadv_multi **out);
void mimadv_multi_destroy(mimadv_multi *m);
const char *mimadv_multi_last_error(const mimadv_multi *m);
/* Tile shape and level chunk per subdomain (arrays of ndom; kchunk 0 =
 * default 6, the maximum: a work item stages its levels' velocities in LDS).  Without a call, a cost model picks the shape that wastes the
 * fewest face evaluations. */
int mimadv_multi_set_tiles(mimadv_multi *m, const int *tile_x, const int *kchunk);
int mimadv_multi_get_tiles(const mimadv_multi *m, int *tile_x, int *tile_y, int *kchunk);
/* Autotuner: times every candidate (tile shape x level chunk in 4, 6)
 * of every subdomain with launches of one stage of that subdomain alone,
 * its work list repeated until the launch fills the GPU (so a candidate is
 * judged by its cost per work item at saturation), reps times each; keeps
 * the fastest and reuses a result for subdomains of equal size, field count
 * and scheme.  The fields are not modified. */
int mimadv_multi_autotune(mimadv_multi *m, int reps);
/* nsteps Shu-Osher steps (MIMADV_INT_SSPRK3 or MIMADV_INT_HEUN) of every
 * subdomain with its own dt[d]; one launch per stage for the whole batch. */
int mimadv_multi_advance(mimadv_multi *m, int integrator, int nsteps, const double *dt);
/* Time of the last mimadv_multi_advance between HIP events on its stream. */
int mimadv_multi_last_ms(mimadv_multi *m, double *ms);

/* ---- adjoint of the plain FP64 'weno' path (no reference counterpart: the
 * reference has no adjoint or tangent-linear model) ----
 * Scheme MIMADV_WENO in flux form with the default floor convention
 * (MIMADV_WENO_BOTTOM_X0_IS_X1), periodic lateral boundaries, 'nnf' floor
 * and lid.  Fields and winds carry periodic ghost images, as the forward
 * reads them; derivatives are taken with respect to interior values: the
 * east/north boundary faces u(nx-1), v(ny-1) are credited to their interior
 * images u(4), v(4).  One-sided derivatives at the limiter kinks: the
 * selected candidate of min(), zero on the clamped side of min(|r|,1), the
 * branch taken at the ratio-20 weight switch, the u >= 0 side at u = 0. */

/* Vector-Jacobian product of the tendency L(q; u, v, w) = -div(q u):
 * qbar = (dL/dq)^T lam on interior cells (ghost entries are not written);
 * ubar, vbar, wbar += (dL/d(u, v, w))^T lam.  Device arrays. */
int mimadv_adjoint_tendency(mimadv_ctx *ctx, const double *u, const double *v, const double *w,
                            const double *q, const double *lam, double *qbar, double *ubar, double *vbar,
                            double *wbar);

/* Adjoint of mimadv_advance(integrator, nsteps, dt, u, v, w) of one field
 * in flux form started from q0.  On entry qbar holds dJ/dq(T) on interior
 * cells; on return it holds dJ/dq0 (ghost entries zero).  ubar, vbar, wbar
 * are accumulated.  The trajectory is recomputed from q0 on the device: a
 * first sweep keeps the state of every `checkpoint`-th step, then each
 * segment of `checkpoint` steps is replayed with its stage states held in
 * device memory and swept backwards (checkpoint <= 0 or >= nsteps: one
 * sweep, every stage state held).  Device memory:
 * (3 + nsteps/checkpoint + stages*checkpoint) arrays of nx*ny*nz doubles. */
int mimadv_adjoint_advance(mimadv_ctx *ctx, int integrator, int nsteps, double dt, const double *u,
                           const double *v, const double *w, const double *q0, double *qbar, double *ubar,
                           double *vbar, double *wbar, int checkpoint);

/* The same adjoint split in two, the way a caller that runs the forward
 * anyway would use it: mimadv_advance_record advances q in place exactly as
 * mimadv_advance does for one flux-form field and copies the state at steps
 * 0, K, 2K, ... into ck[0..ceil(nsteps/K)-1] (K = checkpoint, <= 0: nsteps);
 * mimadv_adjoint_advance_recorded then needs no forward sweep of its own:
 * per segment it replays K steps from ck[g] into stage storage
 * (3 + stages*K device arrays) and sweeps them backwards. */
int mimadv_advance_record(mimadv_ctx *ctx, int integrator, int nsteps, double dt, const double *u,
                          const double *v, const double *w, double *q, int checkpoint, double *const *ck);
int mimadv_adjoint_advance_recorded(mimadv_ctx *ctx, int integrator, int nsteps, double dt, const double *u,
                                    const double *v, const double *w, int checkpoint, const double *const *ck,
                                    double *qbar, double *ubar, double *vbar, double *wbar);

/* Kernel time of the last mimadv_advance / mimadv_scalar_tendency call,
 * measured with HIP events on the context stream (ms). */
int mimadv_last_kernel_ms(mimadv_ctx *ctx, double *ms);

#ifdef __cplusplus
}
#endif
#endif
