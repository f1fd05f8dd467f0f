/*
 * mimadv_oracle.h -- CPU restatement of MIMICA's advection operator.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP library is
 * compared against (tests, the smoke check and the CPU-baseline leg of the
 * benchmark).  Nothing in the product path (fdd3260_amd/) links, loads or
 * calls it.
 *
 * It restates, in plain C with the reference's loop structure and operation
 * order, the Fortran routines of the reference model:
 *   src/TOOLS/advq_weno.f90   advs{x,y,z}_weno, weight{x,y,z}, limiter_weno
 *   src/TOOLS/advq_ppm.f90    advs{x,y,z}_ppm, limiter_ppm
 *   src/TOOLS/advq_muscl.f90  advs{x,y,z}_muscl, limiter_muscl
 *   src/TOOLS/advection.f90   advection_{x,y,z} (:556-798), advection_ac (:801-869)
 *   src/INTEG/windadv.f90     qnadv_uvw and adv{u,v,w}{x,y,z}_fd
 *   src/TOOLS/bc.f90          qnbc_p4 (:378-503), lbcu_p4 (:81-344)
 *   src/INTEG/integpack.f90 + INCLUDE/integ.h   Heun / Euler stage combine
 *   src/TOOLS/averages.f90    totalav (:507-551)
 *   src/TOOLS/funcpack.f90    get_momentum/get_velocities/get_advective/
 *                             get_conserved/get_scalars, anelastic (:65-542)
 * plus two schemes the reference does not have (WENO5-JS reconstruction and
 * SSP-RK3 stage coefficients), which are pinned only by closed forms.
 *
 * Pins: every routine is checked against closed forms, invariants or
 * hand-derived known answers in tests/test_oracle_pins.py (advsz_ppm on a
 * stretched grid through the Colella-Woodward interface formula on the grid
 * whose widths are dz1w, the metric the reference passes in).
 *
 * Arrays are Fortran column-major with x fastest: element (i,j,k), 1-based,
 * lives at [(i-1) + nx*((j-1) + ny*(k-1))].  nx = MAXX+5, ny = MAXY+5 (or 1
 * for a 2D run), nz = MAXZ; interior columns are i = 4..nx-2, j = 4..ny-2.
 */
#ifndef MIMADV_ORACLE_H
#define MIMADV_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_BC_PER = 0, ORC_BC_NNF = 1, ORC_BC_OPE = 2 };
enum { ORC_WENO = 0, ORC_PPM = 1, ORC_MUSCL = 2, ORC_WENO5 = 3 };
/* advq_quick.f90, advq_lw.f90 and their unlimited forms (advq_quick_nolim.f90,
 * advq_lw_nolim.f90) */
enum { ORC_QUICK = 4, ORC_LW = 5, ORC_QUICK_NL = 6, ORC_LW_NL = 7 };
enum { ORC_LIM_MINMOD = 0, ORC_LIM_MC = 1, ORC_LIM_VL = 2, ORC_LIM_FROMM = 3 };
/* weno_bottom: what advsz_weno reads for x(0) in its first loop pass. */
enum { ORC_WENO_BOTTOM_X0_IS_X1 = 0, ORC_WENO_BOTTOM_CLOSURE = 1 };

typedef struct {
    int nx, ny, nz;          /* full extents including ghosts */
    double dx, dy, dz;
    const double *fdz0;      /* [nz] cell size fraction */
    const double *dzw;       /* [nz] dz*fdz0 */
    const double *dz1w;      /* [nz] 1/dzw */
    const double *dz1;       /* [nz] 1/(centre distance k..k+1) */
    int bcl[6];              /* ORC_BC_* for west, east, south, north, bottom, top */
} orc_grid;

typedef struct {
    int scheme;              /* ORC_WENO .. ORC_LW_NL: the line kernel to run */
    int muscl_limiter;       /* ORC_LIM_* (reference default: minmod) */
    int ppm_limiter;         /* 1: ppmlimiter='PPM' (reference default), 0: none */
    double lim_tol;          /* reference default 1e-9 */
    double dt0;              /* full step, used by PPM Courant numbers */
    int weno_bottom;         /* ORC_WENO_BOTTOM_* */
    int mom_ord;             /* 2 or 4 */
    int lavisc;              /* artificial viscosity on momentum */
} orc_opts;

/* The line kernel advection_{x,y,z} runs for scal_adv = scheme (ORC_WENO,
 * ORC_PPM, ORC_MUSCL, ORC_QUICK, ORC_LW), limit = !unlimited and sorder. */
int orc_line_scheme(int scheme, int unlimited, int sorder);

/* ---- 1D line kernels (pointers to element 1 of the line, stride 1) ---- */
void orc_advs_h(const orc_opts *o, int n, double d, int bc_lo, int bc_hi,
                const double *x, const double *u, double *xa, int adv);
/* x0 is the value read for x(0) (reference out-of-bounds read, see
 * advq_weno.f90:245-247); ignored by PPM/MUSCL. */
void orc_advs_z(const orc_grid *g, const orc_opts *o, const double *x, double x0,
                const double *w, double *xa, int adv);

/* ---- large-scale vertical advection (advection.f90:918-1042) ---- */
/* grad_1d (gradients.f90:400-473) of one column xz[nz] into gz[nz] */
void orc_grad_1d(const orc_grid *g, const double *xz, double *gz);
/* xs -= dref*w0*grad_1d(xx/dref) on interior columns; dref from dref3 (3D),
 * else dref_prof (levels), else 1 */
void orc_advection_1d(const orc_grid *g, const double *w0, const double *dref_prof, const double *dref3,
                      const double *xx, double *xs);

/* ---- 3D operator ---- */
/* xs = xs - div_dir(q) for dir = 0 (x), 1 (y), 2 (z); like advection_{x,y,z}. */
void orc_advection_dir(const orc_grid *g, const orc_opts *o, int dir,
                       const double *q, const double *vel, double *xs, int adv);
/* xs = ((xs - X) - Y) - Z over the directions in dirmask (bit0 x, bit1 y,
 * bit2 z); y is skipped when ny == 1 (2D model). */
void orc_scalar_tendency(const orc_grid *g, const orc_opts *o, const double *q,
                         const double *u, const double *v, const double *w,
                         double *xs, int adv, int dirmask);

/* Positivity-preserving Shu-Osher stage of the flux-form 'weno' tendency
 * (no reference counterpart): out = c0*q0 + c1*(q + dt*L~(q)) with each
 * cell's outgoing fluxes scaled so that q + dt*L~(q) >= 0; periodic sides,
 * 'nnf' floor and lid.  Returns 0, or -1 if unsupported. */
int orc_positive_stage(const orc_grid *g, const orc_opts *o, const double *q, const double *u, const double *v,
                       const double *w, const double *q0, double c0, double c1, double dt, double *out);

/* qnadv_uvw: ua/va/wa are OVERWRITTEN with +div(rho0 u u) (windadv.f90:46-79).
 * mu/mv/mw: momentum (rho0 u); ut/vt/wt: advecting velocities. */
void orc_qnadv_uvw(const orc_grid *g, const orc_opts *o, int ntime,
                   const double *mu, const double *mv, const double *mw,
                   const double *ut, const double *vt, const double *wt,
                   double *ua, double *va, double *wa);

/* Ghost fill for one scalar (qnbc_p4, serial branch). */
void orc_fill_scalar(const orc_grid *g, double *x);
/* Ghost fill for winds: lbcu_p4 of a single-tile run (bc.f90:81-344). */
void orc_fill_winds(const orc_grid *g, double *u, double *v, double *w);
double orc_sponge_factor(int nz, int sorder, int im, double dt, const double *z0, double zmax, double zdamp,
                         int sponge);
void orc_integ_mom(const orc_grid *g, int sorder, int im, double dt, const double *z0, double zmax, double zdamp,
                   int sponge, double *const *old, const double *const *a, double *const *tend, double *const *x);

/* Stage combine of integ_scal for one field:
 *   sorder > 1:  tend = a + b_lm*tend ; x = xold + a_lm*dt*tend
 *   sorder == 1: x = xold + dt*a
 * with (a_lm, b_lm) from integ.h for stage im (1-based). n elements. */
void orc_integ_scal(long n, int sorder, int im, double dt, const double *xold,
                    const double *a, double *tend, double *x);

/* totalav: unweighted mean over interior columns, all levels. */
double orc_totalav(const orc_grid *g, const double *x);

/* Tangent-linear of the reference 'weno' tendency in flux form with
 * periodic sides (ghosts filled as periodic images) and 'nnf' floor/lid:
 * xs = -div q, dxs = its derivative along (dq, du, dv, dw); interior cells
 * only (x(0) := x(1) convention at the floor).  Kink conventions are listed
 * in mimadv_oracle.c. */
void orc_tlm_weno_tendency(const orc_grid *g, const double *q, const double *dq, const double *u,
                           const double *v, const double *w, const double *du, const double *dv,
                           const double *dw, double *xs, double *dxs);
/* The same for the WENO5-JS extension (x/y WENO5 faces, z WENO5 where the
 * upwind stencil fits the column, the reference face elsewhere). */
void orc_tlm_weno5_tendency(const orc_grid *g, const double *q, const double *dq, const double *u,
                            const double *v, const double *w, const double *du, const double *dv,
                            const double *dw, double *xs, double *dxs);

/* ---- density weighting of the anelastic build (funcpack.f90:65-542) ---- */
/* avden0(1) = den0(1); avden0(k) = (fdz0(k)den0(k) + fdz0(k-1)den0(k-1)) /
 * (fdz0(k) + fdz0(k-1)) (initial.f90:175, :543-546). */
void orc_avden0(int nz, const double *fdz0, const double *den0, double *avden0);
/* get_momentum (inverse = 0; funcpack.f90:99-178 anelastic branch):
 * u,v *= den0(k), w *= avden0(k); get_velocities (inverse = 1; :236-316):
 * the same divisions.  Whole arrays including ghosts; v skipped when ny == 1. */
void orc_get_momentum(const orc_grid *g, const double *den0, const double *avden0, double *u, double *v,
                      double *w, int inverse);
/* get_advective (funcpack.f90:65-92): ua = u/den0 - u0shift,
 * va = v/den0 - v0shift, wa = w/avden0. */
void orc_get_advective(const orc_grid *g, const double *den0, const double *avden0, double u0shift,
                       double v0shift, const double *u, const double *v, const double *w, double *ua,
                       double *va, double *wa);
/* get_conserved (inverse = 0; funcpack.f90:397-454): q *= dens;
 * get_scalars (inverse = 1; :457-514): q /= dens.  dens is a 3D array. */
void orc_get_conserved(const orc_grid *g, const double *dens, double *q, int inverse);

/* OpenMP thread count used by the 3D drivers (1 = serial).  Returns the
 * previous value. */
int orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
