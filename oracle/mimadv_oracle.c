/*
 * mimadv_oracle.c -- CPU restatement of MIMICA's advection operator.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never part of
 * the product path.  See mimadv_oracle.h for the list of reference routines
 * restated here.  Every routine keeps the reference's loop structure and
 * floating-point operation order (Fortran evaluates a*b*c as (a*b)*c and
 * a+b-c as (a+b)-c) so that, built without FMA contraction, it reproduces
 * the reference arithmetic rather than an algebraically equivalent variant.
 *
 * Indexing helpers use the reference's 1-based indices: L(x, i) is x(i).
 */
#include "mimadv_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define L(a, i) ((a)[(i) - 1])
#define IDX(g, i, j, k) ((size_t)((i) - 1) + (size_t)(g)->nx * ((size_t)((j) - 1) + (size_t)(g)->ny * (size_t)((k) - 1)))

static int g_threads = 1;

int orc_set_threads(int n)
{
    int old = g_threads;
    g_threads = n < 1 ? 1 : n;
    return old;
}

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }
/* Fortran SIGN(1., x) with the sign of a negative zero honoured. */
static double sgn1(double x) { return copysign(1.0, x); }

/* Fortran SPACING(x) for REAL(8): b**max(e-p, emin-1); TINY for x == 0. */
static double fspacing(double x)
{
    if (x == 0.0) return ldexp(1.0, -1022);
    int e = ilogb(x) - 52;
    if (e < -1022) e = -1022;
    return ldexp(1.0, e);
}

/* ======================================================================
 * Reference "WENO" (3rd-order switch variant): advq_weno.f90
 * ====================================================================== */

/* weight{x,y,z} body for one face (advq_weno.f90:329-350): ww = linear
 * weights (1/3, 2/3 | 2/3, 1/3) unless the smoothness ratio of a side
 * exceeds 20, then normalised g/(eps+eta)^2 weights. */
static void weno3_weights(double a, double b, double c, double d, double ww[4])
{
    const double g1 = 1.0 / 3.0, g2 = 2.0 / 3.0, eps = 1.e-15, thres = 20.0;
    double eta1 = (b - a) * (b - a);
    double eta2 = (c - b) * (c - b);
    double eta3 = (d - c) * (d - c);
    double s1 = (eps + eta1) * (eps + eta1);
    double s2 = (eps + eta2) * (eps + eta2);
    double s3 = (eps + eta3) * (eps + eta3);
    double w1 = g1 / s1, w2 = g2 / s2, w3 = g2 / s2, w4 = g1 / s3;
    ww[0] = g1; ww[1] = g2; ww[2] = g2; ww[3] = g1;
    if (dmax(eta1, eta2) / (dmin(eta1, eta2) + eps) > thres) {
        double sumw = w1 + w2;
        ww[0] = w1 / sumw;
        ww[1] = w2 / sumw;
    }
    if (dmax(eta2, eta3) / (dmin(eta2, eta3) + eps) > thres) {
        double sumw = w3 + w4;
        ww[2] = w3 / sumw;
        ww[3] = w4 / sumw;
    }
}

/* limiter_weno (advq_weno.f90:462-491): rescale the candidate face values
 * about the left/lower cell value x. */
static void weno3_limiter(double *f1, double *f2, double *f3, double *f4, double x)
{
    const double eps = 1.e-15;
    double minp = dmin(dmin(dmin(*f1, *f2), *f3), *f4);
    double phi = fmin(fabs((0.0 - x) / (minp - x + eps)), 1.0);
    *f1 = phi * (*f1 - x) + x;
    *f2 = phi * (*f2 - x) + x;
    *f3 = phi * (*f3 - x) + x;
    *f4 = phi * (*f4 - x) + x;
}

/* One interior face flux, x/y kernels (advq_weno.f90:92-110):
 * a..d = x(i-1..i+2), uu = u(i+1), ww from weno3_weights. */
static double weno3_face(double a, double b, double c, double d, double uu,
                         const double ww[4], int limit)
{
    double ffim1 = -(0.5 * a) + 1.5 * b;
    double ffip1p = 0.5 * b + 0.5 * c;
    double ffip1m = 0.5 * b + 0.5 * c;
    double ffip2 = 1.5 * c - 0.5 * d;
    if (limit) weno3_limiter(&ffim1, &ffip1m, &ffip1p, &ffip2, b);
    double fl = ww[0] * ffim1 + ww[1] * ffip1p;
    double fr = ww[2] * ffip1m + ww[3] * ffip2;
    return 0.5 * fl * (uu + fabs(uu)) + 0.5 * fr * (uu - fabs(uu));
}

/* The face fluxes of advsx_weno / advsy_weno (advq_weno.f90:92-110):
 * flux(i) at face i+1/2 for i = 3..n-2. */
static void weno3_h_flux(int n, const double *x, const double *u, double *flux)
{
    for (int i = 3; i <= n - 2; ++i) {
        double ww[4];
        weno3_weights(L(x, i - 1), L(x, i), L(x, i + 1), L(x, i + 2), ww);
        L(flux, i) = weno3_face(L(x, i - 1), L(x, i), L(x, i + 1), L(x, i + 2), L(u, i + 1), ww, 1);
    }
}

/* advsx_weno / advsy_weno with i1 = 4, i2 = n-2 (advq_weno.f90:61-131). */
static void weno3_h(int n, double d, const double *x, const double *u, double *xa, int adv)
{
    int ii1 = 3, ii2 = n - 2;
    double d1 = 1.0 / d;
    double *flux = (double *)calloc((size_t)n + 1, sizeof(double));
    weno3_h_flux(n, x, u, flux);
    double fold = L(flux, ii1);
    for (int i = ii1 + 1; i <= ii2; ++i) {
        L(xa, i) = L(xa, i) + (L(flux, i) - fold) * d1;
        fold = L(flux, i);
    }
    if (adv) {
        double uold = L(u, ii2 + 1);
        for (int i = ii2; i >= ii1 + 1; --i) {
            L(xa, i) = L(xa, i) - L(x, i) * (uold - L(u, i)) * d1;
            uold = L(u, i);
        }
    }
    free(flux);
}

static double weno5_left(double a, double b, double c, double d, double e);

/* advsz_weno (advq_weno.f90:202-301) with k1 = 1, k2 = nz.  weno5 != 0 is
 * the WENO5-JS extension in z (no reference counterpart): a face of the
 * main loop whose 5-cell upwind stencil lies inside the column (w > 0:
 * 3 <= k <= nz-2, w < 0: 2 <= k <= nz-3) takes the Jiang-Shu face with the
 * uniform-grid coefficients, as the reference scheme also uses in z; every
 * other face, the closures included, stays the reference face. */
/* The face fluxes of advsz_weno: flux(k) at face k+1/2 for k = 1..nz-1
 * (the closures and the main loop of weno3_z below). */
static void weno3_z_flux(const orc_grid *g, const orc_opts *o, const double *x, double x0,
                         const double *w, double *flux, int weno5)
{
    int nz = g->nz;
    const double g1 = 1.0 / 3.0, g2 = 2.0 / 3.0;
    double lin[4] = {g1, g2, g2, g1};
#define XZ(k) ((k) == 0 ? x0 : L(x, k))
    /* bottom closure, k = 1 (:232-241): no limiter, linear right weights */
    {
        int k = 1;
        double www = L(w, k + 1);
        double ffip1p = 0.5 * XZ(k) + 0.5 * XZ(k + 1);
        double ffip1m = 0.5 * XZ(k) + 0.5 * XZ(k + 1);
        double ffip2 = 1.5 * XZ(k + 1) - 0.5 * XZ(k + 2);
        double fl = ffip1p;
        double fr = lin[2] * ffip1m + lin[3] * ffip2;
        L(flux, k) = 0.5 * fl * (www + fabs(www)) + 0.5 * fr * (www - fabs(www));
    }
    /* main loop k = k1..k2-2 (:245-258); k = 1 reads x(0) (F4) unless the
     * closure convention is selected.  weightz(2, nz-2) sets nonlinear
     * weights only for k = 3..nz-3 (:435). */
    int kstart = (o->weno_bottom == ORC_WENO_BOTTOM_CLOSURE) ? 2 : 1;
    for (int k = kstart; k <= nz - 2; ++k) {
        double wk = L(w, k + 1);
        if (weno5 && wk > 0.0 && k >= 3 && k <= nz - 2) {
            double fl = weno5_left(XZ(k - 2), XZ(k - 1), XZ(k), XZ(k + 1), XZ(k + 2));
            L(flux, k) = 0.5 * fl * (wk + fabs(wk));
            continue;
        }
        if (weno5 && wk < 0.0 && k >= 2 && k <= nz - 3) {
            double fr = weno5_left(XZ(k + 3), XZ(k + 2), XZ(k + 1), XZ(k), XZ(k - 1));
            L(flux, k) = 0.5 * fr * (wk - fabs(wk));
            continue;
        }
        double ww[4] = {g1, g2, g2, g1};
        if (k >= 3 && k <= nz - 3) weno3_weights(XZ(k - 1), XZ(k), XZ(k + 1), XZ(k + 2), ww);
        L(flux, k) = weno3_face(XZ(k - 1), XZ(k), XZ(k + 1), XZ(k + 2), wk, ww, 1);
    }
    /* top closure, k = nz-1 (:262-271): no limiter, linear left weights */
    {
        int k = nz - 1;
        double www = L(w, k + 1);
        double ffim1 = -(0.5 * XZ(k - 1)) + 1.5 * XZ(k);
        double ffip1p = 0.5 * XZ(k) + 0.5 * XZ(k + 1);
        double ffip1m = 0.5 * XZ(k) + 0.5 * XZ(k + 1);
        double fl = lin[0] * ffim1 + lin[1] * ffip1p;
        double fr = ffip1m;
        L(flux, k) = 0.5 * fl * (www + fabs(www)) + 0.5 * fr * (www - fabs(www));
    }
#undef XZ
}

static void weno3_z(const orc_grid *g, const orc_opts *o, const double *x, double x0,
                    const double *w, double *xa, int adv, int weno5)
{
    int nz = g->nz;
    double *flux = (double *)calloc((size_t)nz + 1, sizeof(double));
    weno3_z_flux(g, o, x, x0, w, flux, weno5);
    double fold = L(flux, 1);
    for (int k = 2; k <= nz - 1; ++k) {
        L(xa, k) = L(xa, k) + (L(flux, k) - fold) * L(g->dz1w, k);
        fold = L(flux, k);
    }
    if (adv) {
        double wold = L(w, nz);
        for (int k = nz - 1; k >= 2; --k) {
            L(xa, k) = L(xa, k) - L(x, k) * (wold - L(w, k)) * L(g->dz1w, k);
            wold = L(w, k);
        }
    }
    if (g->bcl[4] == ORC_BC_NNF) L(xa, 1) = L(xa, 1) + (L(flux, 1) - 0.0) * L(g->dz1w, 1);
    if (adv && g->bcl[4] == ORC_BC_NNF) L(xa, 1) = L(xa, 1) - L(x, 1) * (L(w, 2) - 0.0) * L(g->dz1w, 1);
    if (g->bcl[5] == ORC_BC_NNF) L(xa, nz) = L(xa, nz) + (0.0 - L(flux, nz - 1)) * L(g->dz1w, nz);
    if (g->bcl[5] == ORC_BC_NNF && adv) L(xa, nz) = L(xa, nz) - L(x, nz) * (0.0 - L(w, nz)) * L(g->dz1w, nz);
    free(flux);
}

/* ======================================================================
 * PPM: advq_ppm.f90
 * ====================================================================== */

typedef struct { double phir, phil, phi1, phi2; } ppm_lim_t;

/* limiter_ppm (advq_ppm.f90:449-530). */
static ppm_lim_t ppm_limiter(int use_lim, double xx, double xp, double xm, double xplus,
                             double xmin, double ffip1, double ffip2, double ffim1, double ffim2)
{
    const double eps = 1.e-15, thres = 10.0;
    ppm_lim_t r;
    double rp1 = 1.0;
    r.phir = xplus;
    r.phil = xmin;
    double g1a = ffip1 - ffim1, g1b = ffip1 + ffim1;
    double g2a = ffip2 - ffip1, g2b = ffip2 + ffip1;
    double g3a = ffim1 - ffim2, g3b = ffim1 + ffim2;
    double gamma1 = 0.5 * (g1a * g1a + g1b * g1b);
    double gamma2 = 0.5 * (g2a * g2a + g2b * g2b);
    double gamma3 = 0.5 * (g3a * g3a + g3b * g3b);
    double lambda = dmax(dmax(gamma1, gamma2), gamma3) / (dmin(dmin(gamma1, gamma2), gamma3) + eps);
    r.phi1 = r.phir - r.phil;
    r.phi2 = 3.0 * (r.phil - 2.0 * xx + r.phir);
    if (use_lim && lambda > thres) {
        if (r.phir < 0.0 || ((r.phir - xx) * (xp - r.phir) < 0.0 &&
                             (ffim1 * ffip2 >= 0.0 || ffim1 * ffip1 <= 0.0 || ffip2 * ffip1 <= 0.0))) {
            if (fabs(r.phir - xx) < fabs(r.phir - xp)) r.phir = xx;
            else r.phir = xp;
        }
        if (r.phil < 0.0 || ((xx - r.phil) * (r.phil - xm) < 0.0 &&
                             (ffim2 * ffip1 >= 0.0 || ffim2 * ffim1 <= 0.0 || ffip1 * ffim1 <= 0.0))) {
            if (fabs(r.phil - xx) < fabs(r.phil - xm)) r.phil = xx;
            else r.phil = xm;
        }
        r.phi1 = r.phir - r.phil;
        r.phi2 = 3.0 * (r.phil - 2.0 * xx + r.phir);
        double slope = -0.5 * r.phi1 / (sgn1(r.phi2) * dmax(fabs(r.phi2), eps));
        if (fabs(ffim1) > sqrt(fspacing(ffim1))) rp1 = ffip1 / ffim1;
        /* (slope > -1/2 .or. slope < 1/2) is true for every non-NaN slope (F7) */
        if (fabs(r.phi2) > eps && (slope > -0.5 || slope < 0.5)) {
            double lim = dmax(0.0, dmin(dmin(2.0, 0.5 * (rp1 + 1.0)), 2.0 * rp1));
            r.phir = xx + 0.5 * ffim1 * lim;
            r.phil = xx - 0.5 * ffim1 * lim;
            if ((xx - r.phil) * (r.phir - xx) < 0.0) {
                r.phil = xx;
                r.phir = xx;
            }
            r.phi1 = r.phir - r.phil;
            r.phi2 = 0.0;
        }
    }
    return r;
}

static double ppm_fr(const ppm_lim_t *l, double mup)
{
    return l->phir - 0.5 * mup * (l->phi1 + (1.0 - (2.0 / 3.0) * mup) * l->phi2);
}
static double ppm_fl(const ppm_lim_t *l, double muc)
{
    return l->phil + 0.5 * muc * (l->phi1 - (1.0 - (2.0 / 3.0) * muc) * l->phi2);
}

/* advsx_ppm / advsy_ppm with i1 = 4, i2 = n-2 (advq_ppm.f90:55-169). */
static void ppm_h(const orc_opts *o, int n, double d, int bc_hi, const double *x,
                  const double *u, double *xa, int adv)
{
    int ii1 = 3, ii2 = n - 2;
    double d1 = 1.0 / d, dt0 = o->dt0;
    double *fr = (double *)calloc((size_t)n + 2, sizeof(double));
    double *fl = (double *)calloc((size_t)n + 2, sizeof(double));
    double mup, muc, ffip1, ffip2, ffim1, ffim2, xmin, xplus;
    ppm_lim_t l;
    /* left boundary cell ii1 (:81-94) */
    mup = fabs(L(u, ii1 + 1)) * dt0 * d1;
    muc = fabs(L(u, ii1)) * dt0 * d1;
    ffip1 = L(x, ii1 + 1) - L(x, ii1);
    ffip2 = L(x, ii1 + 2) - L(x, ii1 + 1);
    ffim1 = L(x, ii1) - L(x, ii1 - 1);
    ffim2 = L(x, ii1) - L(x, ii1 - 1);
    xmin = 0.5 * (L(x, ii1 + 1) + L(x, ii1));
    xplus = L(x, ii1) + 0.5 * ffip1;
    l = ppm_limiter(o->ppm_limiter, L(x, ii1), L(x, ii1 + 1), L(x, ii1 - 1), xplus, xmin, ffip1, ffip2, ffim1, ffim2);
    L(fr, ii1) = ppm_fr(&l, mup);
    L(fl, ii1) = ppm_fl(&l, muc);
    /* core (:98-118) */
    xmin = xplus;
    for (int i = ii1 + 1; i <= ii2; ++i) {
        mup = fabs(L(u, i + 1)) * dt0 * d1;
        muc = fabs(L(u, i)) * dt0 * d1;
        ffip1 = L(x, i + 1) - L(x, i);
        ffip2 = L(x, i + 2) - L(x, i + 1);
        ffim1 = L(x, i) - L(x, i - 1);
        ffim2 = L(x, i - 1) - L(x, i - 2);
        double sl = 0.5 * (ffip1 + ffim1);
        double slp1 = 0.5 * (ffip2 + ffip1);
        xplus = L(x, i) + 0.5 * ffip1 - (1.0 / 6.0) * (slp1 - sl);
        l = ppm_limiter(o->ppm_limiter, L(x, i), L(x, i + 1), L(x, i - 1), xplus, xmin, ffip1, ffip2, ffim1, ffim2);
        L(fr, i) = ppm_fr(&l, mup);
        L(fl, i) = ppm_fl(&l, muc);
        xmin = xplus;
    }
    /* right boundary cell ii2+1 (:122-134) */
    mup = fabs(L(u, ii2 + 2)) * dt0 * d1;
    muc = fabs(L(u, ii2 + 1)) * dt0 * d1;
    ffip1 = L(x, ii2 + 2) - L(x, ii2 + 1);
    ffip2 = L(x, ii2 + 2) - L(x, ii2 + 1);
    ffim1 = L(x, ii2 + 1) - L(x, ii2);
    ffim2 = L(x, ii2) - L(x, ii2 - 1);
    xplus = L(x, ii2 + 1) + 0.5 * ffip1;
    l = ppm_limiter(o->ppm_limiter, L(x, ii2 + 1), L(x, ii2 + 2), L(x, ii2), xplus, xmin, ffip1, ffip2, ffim1, ffim2);
    L(fr, ii2 + 1) = ppm_fr(&l, mup);
    L(fl, ii2 + 1) = ppm_fl(&l, muc);
    /* non-periodic fix-ups (:138-146); the left one needs ii1 == 2 and can
     * never fire with i1 = 4, so only the right one is restated. */
    if (bc_hi != ORC_BC_PER) {
        L(fr, ii2) = L(fl, ii2);
        L(fl, ii2 + 1) = L(fr, ii2 - 1);
    }
    for (int i = ii1 + 1; i <= ii2; ++i) {
        double ui = L(u, i), up = L(u, i + 1);
        double flux_old = 0.5 * L(fr, i - 1) * (ui + fabs(ui)) + 0.5 * L(fl, i) * (ui - fabs(ui));
        double flux = 0.5 * L(fr, i) * (up + fabs(up)) + 0.5 * L(fl, i + 1) * (up - fabs(up));
        L(xa, i) = L(xa, i) + (flux - flux_old) * d1;
    }
    if (adv) {
        double uold = L(u, ii2 + 1);
        for (int i = ii2; i >= ii1 + 1; --i) {
            L(xa, i) = L(xa, i) - L(x, i) * (uold - L(u, i)) * d1;
            uold = L(u, i);
        }
    }
    free(fr);
    free(fl);
}

/* advsz_ppm (advq_ppm.f90:289-446) with k1 = 1, k2 = nz. */
static void ppm_z(const orc_grid *g, const orc_opts *o, const double *x, const double *w,
                  double *xa, int adv)
{
    int nz = g->nz;
    double dt0 = o->dt0;
    const double *dz1 = g->dz1, *dz1w = g->dz1w;
    double bc51 = (g->bcl[4] == ORC_BC_PER) ? 1.0 : 0.0;
    double bc61 = (g->bcl[5] == ORC_BC_PER) ? 1.0 : 0.0;
    double *fr = (double *)calloc((size_t)nz + 1, sizeof(double));
    double *fl = (double *)calloc((size_t)nz + 1, sizeof(double));
    double mup, muc, ffip1, ffip2, ffim1, ffim2, xmin, xplus;
    ppm_lim_t l;
    int k;
    /* bottom, k = 1 (:328-343) */
    k = 1;
    mup = fabs(L(w, k + 1)) * dt0 * L(dz1, k);
    muc = fabs(L(w, k)) * dt0 * L(dz1, k);
    ffip1 = L(x, k + 1) - L(x, k);
    ffip2 = L(x, k + 2) - L(x, k + 1);
    ffim1 = bc51 * (L(x, nz - 4) - L(x, nz - 5));
    ffim2 = bc51 * (L(x, nz - 5) - L(x, nz - 6));
    xmin = L(x, k) - 0.5 * ffip1;
    xplus = L(x, k) + 0.5 * ffip1;
    l = ppm_limiter(o->ppm_limiter, L(x, k), L(x, k + 1), L(x, k + 1), xplus, xmin, ffip1, ffip2, ffim1, ffim2);
    L(fr, k) = ppm_fr(&l, mup);
    L(fl, k) = ppm_fl(&l, muc);
    /* k = 2 (:345-360) */
    k = 2;
    xmin = xplus;
    mup = fabs(L(w, k + 1)) * dt0 * L(dz1, k);
    muc = fabs(L(w, k)) * dt0 * L(dz1, k);
    ffip1 = L(x, k + 1) - L(x, k);
    ffip2 = L(x, k + 2) - L(x, k + 1);
    ffim1 = L(x, k) - L(x, k - 1);
    ffim2 = bc51 * (L(x, nz - 4) - L(x, nz - 5));
    xplus = L(x, k) + 0.5 * ffip1;
    l = ppm_limiter(o->ppm_limiter, L(x, k), L(x, k + 1), L(x, k - 1), xplus, xmin, ffip1, ffip2, ffim1, ffim2);
    L(fr, k) = ppm_fr(&l, mup);
    L(fl, k) = ppm_fl(&l, muc);
    /* core k = 3..nz-2 with the non-uniform face formula (:364-391) */
    xmin = xplus;
    for (k = 3; k <= nz - 2; ++k) {
        mup = fabs(L(w, k + 1)) * dt0 * L(dz1, k);
        muc = fabs(L(w, k)) * dt0 * L(dz1, k);
        ffip1 = L(x, k + 1) - L(x, k);
        ffip2 = L(x, k + 2) - L(x, k + 1);
        ffim1 = L(x, k) - L(x, k - 1);
        ffim2 = L(x, k - 1) - L(x, k - 2);
        double delta = L(dz1w, k), deltam1 = L(dz1w, k - 1), deltap1 = L(dz1w, k + 1), deltap2 = L(dz1w, k + 2);
        double sl = delta / (deltam1 + delta + deltap1) *
                    ((2.0 * deltam1 + delta) / (deltap1 + delta) * ffip1 + (delta + 2.0 * deltap1) / (deltam1 + delta) * ffim1);
        double slp1 = deltap1 / (delta + deltap1 + deltap2) *
                      ((2.0 * delta + deltap1) / (deltap2 + deltap1) * ffip2 + (deltap1 + 2.0 * deltap2) / (delta + deltap1) * ffip1);
        xplus = L(x, k) + delta / (delta + deltap1) * ffip1 +
                1.0 / (deltam1 + delta + deltap1 + deltap2) *
                    (2.0 * deltap1 * delta / (delta + deltap1) *
                         ((deltam1 + delta) / (2.0 * delta + deltap1) - (deltap2 + deltap1) / (2.0 * deltap1 + delta)) * ffip1 -
                     delta * (deltam1 + delta) / (2.0 * delta + deltap1) * slp1 +
                     deltap1 * (deltap1 + deltap2) / (2.0 * deltap1 + delta) * sl);
        l = ppm_limiter(o->ppm_limiter, L(x, k), L(x, k + 1), L(x, k - 1), xplus, xmin, ffip1, ffip2, ffim1, ffim2);
        L(fr, k) = ppm_fr(&l, mup);
        L(fl, k) = ppm_fl(&l, muc);
        xmin = xplus;
    }
    /* top, k = nz-1 (:395-409) */
    k = nz - 1;
    mup = fabs(L(w, k + 1)) * dt0 * L(dz1, k);
    muc = fabs(L(w, k)) * dt0 * L(dz1, k);
    ffip1 = L(x, k + 1) - L(x, k);
    ffim1 = L(x, k) - L(x, k - 1);
    ffim2 = L(x, k - 1) - L(x, k - 2);
    ffip2 = bc61 * (L(x, 6) - L(x, 5)) + (1.0 - bc61) * ffip1;
    xplus = L(x, k) + 0.5 * ffip1;
    l = ppm_limiter(o->ppm_limiter, L(x, k), L(x, k + 1), L(x, k - 1), xplus, xmin, ffip1, ffip2, ffim1, ffim2);
    L(fr, k) = ppm_fr(&l, mup);
    L(fl, k) = ppm_fl(&l, muc);
    /* k = nz (:411-414) */
    k = nz;
    muc = fabs(L(w, k)) * dt0 * L(dz1, k);
    ffim1 = L(x, k) - L(x, k - 1);
    L(fl, k) = L(x, k) + 0.5 * ffim1 * (1.0 - muc);
    /* divergence (:418-440) */
    double w2 = L(w, 2);
    double flux_old = 0.5 * L(fr, 1) * (w2 + fabs(w2)) + 0.5 * L(fl, 2) * (w2 - fabs(w2));
    for (k = 2; k <= nz - 1; ++k) {
        double wp = L(w, k + 1);
        double flux = 0.5 * L(fr, k) * (wp + fabs(wp)) + 0.5 * L(fl, k + 1) * (wp - fabs(wp));
        L(xa, k) = L(xa, k) + (flux - flux_old) * L(dz1w, k);
        flux_old = flux;
    }
    if (adv) {
        double wold = L(w, nz);
        for (k = nz - 1; k >= 2; --k) {
            L(xa, k) = L(xa, k) - L(x, k) * (wold - L(w, k)) * L(dz1w, k);
            wold = L(w, k);
        }
    }
    if (g->bcl[4] == ORC_BC_NNF)
        L(xa, 1) = L(xa, 1) + (0.5 * L(fr, 1) * (w2 + fabs(w2)) + 0.5 * L(fl, 2) * (w2 - fabs(w2)) - 0.0) * L(dz1w, 1);
    if (g->bcl[4] == ORC_BC_NNF && adv) L(xa, 1) = L(xa, 1) - L(x, 1) * (L(w, 2) - 0.0) * L(dz1w, 1);
    if (g->bcl[5] == ORC_BC_NNF) L(xa, nz) = L(xa, nz) + (0.0 - flux_old) * L(dz1w, nz);
    if (g->bcl[5] == ORC_BC_NNF && adv) L(xa, nz) = L(xa, nz) - L(x, nz) * (0.0 - L(w, nz)) * L(dz1w, nz);
    free(fr);
    free(fl);
}

/* ======================================================================
 * MUSCL: advq_muscl.f90
 * ====================================================================== */

/* limiter_muscl (advq_muscl.f90:335-383) */
static void muscl_limiter(const orc_opts *o, double xx, double ffip1, double ffip2, double ffim1,
                          double *phir, double *phil)
{
    double lt = o->lim_tol;
    double rp1 = 1.0, rp2 = 1.0;
    if (fabs(ffim1) > lt * fabs(xx) && fabs(ffip1) < fabs(ffim1) / lt) rp1 = ffip1 / ffim1;
    if (fabs(ffip2) > lt * fabs(xx) && fabs(ffip1) < fabs(ffip2) / lt) rp2 = ffip1 / ffip2;
    switch (o->muscl_limiter) {
    case ORC_LIM_VL:
        *phil = (rp1 + fabs(rp1)) / (1.0 + fabs(rp1));
        *phir = (rp2 + fabs(rp2)) / (1.0 + fabs(rp2));
        break;
    case ORC_LIM_MC:
        *phil = dmax(0.0, dmin(dmin(2.0 * rp1, 0.5 * (rp1 + 1.0)), 2.0));
        *phir = dmax(0.0, dmin(dmin(2.0 * rp2, 0.5 * (rp2 + 1.0)), 2.0));
        break;
    case ORC_LIM_MINMOD:
        *phil = dmax(0.0, dmin(1.0, rp1));
        *phir = dmax(0.0, dmin(1.0, rp2));
        break;
    default:
        *phil = 0.5 * (rp1 + 1.0);
        *phir = 0.5 * (rp2 + 1.0);
        break;
    }
}

static double muscl_upwind(double uu, double fl, double fr)
{
    return 0.5 * uu * (fl * (1.0 + sgn1(uu)) + fr * (1.0 - sgn1(uu)));
}

/* advsx_muscl / advsy_muscl with i1 = 4, i2 = n-2 (advq_muscl.f90:56-122). */
static void muscl_h(const orc_opts *o, int n, double d, const double *x, const double *u,
                    double *xa, int adv)
{
    const double eps = 1.e-18;
    int i1 = 4, i2 = n - 2;
    double d1 = 1.0 / d, lt = o->lim_tol;
    double *flux = (double *)calloc((size_t)n + 1, sizeof(double));
    for (int i = i1 - 1; i <= i2; ++i) {
        double uu = L(u, i + 1), xx = L(x, i);
        double slope_p = (L(x, i + 1) - xx) * d1;
        double slope_p2 = (L(x, i + 2) - L(x, i + 1)) * d1;
        double slope_m = (xx - L(x, i - 1)) * d1;
        double fl, fr;
        if (d * fabs(slope_m) > lt * fabs(xx) + eps || d * fabs(slope_p2) > lt * fabs(xx) + eps) {
            double phir, phil;
            muscl_limiter(o, xx, slope_p, slope_p2, slope_m, &phir, &phil);
            fr = L(x, i + 1) - 0.5 * phir * slope_p2 * d;
            fl = xx + 0.5 * phil * slope_m * d;
        } else {
            fr = L(x, i + 1) - 0.5 * slope_p * d;
            fl = xx + 0.5 * slope_p * d;
        }
        L(flux, i) = muscl_upwind(uu, fl, fr);
    }
    double fold = L(flux, i1 - 1);
    for (int i = i1; i <= i2; ++i) {
        L(xa, i) = L(xa, i) + (L(flux, i) - fold) * d1;
        fold = L(flux, i);
    }
    if (adv) {
        double uold = L(u, i2 + 1);
        for (int i = i2; i >= i1; --i) {
            L(xa, i) = L(xa, i) - L(x, i) * (uold - L(u, i)) * d1;
            uold = L(u, i);
        }
    }
    free(flux);
}

/* advsz_muscl (advq_muscl.f90:208-332). */
static void muscl_z(const orc_grid *g, const orc_opts *o, const double *x, const double *w,
                    double *xa, int adv)
{
    const double eps = 1.e-18;
    int nz = g->nz;
    double lt = o->lim_tol, dz = g->dz;
    const double *dz1 = g->dz1, *dz1w = g->dz1w, *fdz0 = g->fdz0;
    double bc51 = (g->bcl[4] == ORC_BC_PER) ? 1.0 : 0.0;
    double bc61 = (g->bcl[5] == ORC_BC_PER) ? 1.0 : 0.0;
    double *flux = (double *)calloc((size_t)nz + 1, sizeof(double));
    double ww, slope_p, slope_p2, slope_m, phir, phil, fl, fr;
    int k;
    /* k = 1 (:245-261) */
    k = 1;
    ww = L(w, 2);
    slope_p = (L(x, 2) - L(x, 1)) * L(dz1, 1);
    slope_p2 = (L(x, 3) - L(x, 2)) * L(dz1, 2);
    slope_m = bc51 * (L(x, nz - 4) - L(x, nz - 5)) * L(dz1, nz - 5) + (1.0 - bc51) * slope_p;
    if (L(fdz0, 2) * dz * fabs(slope_p2) > lt * fabs(L(x, 1)) + eps) {
        muscl_limiter(o, L(x, 1), slope_p, slope_p2, slope_m, &phir, &phil);
        fl = L(x, 1) + 0.5 * phil * slope_m / L(dz1w, 1);
        fr = L(x, 2) - 0.5 * phir * slope_p2 / L(dz1w, 1);
    } else {
        fl = L(x, 1) + 0.5 * slope_p / L(dz1w, 1);
        fr = L(x, 2) - 0.5 * slope_p / L(dz1w, 1);
    }
    L(flux, 1) = muscl_upwind(ww, fl, fr);
    /* core k = 2..nz-2 (:265-282) */
    for (k = 2; k <= nz - 2; ++k) {
        ww = L(w, k + 1);
        slope_p = (L(x, k + 1) - L(x, k)) * L(dz1, k);
        slope_p2 = (L(x, k + 2) - L(x, k + 1)) * L(dz1, k + 1);
        slope_m = (L(x, k) - L(x, k - 1)) * L(dz1, k - 1);
        if (L(fdz0, k - 1) * dz * fabs(slope_m) > lt * fabs(L(x, k)) + eps ||
            L(fdz0, k + 1) * dz * fabs(slope_p2) > lt * fabs(L(x, k)) + eps) {
            muscl_limiter(o, L(x, k), slope_p, slope_p2, slope_m, &phir, &phil);
            fr = L(x, k + 1) - 0.5 * phir * slope_p2 / L(dz1w, k);
            fl = L(x, k) + 0.5 * phil * slope_m / L(dz1w, k);
        } else {
            fr = L(x, k + 1) - 0.5 * slope_p / L(dz1w, k);
            fl = L(x, k) + 0.5 * slope_p / L(dz1w, k);
        }
        L(flux, k) = muscl_upwind(ww, fl, fr);
    }
    /* top k = nz-1 (:286-302) */
    k = nz - 1;
    ww = L(w, k + 1);
    slope_p = (L(x, k + 1) - L(x, k)) * L(dz1, k);
    slope_p2 = bc61 * (L(x, 6) - L(x, 5)) * L(dz1, 5) + (1.0 - bc61) * slope_p;
    slope_m = (L(x, k) - L(x, k - 1)) * L(dz1, k - 1);
    if (L(fdz0, k - 1) * dz * fabs(slope_m) > lt * fabs(L(x, nz - 1)) + eps) {
        muscl_limiter(o, L(x, k), slope_p, slope_p2, slope_m, &phir, &phil);
        fr = L(x, k + 1) - 0.5 * phir * slope_p2 / L(dz1w, k);
        fl = L(x, k) + 0.5 * phil * slope_m / L(dz1w, k);
    } else {
        fr = L(x, k + 1) - 0.5 * slope_p / L(dz1w, k);
        fl = L(x, k) + 0.5 * slope_p / L(dz1w, k);
    }
    L(flux, k) = muscl_upwind(ww, fl, fr);
    /* divergence (:306-326) */
    double fold = L(flux, 1);
    for (k = 2; k <= nz - 1; ++k) {
        L(xa, k) = L(xa, k) + (L(flux, k) - fold) * L(dz1w, k);
        fold = L(flux, k);
    }
    if (adv) {
        double wold = L(w, nz);
        for (k = nz - 1; k >= 2; --k) {
            L(xa, k) = L(xa, k) - L(x, k) * (wold - L(w, k)) * L(dz1w, k);
            wold = L(w, k);
        }
    }
    if (g->bcl[4] == ORC_BC_NNF) L(xa, 1) = L(xa, 1) + (L(flux, 1) - 0.0) * L(dz1w, 1);
    if (g->bcl[4] == ORC_BC_NNF && adv) L(xa, 1) = L(xa, 1) - L(x, 1) * (L(w, 2) - 0.0) * L(dz1w, 1);
    if (g->bcl[5] == ORC_BC_NNF) L(xa, nz) = L(xa, nz) + (0.0 - L(flux, nz - 1)) * L(dz1w, nz);
    if (g->bcl[5] == ORC_BC_NNF && adv) L(xa, nz) = L(xa, nz) - L(x, nz) * (0.0 - L(w, nz)) * L(dz1w, nz);
    free(flux);
}

/* ======================================================================
 * QUICK: advq_quick.f90 (limited) and advq_quick_nolim.f90 (unlimited)
 * ====================================================================== */

/* limiter_quick (advq_quick.f90:313-352); phic is the value the guards
 * compare with (x in x/y, x/dz in z) */
static void quick_limiter(double lt, double phic, double ffip1, double ffip2, double ffim1, double *phil,
                          double *phir)
{
    const double eps = 1.e-18, c1 = 1.0 / 3.0;
    double rp1 = 1.0, rp2 = 1.0, rp11 = 1.0, rp21 = 1.0;
    if (fabs(ffim1) > lt * fabs(phic) + eps) rp1 = ffip1 / ffim1;
    if (fabs(ffip2) > lt * fabs(phic) + eps) rp2 = ffip1 / ffip2;
    if (fabs(ffip1) > lt * fabs(phic) + eps) {
        rp11 = ffim1 / ffip1;
        rp21 = ffip2 / ffip1;
    }
    *phil = rp1 * (1.0 - c1 * (1.0 - rp11));
    *phir = rp2 * (1.0 - c1 * (1.0 - rp21));
    if (fabs(ffim1) > lt * fabs(phic) + eps || fabs(ffip2) > lt * fabs(phic) + eps) {
        *phil = dmax(0.0, dmin(2.0 * rp1, dmin(*phil, 2.0)));
        *phir = dmax(0.0, dmin(2.0 * rp2, dmin(*phir, 2.0)));
    }
}

static double half_upwind(double uu, double fl, double fr)
{
    return 0.5 * fl * (uu + fabs(uu)) + 0.5 * fr * (uu - fabs(uu));
}

/* divergence of line fluxes flux(i1-1..i2) into xa(i1..i2), plus the
 * advective-form correction (the common tail of the x/y line kernels) */
static void h_divergence(int n, double d1, const double *x, const double *u, const double *flux, double *xa,
                         int adv)
{
    int i1 = 4, i2 = n - 2;
    double fold = L(flux, i1 - 1);
    for (int i = i1; i <= i2; ++i) {
        L(xa, i) = L(xa, i) + (L(flux, i) - fold) * d1;
        fold = L(flux, i);
    }
    if (adv) {
        double uold = L(u, i2 + 1);
        for (int i = i2; i >= i1; --i) {
            L(xa, i) = L(xa, i) - L(x, i) * (uold - L(u, i)) * d1;
            uold = L(u, i);
        }
    }
}

/* z divergence of flux(1..nz-1) with the 'nnf' closures (the common tail of
 * the QUICK/LW/MUSCL z kernels, e.g. advq_quick.f90:286-305) */
static void z_divergence(const orc_grid *g, const double *x, const double *w, const double *flux, double *xa,
                         int adv)
{
    int nz = g->nz;
    const double *dz1w = g->dz1w;
    double fold = L(flux, 1);
    for (int k = 2; k <= nz - 1; ++k) {
        L(xa, k) = L(xa, k) + (L(flux, k) - fold) * L(dz1w, k);
        fold = L(flux, k);
    }
    if (adv) {
        double wold = L(w, nz);
        for (int k = nz - 1; k >= 2; --k) {
            L(xa, k) = L(xa, k) - L(x, k) * (wold - L(w, k)) * L(dz1w, k);
            wold = L(w, k);
        }
    }
    if (g->bcl[4] == ORC_BC_NNF) L(xa, 1) = L(xa, 1) + (L(flux, 1) - 0.0) * L(dz1w, 1);
    if (g->bcl[4] == ORC_BC_NNF && adv) L(xa, 1) = L(xa, 1) - L(x, 1) * (L(w, 2) - 0.0) * L(dz1w, 1);
    if (g->bcl[5] == ORC_BC_NNF) L(xa, nz) = L(xa, nz) + (0.0 - L(flux, nz - 1)) * L(dz1w, nz);
    if (g->bcl[5] == ORC_BC_NNF && adv) L(xa, nz) = L(xa, nz) - L(x, nz) * (0.0 - L(w, nz)) * L(dz1w, nz);
}

/* advsx_quick / advsy_quick (advq_quick.f90:54-122) and the unlimited
 * advsx_quick_nl (advq_quick_nolim.f90:54-111) */
static void quick_h(const orc_opts *o, int unlimited, int n, double d, const double *x, const double *u,
                    double *xa, int adv)
{
    int i1 = 4, i2 = n - 2;
    double d1 = 1.0 / d;
    double *flux = (double *)calloc((size_t)n + 1, sizeof(double));
    for (int i = i1 - 1; i <= i2; ++i) {
        double uu = L(u, i + 1), xx = L(x, i);
        double ffip1 = L(x, i + 1) - xx;
        double ffip2 = L(x, i + 2) - L(x, i + 1);
        double ffim1 = xx - L(x, i - 1);
        double phil, phir, fl, fr;
        if (unlimited) {
            const double cc = 1.0 / 4.0;
            phil = ffip1 - cc * (ffip1 - ffim1);
            phir = ffip1 + cc * (ffip2 - ffip1);
            fl = xx + 0.5 * phil;
            fr = L(x, i + 1) - 0.5 * phir;
        } else {
            quick_limiter(o->lim_tol, xx, ffip1, ffip2, ffim1, &phil, &phir);
            fl = xx + 0.5 * phil * ffim1;
            fr = L(x, i + 1) - 0.5 * phir * ffip2;
        }
        L(flux, i) = half_upwind(uu, fl, fr);
    }
    h_divergence(n, d1, x, u, flux, xa, adv);
    free(flux);
}

/* advsz_quick (advq_quick.f90:197-311), advsz_quick_nl
 * (advq_quick_nolim.f90:212-318): per-metre slopes, face offsets by
 * dzw(k) on both sides, guards against x/dz */
static void quick_z(const orc_grid *g, const orc_opts *o, int unlimited, const double *x, const double *w,
                    double *xa, int adv)
{
    int nz = g->nz;
    double dz = g->dz;
    const double *dz1 = g->dz1, *dz1w = g->dz1w;
    double bc51 = (g->bcl[4] == ORC_BC_PER) ? 1.0 : 0.0;
    double bc61 = (g->bcl[5] == ORC_BC_PER) ? 1.0 : 0.0;
    double *flux = (double *)calloc((size_t)nz + 1, sizeof(double));
    for (int k = 1; k <= nz - 1; ++k) {
        double xx = L(x, k), ww = L(w, k + 1);
        double ffip1 = (L(x, k + 1) - xx) * L(dz1, k);
        double ffip2, ffim1;
        if (k == 1) {
            ffip2 = (L(x, 3) - L(x, 2)) * L(dz1, 2);
            ffim1 = bc51 * (L(x, nz - 4) - L(x, nz - 5)) * L(dz1, nz - 5) + (1.0 - bc51) * ffip1;
        } else if (k == nz - 1) {
            ffim1 = (xx - L(x, k - 1)) * L(dz1, k - 1);
            ffip2 = bc61 * (L(x, 6) - L(x, 5)) * L(dz1, 5) + (1.0 - bc61) * ffip1;
        } else {
            ffip2 = (L(x, k + 2) - L(x, k + 1)) * L(dz1, k + 1);
            ffim1 = (xx - L(x, k - 1)) * L(dz1, k - 1);
        }
        double phil, phir, fl, fr;
        if (unlimited) {
            const double cc = 1.0 / 4.0;
            phil = ffip1 - cc * (ffip1 - ffim1);
            phir = ffip1 + cc * (ffip2 - ffip1);
            fl = xx + 0.5 * phil / L(dz1w, k);
            fr = L(x, k + 1) - 0.5 * phir / L(dz1w, k);
        } else {
            quick_limiter(o->lim_tol, xx / dz, ffip1, ffip2, ffim1, &phil, &phir);
            fl = xx + 0.5 * phil * ffim1 / L(dz1w, k);
            fr = L(x, k + 1) - 0.5 * phir * ffip2 / L(dz1w, k);
        }
        L(flux, k) = half_upwind(ww, fl, fr);
    }
    z_divergence(g, x, w, flux, xa, adv);
    free(flux);
}

/* ======================================================================
 * Lax-Wendroff: advq_lw.f90 (flux-limited, lwlimiter = 'MC', the module's
 * value, which nothing in the model changes) and advq_lw_nolim.f90
 * ====================================================================== */

/* limiter_lw with lwlimiter = 'MC' (advq_lw.f90:349-373,389-391) */
static void lw_limiter(double lt, double xx, double ffip1, double ffip2, double ffim1, double *phir,
                       double *phil)
{
    const double eps = 1.e-18;
    double rp1 = 1.0, rp2 = 1.0;
    if (fabs(ffim1) > lt * fabs(xx) + eps) rp1 = ffip1 / ffim1;
    if (fabs(ffip2) > lt * fabs(xx) + eps) rp2 = ffip1 / ffip2;
    *phil = dmax(dmin(dmin(2.0 * rp1, 0.5 * (rp1 + 1.0)), 2.0), 0.0);
    *phir = dmax(dmin(dmin(2.0 * rp2, 0.5 * (rp2 + 1.0)), 2.0), 0.0);
}

static double lw_flux(double uu, double fl, double fr)
{
    return 0.5 * uu * (fl * (sgn1(uu) + 1.0) - fr * (sgn1(uu) - 1.0));
}

/* advsx_lw / advsy_lw (advq_lw.f90:57-126), advsx_lw_nl
 * (advq_lw_nolim.f90:54-111); Courant number from the full step dt0 */
static void lw_h(const orc_opts *o, int unlimited, int n, double d, const double *x, const double *u,
                 double *xa, int adv)
{
    const double eps = 1.e-18;
    int i1 = 4, i2 = n - 2;
    double d1 = 1.0 / d, lt = o->lim_tol;
    double *flux = (double *)calloc((size_t)n + 1, sizeof(double));
    for (int i = i1 - 1; i <= i2; ++i) {
        double uu = L(u, i + 1), xx = L(x, i);
        double ffip1 = L(x, i + 1) - xx;
        double mu = fabs(uu) * o->dt0 * d1;
        double fl, fr;
        if (unlimited) {
            fl = xx + 0.5 * ffip1 * (1.0 - mu);
            fr = L(x, i + 1) - 0.5 * ffip1 * (1.0 - mu);
            L(flux, i) = half_upwind(uu, fl, fr);
            continue;
        }
        double ffim1 = xx - L(x, i - 1);
        double ffip2 = L(x, i + 2) - L(x, i + 1);
        if (fabs(ffim1) > lt * fabs(xx) + eps || fabs(ffip2) > lt * fabs(xx) + eps) {
            double phir, phil;
            lw_limiter(lt, xx, ffip1, ffip2, ffim1, &phir, &phil);
            fl = xx + 0.5 * phil * ffim1 * (1.0 - mu);
            fr = L(x, i + 1) - 0.5 * phir * ffip2 * (1.0 - mu);
        } else {
            fl = xx + 0.5 * ffip1 * (1.0 - mu);
            fr = L(x, i + 1) - 0.5 * ffip1 * (1.0 - mu);
        }
        L(flux, i) = lw_flux(uu, fl, fr);
    }
    h_divergence(n, d1, x, u, flux, xa, adv);
    free(flux);
}

/* advsz_lw (advq_lw.f90:210-332), advsz_lw_nl (advq_lw_nolim.f90:188-265) */
static void lw_z(const orc_grid *g, const orc_opts *o, int unlimited, const double *x, const double *w,
                 double *xa, int adv)
{
    const double eps = 1.e-18;
    int nz = g->nz;
    double dz = g->dz, lt = o->lim_tol;
    const double *dz1 = g->dz1, *dz1w = g->dz1w;
    double bc51 = (g->bcl[4] == ORC_BC_PER) ? 1.0 : 0.0;
    double bc61 = (g->bcl[5] == ORC_BC_PER) ? 1.0 : 0.0;
    double *flux = (double *)calloc((size_t)nz + 1, sizeof(double));
    for (int k = 1; k <= nz - 1; ++k) {
        double xx = L(x, k), ww = L(w, k + 1);
        double mu = fabs(ww) * o->dt0 * L(dz1, k);
        double ffip1 = (L(x, k + 1) - xx) * L(dz1, k);
        double fl, fr;
        if (unlimited) {
            fl = xx + 0.5 * ffip1 * (1.0 - mu) / L(dz1w, k);
            fr = L(x, k + 1) - 0.5 * ffip1 * (1.0 - mu) / L(dz1w, k);
            L(flux, k) = half_upwind(ww, fl, fr);
            continue;
        }
        double ffip2, ffim1;
        if (k == 1) {
            ffip2 = (L(x, 3) - L(x, 2)) * L(dz1, 2);
            ffim1 = bc51 * (L(x, nz - 4) - L(x, nz - 5)) * L(dz1, nz - 5) + (1.0 - bc51) * ffip1;
        } else if (k == nz - 1) {
            ffim1 = (xx - L(x, k - 1)) * L(dz1, k - 1);
            ffip2 = bc61 * (L(x, 6) - L(x, 5)) * L(dz1, 5) + (1.0 - bc61) * ffip1;
        } else {
            ffim1 = (xx - L(x, k - 1)) * L(dz1, k - 1);
            ffip2 = (L(x, k + 2) - L(x, k + 1)) * L(dz1, k + 1);
        }
        if (dz * fabs(ffim1) > lt * fabs(xx) + eps || dz * fabs(ffip2) > lt * fabs(xx) + eps) {
            double phir, phil;
            lw_limiter(lt, xx, ffip1, ffip2, ffim1, &phir, &phil);
            fl = xx + 0.5 * phil * ffim1 * (1.0 - mu) / L(dz1w, k);
            fr = L(x, k + 1) - 0.5 * phir * ffip2 * (1.0 - mu) / L(dz1w, k);
        } else {
            fl = xx + 0.5 * ffip1 * (1.0 - mu) / L(dz1w, k);
            fr = L(x, k + 1) - 0.5 * ffip1 * (1.0 - mu) / L(dz1w, k);
        }
        L(flux, k) = lw_flux(ww, fl, fr);
    }
    z_divergence(g, x, w, flux, xa, adv);
    free(flux);
}

/* The line kernel advection_{x,y,z} selects (advection.f90:580-620,
 * :661-701, :743-782): unlimited ('limit' false) runs unlimited QUICK for
 * scal_adv = 'quick' and unlimited Lax-Wendroff for every other value;
 * limited with sorder = 1 runs the limited Lax-Wendroff whatever scal_adv
 * is; otherwise scal_adv's own scheme. */
int orc_line_scheme(int scheme, int unlimited, int sorder)
{
    if (unlimited) return scheme == ORC_QUICK ? ORC_QUICK_NL : ORC_LW_NL;
    if (sorder == 1) return ORC_LW;
    return scheme;
}

/* ======================================================================
 * Extension (not in the reference): WENO5-JS on x/y lines, periodic data
 * with the reference ghost layout (3 west, 2 east), and in z inside
 * weno3_z where the upwind stencil fits the column.  Jiang & Shu (1996):
 * candidate stencils (2,-7,11)/6, (-1,5,2)/6, (2,5,-1)/6; linear weights
 * (0.1, 0.6, 0.3); smoothness indicators beta_r; eps = 1e-6; p = 2.
 * The x(n+1) value needed by the last face for negative velocity is taken
 * from its periodic image x(6).
 * ====================================================================== */
static double weno5_left(double a, double b, double c, double d, double e)
{
    /* reconstruction at the right face of c from a..e = x(i-2..i+2) */
    const double eps = 1e-6;
    double q0 = (2.0 * a - 7.0 * b + 11.0 * c) / 6.0;
    double q1 = (-b + 5.0 * c + 2.0 * d) / 6.0;
    double q2 = (2.0 * c + 5.0 * d - e) / 6.0;
    double b0 = 13.0 / 12.0 * (a - 2.0 * b + c) * (a - 2.0 * b + c) + 0.25 * (a - 4.0 * b + 3.0 * c) * (a - 4.0 * b + 3.0 * c);
    double b1 = 13.0 / 12.0 * (b - 2.0 * c + d) * (b - 2.0 * c + d) + 0.25 * (b - d) * (b - d);
    double b2 = 13.0 / 12.0 * (c - 2.0 * d + e) * (c - 2.0 * d + e) + 0.25 * (3.0 * c - 4.0 * d + e) * (3.0 * c - 4.0 * d + e);
    double a0 = 0.1 / ((eps + b0) * (eps + b0));
    double a1 = 0.6 / ((eps + b1) * (eps + b1));
    double a2 = 0.3 / ((eps + b2) * (eps + b2));
    return (a0 * q0 + a1 * q1 + a2 * q2) / (a0 + a1 + a2);
}

static void weno5_h(int n, double d, const double *x, const double *u, double *xa, int adv)
{
    int ii1 = 3, ii2 = n - 2;
    double d1 = 1.0 / d;
    double *flux = (double *)calloc((size_t)n + 1, sizeof(double));
#define XP(i) ((i) == n + 1 ? L(x, 6) : L(x, i))
    for (int i = ii1; i <= ii2; ++i) {
        double uu = L(u, i + 1);
        double fl = weno5_left(XP(i - 2), XP(i - 1), XP(i), XP(i + 1), XP(i + 2));
        double fr = weno5_left(XP(i + 3), XP(i + 2), XP(i + 1), XP(i), XP(i - 1));
        L(flux, i) = 0.5 * fl * (uu + fabs(uu)) + 0.5 * fr * (uu - fabs(uu));
    }
#undef XP
    double fold = L(flux, ii1);
    for (int i = ii1 + 1; i <= ii2; ++i) {
        L(xa, i) = L(xa, i) + (L(flux, i) - fold) * d1;
        fold = L(flux, i);
    }
    if (adv) {
        double uold = L(u, ii2 + 1);
        for (int i = ii2; i >= ii1 + 1; --i) {
            L(xa, i) = L(xa, i) - L(x, i) * (uold - L(u, i)) * d1;
            uold = L(u, i);
        }
    }
    free(flux);
}

/* ======================================================================
 * Dispatch and 3D drivers
 * ====================================================================== */

void orc_advs_h(const orc_opts *o, int n, double d, int bc_lo, int bc_hi,
                const double *x, const double *u, double *xa, int adv)
{
    (void)bc_lo;
    switch (o->scheme) {
    case ORC_WENO: weno3_h(n, d, x, u, xa, adv); break;
    case ORC_PPM: ppm_h(o, n, d, bc_hi, x, u, xa, adv); break;
    case ORC_MUSCL: muscl_h(o, n, d, x, u, xa, adv); break;
    case ORC_WENO5: weno5_h(n, d, x, u, xa, adv); break;
    case ORC_QUICK: quick_h(o, 0, n, d, x, u, xa, adv); break;
    case ORC_QUICK_NL: quick_h(o, 1, n, d, x, u, xa, adv); break;
    case ORC_LW: lw_h(o, 0, n, d, x, u, xa, adv); break;
    case ORC_LW_NL: lw_h(o, 1, n, d, x, u, xa, adv); break;
    default: break;
    }
}

void orc_advs_z(const orc_grid *g, const orc_opts *o, const double *x, double x0,
                const double *w, double *xa, int adv)
{
    switch (o->scheme) {
    case ORC_WENO: weno3_z(g, o, x, x0, w, xa, adv, 0); break;
    case ORC_WENO5: weno3_z(g, o, x, x0, w, xa, adv, 1); break;
    case ORC_PPM: ppm_z(g, o, x, w, xa, adv); break;
    case ORC_MUSCL: muscl_z(g, o, x, w, xa, adv); break;
    case ORC_QUICK: quick_z(g, o, 0, x, w, xa, adv); break;
    case ORC_QUICK_NL: quick_z(g, o, 1, x, w, xa, adv); break;
    case ORC_LW: lw_z(g, o, 0, x, w, xa, adv); break;
    case ORC_LW_NL: lw_z(g, o, 1, x, w, xa, adv); break;
    default: break;
    }
}

/* advection_{x,y,z} (advection.f90:556-798): xx = 0; per-line kernels over
 * the interior tile; xs = xs - xx. */
void orc_advection_dir(const orc_grid *g, const orc_opts *o, int dir,
                       const double *q, const double *vel, double *xs, int adv)
{
    int nx = g->nx, ny = g->ny, nz = g->nz;
    int three_d = ny > 1;
    int jt0 = three_d ? 4 : 1, jt1 = three_d ? ny - 2 : 1;
    size_t ntot = (size_t)nx * ny * nz;
    double *xx = (double *)calloc(ntot, sizeof(double));
    if (dir == 0) {
#pragma omp parallel for collapse(2) num_threads(g_threads) schedule(static)
        for (int k = 1; k <= nz; ++k)
            for (int j = jt0; j <= jt1; ++j) {
                size_t b = IDX(g, 1, j, k);
                orc_advs_h(o, nx, g->dx, g->bcl[0], g->bcl[1], q + b, vel + b, xx + b, adv);
            }
    } else if (dir == 1) {
#pragma omp parallel num_threads(g_threads)
        {
            double *lx = (double *)malloc(sizeof(double) * (size_t)ny * 3);
            double *lv = lx + ny, *la = lx + 2 * ny;
#pragma omp for collapse(2) schedule(static)
            for (int k = 1; k <= nz; ++k)
                for (int i = 4; i <= nx - 2; ++i) {
                    for (int j = 1; j <= ny; ++j) {
                        size_t id = IDX(g, i, j, k);
                        L(lx, j) = q[id];
                        L(lv, j) = vel[id];
                        L(la, j) = 0.0;
                    }
                    orc_advs_h(o, ny, g->dy, g->bcl[2], g->bcl[3], lx, lv, la, adv);
                    for (int j = 1; j <= ny; ++j) xx[IDX(g, i, j, k)] = L(la, j);
                }
            free(lx);
        }
    } else {
#pragma omp parallel num_threads(g_threads)
        {
            double *lx = (double *)malloc(sizeof(double) * (size_t)nz * 3);
            double *lw = lx + nz, *la = lx + 2 * nz;
#pragma omp for collapse(2) schedule(static)
            for (int j = jt0; j <= jt1; ++j)
                for (int i = 4; i <= nx - 2; ++i) {
                    for (int k = 1; k <= nz; ++k) {
                        size_t id = IDX(g, i, j, k);
                        L(lx, k) = q[id];
                        L(lw, k) = vel[id];
                        L(la, k) = 0.0;
                    }
                    /* x(0) convention for the reference's out-of-bounds read */
                    orc_advs_z(g, o, lx, L(lx, 1), lw, la, adv);
                    for (int k = 1; k <= nz; ++k) xx[IDX(g, i, j, k)] = L(la, k);
                }
            free(lx);
        }
    }
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (size_t n = 0; n < ntot; ++n) xs[n] = xs[n] - xx[n];
    free(xx);
}

void orc_scalar_tendency(const orc_grid *g, const orc_opts *o, const double *q,
                         const double *u, const double *v, const double *w,
                         double *xs, int adv, int dirmask)
{
    if (dirmask & 1) orc_advection_dir(g, o, 0, q, u, xs, adv);
    if ((dirmask & 2) && g->ny > 1) orc_advection_dir(g, o, 1, q, v, xs, adv);
    if (dirmask & 4) orc_advection_dir(g, o, 2, q, w, xs, adv);
}

/* ======================================================================
 * Positivity-preserving stage (no reference counterpart; parity pinned by
 * tests/test_oracle_positive.py)
 * ======================================================================
 *
 * One Shu-Osher stage out = c0*q0 + c1*(q + dt*L(q)) of the flux-form
 * 'weno' tendency in which every cell's outgoing face fluxes are scaled by
 * one factor theta so that the Euler step q + dt*L(q) cannot take the cell
 * below zero:
 *   O     = dt * sum over the cell's faces of the outgoing flux / width
 *           (x: (max(F_i,0) + max(-F_{i-1},0)) dx1, then y, then z with
 *           dz1w(k); the floor and lid faces carry no flux),
 *   theta = O > max(q,0) ? (max(q,0) / O) * (1 - 2^-40) : 1
 *           (the margin keeps the limited cell at +2^-40 q instead of an
 *           ulp below zero after the rounding of the divergence),
 * and a face flux F is multiplied by theta of its donor cell (left/lower
 * if F > 0, right/upper if F < 0; across the periodic sides the donor is
 * the periodic image).  Each flux enters both neighbours with one value,
 * so mass is conserved; inflows are never scaled, so a non-negative q
 * stays non-negative; with theta = 1 everywhere the stage equals the
 * reference stage bit for bit (the divergence is formed in its order).
 * Periodic sides, 'nnf' floor and lid, advective form off.  Returns 0, or
 * -1 for an unsupported scheme or boundary. */
int orc_positive_stage(const orc_grid *g, const orc_opts *o, const double *q, const double *u, const double *v,
                       const double *w, const double *q0, double c0, double c1, double dt, double *out)
{
    int nx = g->nx, ny = g->ny, nz = g->nz;
    int three_d = ny > 1;
    if (o->scheme != ORC_WENO || g->bcl[4] != ORC_BC_NNF || g->bcl[5] != ORC_BC_NNF) return -1;
    if (g->bcl[0] != ORC_BC_PER || g->bcl[1] != ORC_BC_PER) return -1;
    if (three_d && (g->bcl[2] != ORC_BC_PER || g->bcl[3] != ORC_BC_PER)) return -1;
    int mx = nx - 5, my = three_d ? ny - 5 : 1;
    int jt0 = three_d ? 4 : 1, jt1 = three_d ? ny - 2 : 1;
    size_t n3 = (size_t)nx * ny * nz;
    double *fx = (double *)calloc(n3, sizeof(double));
    double *fy = (double *)calloc(n3, sizeof(double));
    double *fz = (double *)calloc(n3, sizeof(double));
    double *th = (double *)calloc(n3, sizeof(double));
    double dx1 = 1.0 / g->dx, dy1 = 1.0 / g->dy;
    /* face fluxes: x faces i+1/2 (i = 3..nx-2) of the interior rows */
    for (int k = 1; k <= nz; ++k)
        for (int j = jt0; j <= jt1; ++j) {
            size_t b = IDX(g, 1, j, k);
            weno3_h_flux(nx, q + b, u + b, fx + b);
        }
    /* y faces j+1/2 (j = 3..ny-2) of the interior columns */
    if (three_d) {
        double *lx = (double *)malloc(sizeof(double) * (size_t)ny * 3);
        double *lv = lx + ny, *lf = lx + 2 * ny;
        for (int k = 1; k <= nz; ++k)
            for (int i = 4; i <= nx - 2; ++i) {
                for (int j = 1; j <= ny; ++j) {
                    L(lx, j) = q[IDX(g, i, j, k)];
                    L(lv, j) = v[IDX(g, i, j, k)];
                    L(lf, j) = 0.0;
                }
                weno3_h_flux(ny, lx, lv, lf);
                for (int j = 1; j <= ny; ++j) fy[IDX(g, i, j, k)] = L(lf, j);
            }
        free(lx);
    }
    /* z faces k+1/2 (k = 1..nz-1), x(0) = x(1) */
    {
        double *lx = (double *)malloc(sizeof(double) * (size_t)nz * 3);
        double *lw = lx + nz, *lf = lx + 2 * nz;
        for (int j = jt0; j <= jt1; ++j)
            for (int i = 4; i <= nx - 2; ++i) {
                for (int k = 1; k <= nz; ++k) {
                    L(lx, k) = q[IDX(g, i, j, k)];
                    L(lw, k) = w[IDX(g, i, j, k)];
                    L(lf, k) = 0.0;
                }
                weno3_z_flux(g, o, lx, L(lx, 1), lw, lf, 0);
                for (int k = 1; k <= nz; ++k) fz[IDX(g, i, j, k)] = L(lf, k);
            }
        free(lx);
    }
#define POS(f) ((f) > 0.0 ? (f) : 0.0)
#define NEG(f) ((f) < 0.0 ? -(f) : 0.0)
    /* theta of every interior cell */
    for (int k = 1; k <= nz; ++k)
        for (int j = jt0; j <= jt1; ++j)
            for (int i = 4; i <= nx - 2; ++i) {
                size_t id = IDX(g, i, j, k);
                double ox = (POS(fx[id]) + NEG(fx[IDX(g, i - 1, j, k)])) * dx1;
                double o = ox;
                if (three_d) o = o + (POS(fy[id]) + NEG(fy[IDX(g, i, j - 1, k)])) * dy1;
                double fzt = k < nz ? fz[id] : 0.0, fzb = k > 1 ? fz[IDX(g, i, j, k - 1)] : 0.0;
                o = o + (POS(fzt) + NEG(fzb)) * L(g->dz1w, k);
                o = dt * o;
                double qp = q[id] > 0.0 ? q[id] : 0.0;
                th[id] = o > qp ? (qp / o) * (1.0 - 0x1p-40) : 1.0;
            }
#undef POS
#undef NEG
#define WX(i) ((i) < 4 ? (i) + mx : ((i) > nx - 2 ? (i) - mx : (i)))
#define WY(j) (!three_d ? 1 : ((j) < 4 ? (j) + my : ((j) > ny - 2 ? (j) - my : (j))))
    /* flux of x face f+1/2 of row (j, k), scaled by its donor's theta */
#define FXS(f, j, k) (fx[IDX(g, f, j, k)] * (fx[IDX(g, f, j, k)] > 0.0 ? th[IDX(g, WX(f), j, k)] \
                                                                       : th[IDX(g, WX((f) + 1), j, k)]))
#define FYS(i, f, k) (fy[IDX(g, i, f, k)] * (fy[IDX(g, i, f, k)] > 0.0 ? th[IDX(g, i, WY(f), k)] \
                                                                       : th[IDX(g, i, WY((f) + 1), k)]))
#define FZS(i, j, f) (fz[IDX(g, i, j, f)] * (fz[IDX(g, i, j, f)] > 0.0 ? th[IDX(g, i, j, f)] : th[IDX(g, i, j, (f) + 1)]))
    for (size_t n = 0; n < n3; ++n) out[n] = q[n];
    for (int k = 1; k <= nz; ++k)
        for (int j = jt0; j <= jt1; ++j)
            for (int i = 4; i <= nx - 2; ++i) {
                size_t id = IDX(g, i, j, k);
                double X = (FXS(i, j, k) - FXS(i - 1, j, k)) * dx1;
                double Y = three_d ? (FYS(i, j, k) - FYS(i, j - 1, k)) * dy1 : 0.0;
                double Z;
                if (k == 1) Z = (FZS(i, j, 1) - 0.0) * L(g->dz1w, 1);
                else if (k == nz) Z = (0.0 - FZS(i, j, nz - 1)) * L(g->dz1w, nz);
                else Z = (FZS(i, j, k) - FZS(i, j, k - 1)) * L(g->dz1w, k);
                double A = 0.0;
                A = A - X;
                if (three_d) A = A - Y;
                A = A - Z;
                out[id] = c0 * q0[id] + c1 * (q[id] + dt * A);
            }
#undef FXS
#undef FYS
#undef FZS
#undef WX
#undef WY
    orc_fill_scalar(g, out);
    free(fx);
    free(fy);
    free(fz);
    free(th);
    return 0;
}

/* ======================================================================
 * Large-scale vertical advection: advection_1d / advection_ac_1d
 * (advection.f90:918-1042) with grad_1d (gradients.f90:400-473)
 * ====================================================================== */

/* grad_1d: gz(k) for k = 1..nz-1 from the quadratic through levels k-1, k,
 * k+1 (the floor reads the value of level 1 at distance dz*fdz0(1));
 * gz(nz) = 0 */
void orc_grad_1d(const orc_grid *g, const double *xz, double *gz)
{
    int nz = g->nz;
    double dz = g->dz;
    const double *fdz0 = g->fdz0;
    for (int k = 1; k <= nz; ++k) L(gz, k) = 0.0;
    for (int k = 1; k <= nz - 1; ++k) {
        double deltad = 0.5 * dz * (L(fdz0, k + 1) + L(fdz0, k));
        double deltau = k == 1 ? dz * L(fdz0, 1) : 0.5 * dz * (L(fdz0, k) + L(fdz0, k - 1));
        double deltac = (deltad + deltau) * (deltad * deltad + deltau * deltau) -
                        (deltad - deltau) * (deltad * deltad - deltau * deltau);
        double ffip1 = L(xz, k + 1) - L(xz, k);
        double ffim1 = k == 1 ? 0.0 : L(xz, k) - L(xz, k - 1);
        L(gz, k) = ((ffip1 + ffim1) * (deltad * deltad + deltau * deltau) -
                    (ffip1 - ffim1) * (deltad * deltad - deltau * deltau)) / deltac;
    }
}

/* advection_ac_1d: xs(i,j,:) = xs(i,j,:) - dref*w*grad_1d(xx/dref) on the
 * interior columns; dref is the level profile dref_prof[nz] (den0 of the
 * anelastic conservative build), the 3D array dref3 (pressure%dens of the
 * compressible conservative build), or 1 when both are NULL */
void orc_advection_1d(const orc_grid *g, const double *w0, const double *dref_prof, const double *dref3,
                      const double *xx, double *xs)
{
    int nx = g->nx, ny = g->ny, nz = g->nz;
    int three_d = ny > 1;
    int jt0 = three_d ? 4 : 1, jt1 = three_d ? ny - 2 : 1;
#pragma omp parallel num_threads(g_threads)
    {
        double *xz = (double *)malloc(sizeof(double) * (size_t)nz * 3);
        double *gz = xz + nz, *dr = xz + 2 * nz;
#pragma omp for collapse(2) schedule(static)
        for (int j = jt0; j <= jt1; ++j)
            for (int i = 4; i <= nx - 2; ++i) {
                for (int k = 1; k <= nz; ++k) {
                    size_t id = IDX(g, i, j, k);
                    L(dr, k) = dref3 ? dref3[id] : (dref_prof ? L(dref_prof, k) : 1.0);
                    L(xz, k) = xx[id] / L(dr, k);
                }
                orc_grad_1d(g, xz, gz);
                for (int k = 1; k <= nz; ++k) {
                    size_t id = IDX(g, i, j, k);
                    xs[id] = xs[id] - L(dr, k) * L(w0, k) * L(gz, k);
                }
            }
        free(xz);
    }
}

/* ======================================================================
 * Momentum: windadv.f90
 * ====================================================================== */

static const double AVISC = 1.0 / 16.0;

static void mom_coefs(int mom_ord, int self, double *c1, double *c2)
{
    /* self-advection lines use (4/3, 1/3); cross lines (9/8, 1/8) */
    if (mom_ord == 4) {
        *c1 = self ? 4.0 / 3.0 : 9.0 / 8.0;
        *c2 = self ? 1.0 / 3.0 : 1.0 / 8.0;
    } else {
        *c1 = 1.0;
        *c2 = 0.0;
    }
}

static double visc5(const double *u, int i)
{
    return -L(u, i + 2) + 4.0 * L(u, i + 1) - 6.0 * L(u, i) + 4.0 * L(u, i - 1) - L(u, i - 2);
}

/* advux_fd / advvy_fd (windadv.f90:494-554, 754-826): self advection along a
 * line of length n with spacing d, interior i0..i1. */
static void mom_self_h(const orc_opts *o, int n, int i0, int i1, double d, const double *u,
                       const double *ut, double *ua)
{
    double c1, c2;
    mom_coefs(o->mom_ord, 1, &c1, &c2);
    double d1 = 1.0 / d, d2 = 1.0 / (2.0 * d);
    double *flux = (double *)calloc((size_t)n + 1, sizeof(double));
    double *advf = (double *)calloc((size_t)n + 1, sizeof(double));
    for (int i = i0 - 2; i <= i1; ++i) {
        double uu = (L(ut, i + 1) + L(ut, i)) * 0.5;
        double uu2 = (L(ut, i + 2) + L(ut, i)) * 0.5;
        double fr = (L(u, i + 1) + L(u, i)) * 0.5;
        double fr2 = (L(u, i + 2) + L(u, i)) * 0.5;
        L(flux, i) = uu * fr;
        L(advf, i) = uu2 * fr2;
    }
    for (int i = i0; i <= i1; ++i)
        L(ua, i) = L(ua, i) + c1 * (L(flux, i) - L(flux, i - 1)) * d1 - c2 * (L(advf, i) - L(advf, i - 2)) * d2;
    if (o->lavisc)
        for (int i = i0; i <= i1; ++i) {
            double mu = fabs(L(ut, i)) * d1;
            L(ua, i) = L(ua, i) - AVISC * mu * visc5(u, i);
        }
    free(flux);
    free(advf);
}

/* advuy_fd / advvx_fd (windadv.f90:556-626, 696-752): cross advection along a
 * line; t1..t4 are the advecting component at offsets (-1, 0, -2, +1) of the
 * transverse index. */
static void mom_cross_h(const orc_opts *o, int n, int i0, int i1, double d, const double *u,
                        const double *t1, const double *t2, const double *t3, const double *t4,
                        double *ua)
{
    double c1, c2;
    mom_coefs(o->mom_ord, 0, &c1, &c2);
    double d1 = 1.0 / d, d3 = 1.0 / (3.0 * d);
    double *flux = (double *)calloc((size_t)n + 1, sizeof(double));
    for (int i = i0 - 2; i <= i1 + 1; ++i) {
        double vv = c1 * (L(t1, i + 1) + L(t2, i + 1)) * 0.5 - c2 * (L(t3, i + 1) + L(t4, i + 1)) * 0.5;
        double fr = (L(u, i + 1) + L(u, i)) * 0.5;
        L(flux, i) = vv * fr;
    }
    for (int i = i0; i <= i1; ++i)
        L(ua, i) = L(ua, i) + c1 * (L(flux, i) - L(flux, i - 1)) * d1 - c2 * (L(flux, i + 1) - L(flux, i - 2)) * d3;
    if (o->lavisc)
        for (int i = i0; i <= i1; ++i) {
            double mu = 0.25 * fabs(L(t1, i) + L(t2, i) + L(t1, i + 1) + L(t2, i + 1)) * d1;
            L(ua, i) = L(ua, i) - AVISC * mu * visc5(u, i);
        }
    free(flux);
}

/* advuz_fd / advvz_fd (windadv.f90:628-694, 828-894).  ua2_metric_k is the
 * level of dz1w used for the k = 2 boundary line: 1 for u (:674), 2 for v
 * (:874). */
static void mom_cross_z(const orc_grid *g, const orc_opts *o, const double *u, const double *t1,
                        const double *t2, const double *t3, const double *t4, double *ua,
                        int ua2_metric_k)
{
    int nz = g->nz;
    const double *dzw = g->dzw, *dz1w = g->dz1w;
    double c1, c2;
    mom_coefs(o->mom_ord, 0, &c1, &c2);
    double *flux = (double *)calloc((size_t)nz + 1, sizeof(double));
    for (int k = 1; k <= nz - 1; ++k) {
        double ww = c1 * (L(t1, k + 1) + L(t2, k + 1)) * 0.5 - c2 * (L(t3, k + 1) + L(t4, k + 1)) * 0.5;
        double fr = (L(dzw, k + 1) * L(u, k + 1) + L(dzw, k) * L(u, k)) / (L(dzw, k + 1) + L(dzw, k));
        L(flux, k) = ww * fr;
    }
    if (g->bcl[5] == ORC_BC_NNF) L(flux, nz) = 0.0;
    if (g->bcl[5] == ORC_BC_OPE) L(flux, nz) = L(flux, nz - 1);
    for (int k = 3; k <= nz - 1; ++k)
        L(ua, k) = L(ua, k) + c1 * (L(flux, k) - L(flux, k - 1)) * L(dz1w, k) -
                   c2 * (L(flux, k + 1) - L(flux, k - 2)) / (L(dzw, k - 1) + L(dzw, k) + L(dzw, k + 1));
    double m2 = L(dz1w, ua2_metric_k);
    if (g->bcl[4] == ORC_BC_NNF) L(ua, 1) = L(ua, 1) + (L(flux, 1) - 0.0) * L(dz1w, 1);
    if (g->bcl[4] == ORC_BC_NNF)
        L(ua, 2) = L(ua, 2) + c1 * (L(flux, 2) - L(flux, 1)) * m2 - c2 * (L(flux, 3) - 0.0) / (L(dzw, 1) + L(dzw, 2) + L(dzw, 3));
    if (g->bcl[4] == ORC_BC_OPE) L(ua, 1) = L(ua, 1) + (L(flux, 1) - L(flux, 1)) * L(dz1w, 1);
    if (g->bcl[4] == ORC_BC_OPE)
        L(ua, 2) = L(ua, 2) + c1 * (L(flux, 2) - L(flux, 1)) * m2 - c2 * (L(flux, 3) - L(flux, 1)) / (L(dzw, 1) + L(dzw, 2) + L(dzw, 3));
    if (g->bcl[5] == ORC_BC_NNF) L(ua, nz) = L(ua, nz) + (0.0 - L(flux, nz - 1)) * L(dz1w, nz);
    if (g->bcl[5] == ORC_BC_OPE) L(ua, nz) = L(ua, nz) + (L(flux, nz - 1) - L(flux, nz - 1)) * L(dz1w, nz);
    if (o->lavisc)
        for (int k = 3; k <= nz - 2; ++k) {
            double mu = 0.25 * fabs(L(t1, k) + L(t2, k) + L(t1, k + 1) + L(t2, k + 1)) * L(dz1w, k);
            L(ua, k) = L(ua, k) - AVISC * mu * visc5(u, k);
        }
    free(flux);
}

/* advwx_fd / advwy_fd (windadv.f90:896-1056) for one x-z (or y-z) slice:
 * w and the advecting component t are 2D arrays (n, nz) with the line index
 * fastest (stride s between levels). */
static void mom_w_h(const orc_grid *g, const orc_opts *o, int n, int i0, int i1, double d,
                    const double *w, const double *t, double *wa, size_t s)
{
    int nz = g->nz;
    double c1, c2;
    mom_coefs(o->mom_ord, 0, &c1, &c2);
    double d1 = 1.0 / d, d3 = 1.0 / (3.0 * d);
#define W2(i, k) w[(size_t)((i) - 1) + s * (size_t)((k) - 1)]
#define T2(i, k) t[(size_t)((i) - 1) + s * (size_t)((k) - 1)]
#define WA2(i, k) wa[(size_t)((i) - 1) + s * (size_t)((k) - 1)]
    double *flux = (double *)calloc((size_t)(n + 1) * (size_t)(nz + 1), sizeof(double));
#define F2(i, k) flux[(size_t)((i) - 1) + (size_t)n * (size_t)((k) - 1)]
    for (int i = i0 - 2; i <= i1 + 1; ++i) {
        double uu = (T2(i + 1, 2) + T2(i + 1, 1)) * 0.5;
        double fr = (W2(i + 1, 2) + W2(i, 2)) * 0.5;
        F2(i, 2) = uu * fr;
        uu = (T2(i + 1, nz) + T2(i + 1, nz - 1)) * 0.5;
        fr = (W2(i + 1, nz) + W2(i, nz)) * 0.5;
        F2(i, nz) = uu * fr;
    }
    for (int k = 3; k <= nz - 1; ++k)
        for (int i = i0 - 2; i <= i1 + 1; ++i) {
            double uu = c1 * (T2(i + 1, k) + T2(i + 1, k - 1)) * 0.5 - c2 * (T2(i + 1, k + 1) + T2(i + 1, k - 2)) * 0.5;
            double fr = (W2(i + 1, k) + W2(i, k)) * 0.5;
            F2(i, k) = uu * fr;
        }
    for (int k = 2; k <= nz; ++k)
        for (int i = i0; i <= i1; ++i)
            WA2(i, k) = WA2(i, k) + c1 * (F2(i, k) - F2(i - 1, k)) * d1 - c2 * (F2(i + 1, k) - F2(i - 2, k)) * d3;
    if (o->lavisc)
        for (int k = 2; k <= nz; ++k)
            for (int i = i0; i <= i1; ++i) {
                double mu = 0.25 * fabs(T2(i, k) + T2(i, k - 1) + T2(i + 1, k) + T2(i + 1, k - 1)) * d1;
                WA2(i, k) = WA2(i, k) - AVISC * mu *
                            (-W2(i + 2, k) + 4.0 * W2(i + 1, k) - 6.0 * W2(i, k) + 4.0 * W2(i - 1, k) - W2(i - 2, k));
            }
#undef F2
#undef W2
#undef T2
#undef WA2
    free(flux);
}

/* advwz_fd (windadv.f90:1058-1127). */
static void mom_w_z(const orc_grid *g, const orc_opts *o, const double *w, const double *wt, double *wa)
{
    int nz = g->nz;
    const double *dzw = g->dzw, *dz1 = g->dz1;
    double c1, c2;
    mom_coefs(o->mom_ord, 1, &c1, &c2);
    double *flux = (double *)calloc((size_t)nz + 1, sizeof(double));
    double *advf = (double *)calloc((size_t)nz + 1, sizeof(double));
    for (int k = 1; k <= nz - 2; ++k) {
        double ww = (L(wt, k + 1) + L(wt, k)) * 0.5;
        double ww2 = (L(dzw, k + 1) * L(wt, k + 2) + L(dzw, k) * L(wt, k)) / (L(dzw, k + 1) + L(dzw, k));
        double fr = (L(w, k + 1) + L(w, k)) * 0.5;
        double fr2 = (L(dzw, k + 1) * L(w, k + 2) + L(dzw, k) * L(w, k)) / (L(dzw, k + 1) + L(dzw, k));
        L(flux, k) = ww * fr;
        L(advf, k) = ww2 * fr2;
    }
    L(flux, nz - 1) = 0.25 * (L(wt, nz) + L(wt, nz - 1)) * (L(w, nz) + L(w, nz - 1));
    {
        double s = L(dzw, nz) + L(dzw, nz - 1);
        if (g->bcl[5] == ORC_BC_NNF)
            L(advf, nz - 1) = (0.0 + L(dzw, nz - 1) * L(wt, nz - 1)) * (0.0 + L(dzw, nz - 1) * L(w, nz - 1)) / (s * s);
        if (g->bcl[5] == ORC_BC_OPE)
            L(advf, nz - 1) = (L(dzw, nz) * L(wt, nz) + L(dzw, nz - 1) * L(wt, nz - 1)) *
                              (L(dzw, nz) * L(w, nz) + L(dzw, nz - 1) * L(w, nz - 1)) / (s * s);
    }
    for (int k = 3; k <= nz - 1; ++k)
        L(wa, k) = L(wa, k) + c1 * (L(flux, k) - L(flux, k - 1)) * L(dz1, k - 1) -
                   c2 * (L(advf, k) - L(advf, k - 2)) / (L(dzw, k) + L(dzw, k - 1));
    if (g->bcl[4] == ORC_BC_NNF)
        L(wa, 2) = L(wa, 2) + c1 * (L(flux, 2) - L(flux, 1)) * L(dz1, 1) - c2 * (L(advf, 2) - 0.0) / (L(dzw, 2) + L(dzw, 1));
    if (g->bcl[4] == ORC_BC_OPE)
        L(wa, 2) = L(wa, 2) + c1 * (L(flux, 2) - L(flux, 1)) * L(dz1, 1) - c2 * (L(advf, 2) - L(flux, 1)) / (L(dzw, 2) + L(dzw, 1));
    if (g->bcl[5] == ORC_BC_NNF) L(wa, nz) = L(wa, nz) + (0.0 - L(flux, nz - 1)) * L(dz1, nz - 1);
    if (g->bcl[5] == ORC_BC_OPE) L(wa, nz) = L(wa, nz) + (L(flux, nz - 1) - L(flux, nz - 1)) * L(dz1, nz - 1);
    if (o->lavisc)
        for (int k = 3; k <= nz - 2; ++k) {
            double mu = fabs(L(wt, k)) * L(dz1, k - 1);
            L(wa, k) = L(wa, k) - AVISC * mu * visc5(w, k);
        }
    free(flux);
    free(advf);
}

/* One direction of one momentum component (windadv.f90:82-492). comp: 0 u,
 * 1 v, 2 w; dir: 0 x, 1 y, 2 z.  Accumulates into out. */
static void mom_dir(const orc_grid *g, const orc_opts *o, int comp, int dir, const double *m,
                    const double *ut, const double *vt, const double *wt, double *out)
{
    int nx = g->nx, ny = g->ny, nz = g->nz;
    int three_d = ny > 1;
    int jt0 = three_d ? 4 : 1, jt1 = three_d ? ny - 2 : 1;
    size_t sx = 1, sy = (size_t)nx, sz = (size_t)nx * ny;
    if (comp == 2 && dir != 2) {
        /* advwx_fd over x-z slices, advwy_fd over y-z slices */
        if (dir == 0) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
            for (int j = jt0; j <= jt1; ++j) {
                size_t b = IDX(g, 1, j, 1);
                mom_w_h(g, o, nx, 4, nx - 2, g->dx, m + b, ut + b, out + b, sz);
            }
        } else {
#pragma omp parallel num_threads(g_threads)
            {
                double *buf = (double *)malloc(sizeof(double) * (size_t)ny * nz * 3);
                double *lw = buf, *lt = buf + (size_t)ny * nz, *la = buf + 2 * (size_t)ny * nz;
#pragma omp for schedule(static)
                for (int i = 4; i <= nx - 2; ++i) {
                    for (int k = 1; k <= nz; ++k)
                        for (int j = 1; j <= ny; ++j) {
                            size_t id = IDX(g, i, j, k), l = (size_t)(j - 1) + (size_t)ny * (k - 1);
                            lw[l] = m[id];
                            lt[l] = vt[id];
                            la[l] = out[id];
                        }
                    mom_w_h(g, o, ny, 4, ny - 2, g->dy, lw, lt, la, (size_t)ny);
                    for (int k = 1; k <= nz; ++k)
                        for (int j = 1; j <= ny; ++j) out[IDX(g, i, j, k)] = la[(size_t)(j - 1) + (size_t)ny * (k - 1)];
                }
                free(buf);
            }
        }
        return;
    }
    /* line-wise kernels: gather line along dir, call, scatter */
    int n = dir == 0 ? nx : (dir == 1 ? ny : nz);
    size_t st = dir == 0 ? sx : (dir == 1 ? sy : sz);
    /* transverse offsets for the cross kernels: advecting component index
     * shifted along the component's own direction (u: x, v: y) */
    size_t sh = comp == 0 ? sx : sy;
#pragma omp parallel num_threads(g_threads)
    {
        double *buf = (double *)malloc(sizeof(double) * (size_t)n * 7);
        double *lm = buf, *la = buf + n, *l1 = buf + 2 * n, *l2 = buf + 3 * n, *l3 = buf + 4 * n, *l4 = buf + 5 * n;
        /* loop over the two transverse indices */
        int a0, a1, b0, b1;
        if (dir == 0) { a0 = jt0; a1 = jt1; b0 = 1; b1 = nz; }
        else if (dir == 1) { a0 = 4; a1 = nx - 2; b0 = 1; b1 = nz; }
        else { a0 = 4; a1 = nx - 2; b0 = jt0; b1 = jt1; }
#pragma omp for collapse(2) schedule(static)
        for (int b = b0; b <= b1; ++b)
            for (int a = a0; a <= a1; ++a) {
                size_t base;
                if (dir == 0) base = IDX(g, 1, a, b);
                else if (dir == 1) base = IDX(g, a, 1, b);
                else base = IDX(g, a, b, 1);
                const double *adv = dir == 0 ? ut : (dir == 1 ? vt : wt);
                for (int l = 1; l <= n; ++l) {
                    size_t id = base + st * (size_t)(l - 1);
                    L(lm, l) = m[id];
                    L(la, l) = out[id];
                    if (comp == dir) {
                        L(l1, l) = adv[id];
                    } else {
                        L(l1, l) = adv[id - sh];
                        L(l2, l) = adv[id];
                        L(l3, l) = adv[id - 2 * sh];
                        L(l4, l) = adv[id + sh];
                    }
                }
                if (comp == dir && dir == 2) {
                    mom_w_z(g, o, lm, l1, la);
                } else if (comp == dir) {
                    mom_self_h(o, n, 4, n - 2, dir == 0 ? g->dx : g->dy, lm, l1, la);
                } else if (dir == 2) {
                    mom_cross_z(g, o, lm, l1, l2, l3, l4, la, comp == 0 ? 1 : 2);
                } else {
                    mom_cross_h(o, n, 4, n - 2, dir == 0 ? g->dx : g->dy, lm, l1, l2, l3, l4, la);
                }
                for (int l = 1; l <= n; ++l) out[base + st * (size_t)(l - 1)] = L(la, l);
            }
        free(buf);
    }
    (void)sz;
}

void orc_qnadv_uvw(const orc_grid *g, const orc_opts *o, int ntime,
                   const double *mu, const double *mv, const double *mw,
                   const double *ut, const double *vt, const double *wt,
                   double *ua, double *va, double *wa)
{
    size_t ntot = (size_t)g->nx * g->ny * g->nz;
    int three_d = g->ny > 1;
    int even = (ntime % 2) == 0;
    int order_even[3] = {0, 1, 2}, order_odd[3] = {2, 1, 0};
    const int *ord = even ? order_even : order_odd;
    const double *m[3] = {mu, mv, mw};
    double *out[3] = {ua, va, wa};
    for (int c = 0; c < 3; ++c) {
        if (c == 1 && !three_d) continue;
        memset(out[c], 0, ntot * sizeof(double));
        for (int s = 0; s < 3; ++s) {
            int d = ord[s];
            if (d == 1 && !three_d) continue;
            mom_dir(g, o, c, d, m[c], ut, vt, wt, out[c]);
        }
    }
}

/* ======================================================================
 * Boundary fills: bc.f90
 * ====================================================================== */

void orc_fill_scalar(const orc_grid *g, double *x)
{
    int nx = g->nx, ny = g->ny, nz = g->nz;
    int three_d = ny > 1;
#define X3(i, j, k) x[IDX(g, i, j, k)]
    if (g->bcl[0] == ORC_BC_PER || g->bcl[1] == ORC_BC_PER)
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j) {
                X3(1, j, k) = X3(nx - 4, j, k);
                X3(2, j, k) = X3(nx - 3, j, k);
                X3(3, j, k) = X3(nx - 2, j, k);
                X3(nx, j, k) = X3(5, j, k);
                X3(nx - 1, j, k) = X3(4, j, k);
            }
    if (three_d && (g->bcl[2] == ORC_BC_PER || g->bcl[3] == ORC_BC_PER))
        for (int k = 1; k <= nz; ++k)
            for (int i = 1; i <= nx; ++i) {
                X3(i, 1, k) = X3(i, ny - 4, k);
                X3(i, 2, k) = X3(i, ny - 3, k);
                X3(i, 3, k) = X3(i, ny - 2, k);
                X3(i, ny, k) = X3(i, 5, k);
                X3(i, ny - 1, k) = X3(i, 4, k);
            }
    if (g->bcl[4] == ORC_BC_PER || g->bcl[5] == ORC_BC_PER)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                X3(i, j, 1) = X3(i, j, nz - 4);
                X3(i, j, 2) = X3(i, j, nz - 3);
                X3(i, j, 3) = X3(i, j, nz - 2);
                X3(i, j, nz) = X3(i, j, 5);
                X3(i, j, nz - 1) = X3(i, j, 4);
            }
    if (g->bcl[0] == ORC_BC_OPE || g->bcl[0] == ORC_BC_NNF)
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j) {
                X3(2, j, k) = X3(4, j, k);
                X3(1, j, k) = X3(5, j, k);
            }
    if (g->bcl[1] == ORC_BC_OPE || g->bcl[1] == ORC_BC_NNF)
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j) {
                X3(nx, j, k) = X3(nx - 4, j, k);
                X3(nx - 1, j, k) = X3(nx - 3, j, k);
            }
    if (three_d && (g->bcl[2] == ORC_BC_OPE || g->bcl[2] == ORC_BC_NNF))
        for (int k = 1; k <= nz; ++k)
            for (int i = 1; i <= nx; ++i) {
                X3(i, 2, k) = X3(i, 4, k);
                X3(i, 1, k) = X3(i, 5, k);
            }
    if (three_d && (g->bcl[3] == ORC_BC_OPE || g->bcl[3] == ORC_BC_NNF))
        for (int k = 1; k <= nz; ++k)
            for (int i = 1; i <= nx; ++i) {
                X3(i, ny, k) = X3(i, ny - 4, k);
                X3(i, ny - 1, k) = X3(i, ny - 3, k);
            }
#undef X3
}

/* lbcu_p4 (bc.f90:81-344), serial branch of a single-tile run (it_start =
 * jt_start = 4, no nesting, no CHANNEL): periodic copies of u, v, w in x,
 * y and z, then the no-flux ('nnf', negated) and open ('ope') fills of the
 * normal component on non-periodic sides.  The tangential fills of those
 * sides need it_start == 3 (jt_start == 3), which a single tile never has,
 * so v, w (u, w) keep their ghost values there, as in the reference. */
void orc_fill_winds(const orc_grid *g, double *u, double *v, double *w)
{
    int nx = g->nx, ny = g->ny, nz = g->nz;
    int three_d = ny > 1;
    double *f[3] = {u, three_d ? v : NULL, w};
    if (g->bcl[0] == ORC_BC_PER || g->bcl[1] == ORC_BC_PER)
        for (int c = 0; c < 3; ++c) {
            double *x = f[c];
            if (!x) continue;
            for (int k = 1; k <= nz; ++k)
                for (int j = 1; j <= ny; ++j) {
                    x[IDX(g, 1, j, k)] = x[IDX(g, nx - 4, j, k)];
                    x[IDX(g, 2, j, k)] = x[IDX(g, nx - 3, j, k)];
                    x[IDX(g, 3, j, k)] = x[IDX(g, nx - 2, j, k)];
                    x[IDX(g, nx, j, k)] = x[IDX(g, 5, j, k)];
                    x[IDX(g, nx - 1, j, k)] = x[IDX(g, 4, j, k)];
                }
        }
    if (three_d && (g->bcl[2] == ORC_BC_PER || g->bcl[3] == ORC_BC_PER))
        for (int c = 0; c < 3; ++c) {
            double *x = f[c];
            if (!x) continue;
            for (int k = 1; k <= nz; ++k)
                for (int i = 1; i <= nx; ++i) {
                    x[IDX(g, i, 1, k)] = x[IDX(g, i, ny - 4, k)];
                    x[IDX(g, i, 2, k)] = x[IDX(g, i, ny - 3, k)];
                    x[IDX(g, i, 3, k)] = x[IDX(g, i, ny - 2, k)];
                    x[IDX(g, i, ny, k)] = x[IDX(g, i, 5, k)];
                    x[IDX(g, i, ny - 1, k)] = x[IDX(g, i, 4, k)];
                }
        }
    if (g->bcl[4] == ORC_BC_PER || g->bcl[5] == ORC_BC_PER)
        for (int c = 0; c < 3; ++c) {
            double *x = f[c];
            if (!x) continue;
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    x[IDX(g, i, j, 1)] = x[IDX(g, i, j, nz - 4)];
                    x[IDX(g, i, j, 2)] = x[IDX(g, i, j, nz - 3)];
                    x[IDX(g, i, j, 3)] = x[IDX(g, i, j, nz - 2)];
                    x[IDX(g, i, j, nz)] = x[IDX(g, i, j, 5)];
                    x[IDX(g, i, j, nz - 1)] = x[IDX(g, i, j, 4)];
                }
        }
    /* west / east sides: u is the normal component (bc.f90:255-281) */
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j) {
            double s = g->bcl[0] == ORC_BC_NNF ? -1.0 : 1.0;
            if (g->bcl[0] == ORC_BC_NNF || g->bcl[0] == ORC_BC_OPE) {
                u[IDX(g, 3, j, k)] = s * u[IDX(g, 4, j, k)];
                u[IDX(g, 2, j, k)] = s * u[IDX(g, 5, j, k)];
                u[IDX(g, 1, j, k)] = s * u[IDX(g, 6, j, k)];
            }
            s = g->bcl[1] == ORC_BC_NNF ? -1.0 : 1.0;
            if (g->bcl[1] == ORC_BC_NNF || g->bcl[1] == ORC_BC_OPE) {
                u[IDX(g, nx, j, k)] = s * u[IDX(g, nx - 3, j, k)];
                u[IDX(g, nx - 1, j, k)] = s * u[IDX(g, nx - 2, j, k)];
            }
        }
    /* south / north sides: v is the normal component (bc.f90:307-333) */
    if (three_d)
        for (int k = 1; k <= nz; ++k)
            for (int i = 1; i <= nx; ++i) {
                double s = g->bcl[2] == ORC_BC_NNF ? -1.0 : 1.0;
                if (g->bcl[2] == ORC_BC_NNF || g->bcl[2] == ORC_BC_OPE) {
                    v[IDX(g, i, 3, k)] = s * v[IDX(g, i, 4, k)];
                    v[IDX(g, i, 2, k)] = s * v[IDX(g, i, 5, k)];
                    v[IDX(g, i, 1, k)] = s * v[IDX(g, i, 6, k)];
                }
                s = g->bcl[3] == ORC_BC_NNF ? -1.0 : 1.0;
                if (g->bcl[3] == ORC_BC_NNF || g->bcl[3] == ORC_BC_OPE) {
                    v[IDX(g, i, ny, k)] = s * v[IDX(g, i, ny - 3, k)];
                    v[IDX(g, i, ny - 1, k)] = s * v[IDX(g, i, ny - 2, k)];
                }
            }
}

/* ======================================================================
 * Stage combine (integpack.f90:137-227 with integ.h) and totalav
 * ====================================================================== */

void orc_integ_scal(long n, int sorder, int im, double dt, const double *xold,
                    const double *a, double *tend, double *x)
{
    double a_lm, b_lm;
    if (sorder > 1) {
        if (im == 1) { a_lm = 1.0; b_lm = 0.0; }
        else { a_lm = 0.5; b_lm = 1.0; }
        for (long i = 0; i < n; ++i) {
            tend[i] = a[i] + b_lm * tend[i];
            x[i] = xold[i] + a_lm * dt * tend[i];
        }
    } else {
        for (long i = 0; i < n; ++i) x[i] = xold[i] + dt * a[i];
    }
}

/* integ_mom (integpack.f90:54-135) with integ.h: the momentum stage of the
 * winds, old/x/a/tend = (u, v, w) triples of whole arrays (v unused in 2D).
 * The implicit sponge (sponge == 2) follows the reference literally: the
 * loop over levels at or above zdamp assigns the WHOLE tauw array (tauw =
 * taud*clamp(cos(pi/2 (zmax - z0(k))/H)^2), pi = 3.14159265, taud = 0.25),
 * so the last such level sets one factor 1/(1 + a_lm dt tauw) for every
 * level, and with no level above zdamp tauw stays 1 before that step. */
double orc_sponge_factor(int nz, int sorder, int im, double dt, const double *z0, double zmax, double zdamp,
                         int sponge)
{
    const double pi = 3.14159265, taud = 0.25;
    double a_lm = (sorder > 1 && im != 1) ? 0.5 : 1.0;
    double tauw = 1.0;
    if (sponge == 2) {
        double H = zmax - zdamp;
        for (int k = 1; k <= nz; ++k)
            if (L(z0, k) >= zdamp) {
                double c = cos(0.5 * pi * (zmax - L(z0, k)) / H);
                tauw = taud * dmax(dmin(pow(c, 2.0), 1.0), 0.0);
            }
        tauw = 1.0 / (1.0 + a_lm * dt * tauw);
    }
    return tauw;
}

void orc_integ_mom(const orc_grid *g, int sorder, int im, double dt, const double *z0, double zmax, double zdamp,
                   int sponge, double *const *old, const double *const *a, double *const *tend, double *const *x)
{
    size_t n = (size_t)g->nx * g->ny * g->nz, plane = (size_t)g->nx * g->ny;
    double a_lm, b_lm;
    if (sorder > 1) {
        if (im == 1) { a_lm = 1.0; b_lm = 0.0; }
        else { a_lm = 0.5; b_lm = 1.0; }
    } else {
        a_lm = 1.0;
        b_lm = 0.0;
    }
    double tw = orc_sponge_factor(g->nz, sorder, im, dt, z0, zmax, zdamp, sponge);
    int three_d = g->ny > 1;
    for (int c = 0; c < 3; ++c) {
        if (c == 1 && !three_d) continue;
        if (sorder > 1)
            for (size_t i = 0; i < n; ++i) tend[c][i] = a[c][i] + b_lm * tend[c][i];
        for (size_t i = 0; i < n; ++i) {
            double upd = sorder > 1 ? old[c][i] + a_lm * dt * tend[c][i] : old[c][i] + dt * a[c][i];
            x[c][i] = c == 2 ? tw * upd : upd;
        }
    }
    (void)plane;
}

double orc_totalav(const orc_grid *g, const double *x)
{
    int three_d = g->ny > 1;
    int jt0 = three_d ? 4 : 1, jt1 = three_d ? g->ny - 2 : 1;
    double s = 0.0;
    long n = 0;
    for (int k = 1; k <= g->nz; ++k)
        for (int j = jt0; j <= jt1; ++j)
            for (int i = 4; i <= g->nx - 2; ++i) {
                s = s + x[IDX(g, i, j, k)];
                ++n;
            }
    return s / (double)n;
}

/* ======================================================================
 * Tangent-linear model of the reference WENO operator (flux form),
 * forward mode with the branch decisions of the value computation.
 * One-sided choices at the kinks (documented conventions shared with the
 * GPU adjoint):
 *   - min() of the limiter candidates: derivative of the candidate the
 *     value computation selects (ties: the later argument, as dmin);
 *   - phi = min(|r|, 1): the clamped branch (|r| >= 1) has zero derivative;
 *     |r| uses sign(r) with sign(0) = +1;
 *   - the ratio-20 weight switch: derivative of the branch taken;
 *   - the face velocity kink at u = 0: the u >= 0 side (d|u| = du).
 * ====================================================================== */

typedef struct { double v, d; } dual;

static dual dd_(double v, double d) { dual r = {v, d}; return r; }
static dual dadd(dual a, dual b) { return dd_(a.v + b.v, a.d + b.d); }
static dual dsub(dual a, dual b) { return dd_(a.v - b.v, a.d - b.d); }
static dual dmul(dual a, dual b) { return dd_(a.v * b.v, a.d * b.v + a.v * b.d); }
static dual dscale(double s, dual a) { return dd_(s * a.v, s * a.d); }
static dual ddiv(dual a, dual b) { double q = a.v / b.v; return dd_(q, (a.d - q * b.d) / b.v); }
static dual dabs_(dual a) { return a.v < 0.0 ? dd_(-a.v, -a.d) : a; }
static dual dminl(dual a, dual b) { return a.v < b.v ? a : b; }
static dual dsq(dual a) { return dmul(a, a); }

static void tl_weights(dual a, dual b, dual c, dual d, dual ww[4])
{
    const double g1 = 1.0 / 3.0, g2 = 2.0 / 3.0, eps = 1.e-15, thres = 20.0;
    dual e1 = dsq(dsub(b, a)), e2 = dsq(dsub(c, b)), e3 = dsq(dsub(d, c));
    dual epsd = dd_(eps, 0.0);
    ww[0] = dd_(g1, 0); ww[1] = dd_(g2, 0); ww[2] = dd_(g2, 0); ww[3] = dd_(g1, 0);
    if (dmax(e1.v, e2.v) / (dmin(e1.v, e2.v) + eps) > thres) {
        dual w1 = ddiv(dd_(g1, 0), dsq(dadd(epsd, e1)));
        dual w2 = ddiv(dd_(g2, 0), dsq(dadd(epsd, e2)));
        dual s = dadd(w1, w2);
        ww[0] = ddiv(w1, s);
        ww[1] = ddiv(w2, s);
    }
    if (dmax(e2.v, e3.v) / (dmin(e2.v, e3.v) + eps) > thres) {
        dual w3 = ddiv(dd_(g2, 0), dsq(dadd(epsd, e2)));
        dual w4 = ddiv(dd_(g1, 0), dsq(dadd(epsd, e3)));
        dual s = dadd(w3, w4);
        ww[2] = ddiv(w3, s);
        ww[3] = ddiv(w4, s);
    }
}

/* face flux and its derivative; kind: 0 loop body, 1 bottom closure,
 * 2 top closure; nonlin: weights switch allowed */
static dual tl_face(dual a, dual b, dual c, dual d, dual u, int kind, int nonlin)
{
    const double eps = 1.e-15, g1 = 1.0 / 3.0, g2 = 2.0 / 3.0;
    dual ffim1 = dadd(dscale(-0.5, a), dscale(1.5, b));
    dual ffip1 = dadd(dscale(0.5, b), dscale(0.5, c));
    dual ffip2 = dsub(dscale(1.5, c), dscale(0.5, d));
    dual fl, fr;
    if (kind == 2) {
        fl = dadd(dscale(g1, ffim1), dscale(g2, ffip1));
        fr = ffip1;
    } else if (kind == 1) {
        fl = ffip1;
        fr = dadd(dscale(g2, ffip1), dscale(g1, ffip2));
    } else {
        dual ww[4] = {dd_(g1, 0), dd_(g2, 0), dd_(g2, 0), dd_(g1, 0)};
        if (nonlin) tl_weights(a, b, c, d, ww);
        dual minp = dminl(dminl(dminl(ffim1, ffip1), ffip1), ffip2);
        dual r = ddiv(dsub(dd_(0, 0), b), dadd(dsub(minp, b), dd_(eps, 0)));
        dual phi = dabs_(r);
        if (!(phi.v < 1.0)) phi = dd_(1.0, 0.0);
        dual f1 = dadd(dmul(phi, dsub(ffim1, b)), b);
        dual f2 = dadd(dmul(phi, dsub(ffip1, b)), b);
        dual f4 = dadd(dmul(phi, dsub(ffip2, b)), b);
        fl = dadd(dmul(ww[0], f1), dmul(ww[1], f2));
        fr = dadd(dmul(ww[2], f2), dmul(ww[3], f4));
    }
    /* 0.5 fl (u + |u|) + 0.5 fr (u - |u|) with d|u| = du at u = 0 */
    dual au = u.v < 0.0 ? dd_(-u.v, -u.d) : u;
    return dadd(dmul(dscale(0.5, fl), dadd(u, au)), dmul(dscale(0.5, fr), dsub(u, au)));
}

/* weno5_left in dual numbers, same operation order for the value (the
 * WENO5-JS extension: no reference counterpart). */
static dual tl_weno5_left(dual a, dual b, dual c, dual d, dual e)
{
    const double eps = 1e-6;
    dual q0 = dscale(1.0 / 6.0, dadd(dsub(dscale(2.0, a), dscale(7.0, b)), dscale(11.0, c)));
    dual q1 = dscale(1.0 / 6.0, dadd(dadd(dscale(-1.0, b), dscale(5.0, c)), dscale(2.0, d)));
    dual q2 = dscale(1.0 / 6.0, dsub(dadd(dscale(2.0, c), dscale(5.0, d)), e));
    dual s0 = dadd(dsub(a, dscale(2.0, b)), c), t0 = dadd(dsub(a, dscale(4.0, b)), dscale(3.0, c));
    dual s1 = dadd(dsub(b, dscale(2.0, c)), d), t1 = dsub(b, d);
    dual s2 = dadd(dsub(c, dscale(2.0, d)), e), t2 = dadd(dsub(dscale(3.0, c), dscale(4.0, d)), e);
    dual b0 = dadd(dscale(13.0 / 12.0, dsq(s0)), dscale(0.25, dsq(t0)));
    dual b1 = dadd(dscale(13.0 / 12.0, dsq(s1)), dscale(0.25, dsq(t1)));
    dual b2 = dadd(dscale(13.0 / 12.0, dsq(s2)), dscale(0.25, dsq(t2)));
    dual epsd = dd_(eps, 0.0);
    dual a0 = ddiv(dd_(0.1, 0.0), dsq(dadd(epsd, b0)));
    dual a1 = ddiv(dd_(0.6, 0.0), dsq(dadd(epsd, b1)));
    dual a2 = ddiv(dd_(0.3, 0.0), dsq(dadd(epsd, b2)));
    return ddiv(dadd(dadd(dmul(a0, q0), dmul(a1, q1)), dmul(a2, q2)), dadd(dadd(a0, a1), a2));
}

/* WENO5 face flux 0.5 fl (u + |u|) + 0.5 fr (u - |u|), d|u| = du at u = 0 */
static dual tl_weno5_face(dual a, dual b, dual c, dual d, dual e, dual f, dual u)
{
    dual fl = tl_weno5_left(a, b, c, d, e);
    dual fr = tl_weno5_left(f, e, d, c, b);
    dual au = u.v < 0.0 ? dd_(-u.v, -u.d) : u;
    return dadd(dmul(dscale(0.5, fl), dadd(u, au)), dmul(dscale(0.5, fr), dsub(u, au)));
}

/* Tangent-linear tendency of 'weno' (weno5 = 0) or of the WENO5-JS
 * extension (weno5 = 1: x/y faces WENO5 with the periodic image x(6) for
 * x(n+1), z faces WENO5 where the upwind stencil lies in the column and
 * the reference face elsewhere, as weno5_h / weno3_z(weno5 = 1)). */
static void tlm_tendency(const orc_grid *g, int weno5, const double *q, const double *dq, const double *u,
                         const double *v, const double *w, const double *du, const double *dv, const double *dw,
                         double *xs, double *dxs)
{
    int nx = g->nx, ny = g->ny, nz = g->nz;
    int three_d = ny > 1;
    int jt0 = three_d ? 4 : 1, jt1 = three_d ? ny - 2 : 1;
    double dx1 = 1.0 / g->dx, dy1 = 1.0 / g->dy;
#define DQ(i, j, k) dd_(q[IDX(g, i, j, k)], dq[IDX(g, i, j, k)])
    /* x(n+1) of the last face for negative velocity: its periodic image x(6) */
#define PX(i) ((i) == nx + 1 ? 6 : (i))
#define PY(j) ((j) == ny + 1 ? 6 : (j))
#pragma omp parallel for collapse(2) num_threads(g_threads) schedule(static)
    for (int k = 1; k <= nz; ++k)
        for (int j = jt0; j <= jt1; ++j)
            for (int i = 4; i <= nx - 2; ++i) {
                size_t id = IDX(g, i, j, k);
                dual t = dd_(0.0, 0.0);
                /* x */
                if (weno5) {
                    dual fp = tl_weno5_face(DQ(i - 2, j, k), DQ(i - 1, j, k), DQ(i, j, k), DQ(i + 1, j, k),
                                            DQ(i + 2, j, k), DQ(PX(i + 3), j, k),
                                            dd_(u[IDX(g, i + 1, j, k)], du[IDX(g, i + 1, j, k)]));
                    dual fm = tl_weno5_face(DQ(i - 3, j, k), DQ(i - 2, j, k), DQ(i - 1, j, k), DQ(i, j, k),
                                            DQ(i + 1, j, k), DQ(i + 2, j, k), dd_(u[id], du[id]));
                    t = dsub(t, dscale(dx1, dsub(fp, fm)));
                } else {
                    dual fp = tl_face(DQ(i - 1, j, k), DQ(i, j, k), DQ(i + 1, j, k), DQ(i + 2, j, k),
                                      dd_(u[IDX(g, i + 1, j, k)], du[IDX(g, i + 1, j, k)]), 0, 1);
                    dual fm = tl_face(DQ(i - 2, j, k), DQ(i - 1, j, k), DQ(i, j, k), DQ(i + 1, j, k),
                                      dd_(u[id], du[id]), 0, 1);
                    t = dsub(t, dscale(dx1, dsub(fp, fm)));
                }
                if (three_d && weno5) {
                    dual fp = tl_weno5_face(DQ(i, j - 2, k), DQ(i, j - 1, k), DQ(i, j, k), DQ(i, j + 1, k),
                                            DQ(i, j + 2, k), DQ(i, PY(j + 3), k),
                                            dd_(v[IDX(g, i, j + 1, k)], dv[IDX(g, i, j + 1, k)]));
                    dual fm = tl_weno5_face(DQ(i, j - 3, k), DQ(i, j - 2, k), DQ(i, j - 1, k), DQ(i, j, k),
                                            DQ(i, j + 1, k), DQ(i, j + 2, k), dd_(v[id], dv[id]));
                    t = dsub(t, dscale(dy1, dsub(fp, fm)));
                } else if (three_d) {
                    dual fp = tl_face(DQ(i, j - 1, k), DQ(i, j, k), DQ(i, j + 1, k), DQ(i, j + 2, k),
                                      dd_(v[IDX(g, i, j + 1, k)], dv[IDX(g, i, j + 1, k)]), 0, 1);
                    dual fm = tl_face(DQ(i, j - 2, k), DQ(i, j - 1, k), DQ(i, j, k), DQ(i, j + 1, k),
                                      dd_(v[id], dv[id]), 0, 1);
                    t = dsub(t, dscale(dy1, dsub(fp, fm)));
                }
                /* z: faces f = 1..nz-1 with the closures; F(0) = F(nz) = 0 (nnf) */
                {
                    dual fz[2];
                    for (int s = 0; s < 2; ++s) {
                        int f = k - 1 + s; /* face f+1/2 */
                        if (f < 1 || f > nz - 1) { fz[s] = dd_(0.0, 0.0); continue; }
                        dual wf = dd_(w[IDX(g, i, j, f + 1)], dw[IDX(g, i, j, f + 1)]);
                        if (weno5 && wf.v > 0.0 && f >= 3 && f <= nz - 2) {
                            dual fl = tl_weno5_left(DQ(i, j, f - 2), DQ(i, j, f - 1), DQ(i, j, f), DQ(i, j, f + 1),
                                                    DQ(i, j, f + 2));
                            fz[s] = dmul(dscale(0.5, fl), dadd(wf, wf));
                        } else if (weno5 && wf.v < 0.0 && f >= 2 && f <= nz - 3) {
                            dual fr = tl_weno5_left(DQ(i, j, f + 3), DQ(i, j, f + 2), DQ(i, j, f + 1), DQ(i, j, f),
                                                    DQ(i, j, f - 1));
                            fz[s] = dmul(dscale(0.5, fr), dsub(wf, dd_(-wf.v, -wf.d)));
                        } else if (f == nz - 1) {
                            fz[s] = tl_face(DQ(i, j, f - 1), DQ(i, j, f), DQ(i, j, f + 1), dd_(0, 0), wf, 2, 0);
                        } else {
                            dual a = f == 1 ? DQ(i, j, 1) : DQ(i, j, f - 1);
                            fz[s] = tl_face(a, DQ(i, j, f), DQ(i, j, f + 1), DQ(i, j, f + 2), wf, 0,
                                            f >= 3 && f <= nz - 3);
                        }
                    }
                    t = dsub(t, dscale(L(g->dz1w, k), dsub(fz[1], fz[0])));
                }
                xs[id] = t.v;
                dxs[id] = t.d;
            }
#undef DQ
#undef PX
#undef PY
}

void orc_tlm_weno_tendency(const orc_grid *g, const double *q, const double *dq, const double *u,
                           const double *v, const double *w, const double *du, const double *dv,
                           const double *dw, double *xs, double *dxs)
{
    tlm_tendency(g, 0, q, dq, u, v, w, du, dv, dw, xs, dxs);
}

void orc_tlm_weno5_tendency(const orc_grid *g, const double *q, const double *dq, const double *u,
                            const double *v, const double *w, const double *du, const double *dv,
                            const double *dw, double *xs, double *dxs)
{
    tlm_tendency(g, 1, q, dq, u, v, w, du, dv, dw, xs, dxs);
}

/* ======================================================================
 * Density weighting, anelastic build (funcpack.f90:65-542).
 * ====================================================================== */

void orc_avden0(int nz, const double *fdz0, const double *den0, double *avden0)
{
    /* initial.f90:543-546 */
    for (int k = 2; k <= nz; ++k)
        avden0[k - 1] = (fdz0[k - 1] * den0[k - 1] + fdz0[k - 2] * den0[k - 2]) / (fdz0[k - 1] + fdz0[k - 2]);
    avden0[0] = den0[0];
}

static void scale_levels(const orc_grid *g, const double *prof, const double *in, double *out, int inverse,
                         double shift)
{
    size_t plane = (size_t)g->nx * g->ny;
    for (int k = 1; k <= g->nz; ++k) {
        double p = prof[k - 1];
        const double *a = in + (size_t)(k - 1) * plane;
        double *b = out + (size_t)(k - 1) * plane;
        for (size_t n = 0; n < plane; ++n) b[n] = (inverse ? a[n] / p : a[n] * p) - shift;
    }
}

void orc_get_momentum(const orc_grid *g, const double *den0, const double *avden0, double *u, double *v,
                      double *w, int inverse)
{
    /* funcpack.f90:160-168 (momentum), :294-302 (velocities) */
    scale_levels(g, den0, u, u, inverse, 0.0);
    if (g->ny > 1) scale_levels(g, den0, v, v, inverse, 0.0);
    scale_levels(g, avden0, w, w, inverse, 0.0);
}

void orc_get_advective(const orc_grid *g, const double *den0, const double *avden0, double u0shift,
                       double v0shift, const double *u, const double *v, const double *w, double *ua,
                       double *va, double *wa)
{
    /* funcpack.f90:76-86: copy, get_velocities, subtract the translation */
    scale_levels(g, den0, u, ua, 1, u0shift);
    if (g->ny > 1) scale_levels(g, den0, v, va, 1, v0shift);
    scale_levels(g, avden0, w, wa, 1, 0.0);
}

void orc_get_conserved(const orc_grid *g, const double *dens, double *q, int inverse)
{
    /* funcpack.f90:413-416 (x*dref), :473-476 (x/dref) */
    size_t n = (size_t)g->nx * g->ny * g->nz;
    for (size_t i = 0; i < n; ++i) q[i] = inverse ? q[i] / dens[i] : q[i] * dens[i];
}
