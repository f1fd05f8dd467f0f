// schemes_fp64.hpp -- per-face and per-cell device formulas of the MIMICA
// advection schemes, FP64, in the reference's floating-point operation order.
//
// The reference computes each direction as a sequential line sweep
// (face fluxes into a line array, then differences).  Here every formula
// is a pure function of a face or cell index and accessors Q(i), U(i)
// (1-based line indices), so a thread can evaluate one face (the marching
// kernels) or one cell (the PPM kernel) without the line sweep.  Position-dependent closures (tile
// edges, bottom/top levels) are selected by index exactly where the
// reference's loops switch formulas.  Built without FMA contraction these
// reproduce the reference arithmetic bit for bit.
#pragma once
#include <hip/hip_runtime.h>

namespace mimadv {

struct ZMetrics {
    const double *fdz0, *dzw, *dz1w, *dz1;
    const double *ppmc;  // [nz][12] PPM non-uniform coefficients per level
    const float *momz;   // [nz][MOMZ_N] level coefficients of the FP32 batched momentum operator
    double dz;
};

struct SchemeParams {
    int scheme;
    int muscl_limiter;
    int ppm_limiter;
    int weno_bottom;
    double lim_tol;
    double dt0;
};

__device__ __forceinline__ double dmaxd(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmind(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double sgn1(double x) { return copysign(1.0, x); }

// ---------------------------------------------------------------------------
// WENO5-JS (extension, no reference counterpart): Jiang & Shu (1996) left
// face value from a..e = x(f-2..f+2); eps = 1e-6, p = 2
// ---------------------------------------------------------------------------
__device__ __forceinline__ double weno5_left(double a, double b, double c, double d, double e)
{
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

// ---------------------------------------------------------------------------
// MUSCL (advq_muscl.f90)
// ---------------------------------------------------------------------------

// limiter_muscl (advq_muscl.f90:335-383)
__device__ __forceinline__ void muscl_limiter(int lim, double lt, double xx, double ffip1, double ffip2,
                                              double ffim1, double &phir, double &phil)
{
    double rp1 = 1.0, rp2 = 1.0;
    if (fabs(ffim1) > lt * fabs(xx) && fabs(ffip1) < fabs(ffim1) / lt) rp1 = ffip1 / ffim1;
    if (fabs(ffip2) > lt * fabs(xx) && fabs(ffip1) < fabs(ffip2) / lt) rp2 = ffip1 / ffip2;
    switch (lim) {
    case 2:  // VL
        phil = (rp1 + fabs(rp1)) / (1.0 + fabs(rp1));
        phir = (rp2 + fabs(rp2)) / (1.0 + fabs(rp2));
        break;
    case 1:  // MC
        phil = dmaxd(0.0, dmind(dmind(2.0 * rp1, 0.5 * (rp1 + 1.0)), 2.0));
        phir = dmaxd(0.0, dmind(dmind(2.0 * rp2, 0.5 * (rp2 + 1.0)), 2.0));
        break;
    case 0:  // minmod
        phil = dmaxd(0.0, dmind(1.0, rp1));
        phir = dmaxd(0.0, dmind(1.0, rp2));
        break;
    default:  // Fromm
        phil = 0.5 * (rp1 + 1.0);
        phir = 0.5 * (rp2 + 1.0);
        break;
    }
}

__device__ __forceinline__ double muscl_upwind(double uu, double fl, double fr)
{
    return 0.5 * uu * (fl * (1.0 + sgn1(uu)) + fr * (1.0 - sgn1(uu)));
}

// face flux i+1/2, x/y (advq_muscl.f90:81-100)
template <class QF, class UF>
__device__ __forceinline__ double muscl_hface(int i, double d, double d1, const SchemeParams &sp, QF Q, UF U)
{
    const double eps = 1.e-18;
    double uu = U(i + 1), xx = Q(i);
    double xp = Q(i + 1);
    double slope_p = (xp - xx) * d1;
    double slope_p2 = (Q(i + 2) - xp) * d1;
    double slope_m = (xx - Q(i - 1)) * d1;
    double lt = sp.lim_tol;
    double fl, fr;
    if (d * fabs(slope_m) > lt * fabs(xx) + eps || d * fabs(slope_p2) > lt * fabs(xx) + eps) {
        double phir, phil;
        muscl_limiter(sp.muscl_limiter, lt, xx, slope_p, slope_p2, slope_m, phir, phil);
        fr = xp - 0.5 * phir * slope_p2 * d;
        fl = xx + 0.5 * phil * slope_m * d;
    } else {
        fr = xp - 0.5 * slope_p * d;
        fl = xx + 0.5 * slope_p * d;
    }
    return muscl_upwind(uu, fl, fr);
}

// z face flux k+1/2 (advsz_muscl, advq_muscl.f90:245-302)
template <class QF, class WF>
__device__ __forceinline__ double muscl_zface(int f, int nz, const ZMetrics &zm, const SchemeParams &sp,
                                              double bc51, double bc61, QF Q, WF W)
{
    const double eps = 1.e-18;
    double lt = sp.lim_tol, dz = zm.dz;
    const double *dz1 = zm.dz1, *dz1w = zm.dz1w, *fdz0 = zm.fdz0;
    double ww = W(f + 1);
    double slope_p, slope_p2, slope_m, phir, phil, fl, fr;
    double m = dz1w[f - 1];
    if (f == 1) {
        slope_p = (Q(2) - Q(1)) * dz1[0];
        slope_p2 = (Q(3) - Q(2)) * dz1[1];
        slope_m = bc51 * (Q(nz - 4) - Q(nz - 5)) * dz1[nz - 6] + (1.0 - bc51) * slope_p;
        if (fdz0[1] * dz * fabs(slope_p2) > lt * fabs(Q(1)) + eps) {
            muscl_limiter(sp.muscl_limiter, lt, Q(1), slope_p, slope_p2, slope_m, phir, phil);
            fl = Q(1) + 0.5 * phil * slope_m / m;
            fr = Q(2) - 0.5 * phir * slope_p2 / m;
        } else {
            fl = Q(1) + 0.5 * slope_p / m;
            fr = Q(2) - 0.5 * slope_p / m;
        }
        return muscl_upwind(ww, fl, fr);
    }
    double xk = Q(f), xp = Q(f + 1);
    slope_p = (xp - xk) * dz1[f - 1];
    slope_m = (xk - Q(f - 1)) * dz1[f - 2];
    bool limited;
    if (f == nz - 1) {
        slope_p2 = bc61 * (Q(6) - Q(5)) * dz1[4] + (1.0 - bc61) * slope_p;
        limited = fdz0[f - 2] * dz * fabs(slope_m) > lt * fabs(xk) + eps;
    } else {
        slope_p2 = (Q(f + 2) - xp) * dz1[f];
        limited = fdz0[f - 2] * dz * fabs(slope_m) > lt * fabs(xk) + eps ||
                  fdz0[f] * dz * fabs(slope_p2) > lt * fabs(xk) + eps;
    }
    if (limited) {
        muscl_limiter(sp.muscl_limiter, lt, xk, slope_p, slope_p2, slope_m, phir, phil);
        fr = xp - 0.5 * phir * slope_p2 / m;
        fl = xk + 0.5 * phil * slope_m / m;
    } else {
        fr = xp - 0.5 * slope_p / m;
        fl = xk + 0.5 * slope_p / m;
    }
    return muscl_upwind(ww, fl, fr);
}

// ---------------------------------------------------------------------------
// QUICK (advq_quick.f90) and its unlimited form (advq_quick_nolim.f90)
// ---------------------------------------------------------------------------

// limiter_quick (advq_quick.f90:313-352); phic: x in x/y, x/dz in z
__device__ __forceinline__ void quick_limiter(double lt, double phic, double ffip1, double ffip2, double ffim1,
                                              double &phil, double &phir)
{
    const double eps = 1.e-18, c1 = 1.0 / 3.0;
    double rp1 = 1.0, rp2 = 1.0, rp11 = 1.0, rp21 = 1.0;
    if (fabs(ffim1) > lt * fabs(phic) + eps) rp1 = ffip1 / ffim1;
    if (fabs(ffip2) > lt * fabs(phic) + eps) rp2 = ffip1 / ffip2;
    if (fabs(ffip1) > lt * fabs(phic) + eps) {
        rp11 = ffim1 / ffip1;
        rp21 = ffip2 / ffip1;
    }
    phil = rp1 * (1.0 - c1 * (1.0 - rp11));
    phir = rp2 * (1.0 - c1 * (1.0 - rp21));
    if (fabs(ffim1) > lt * fabs(phic) + eps || fabs(ffip2) > lt * fabs(phic) + eps) {
        phil = dmaxd(0.0, dmind(2.0 * rp1, dmind(phil, 2.0)));
        phir = dmaxd(0.0, dmind(2.0 * rp2, dmind(phir, 2.0)));
    }
}

__device__ __forceinline__ double half_upwind(double uu, double fl, double fr)
{
    return 0.5 * fl * (uu + fabs(uu)) + 0.5 * fr * (uu - fabs(uu));
}

// face flux i+1/2, x/y (advq_quick.f90:75-88, advq_quick_nolim.f90:77-90)
template <bool NL, class QF, class UF>
__device__ __forceinline__ double quick_hface(int i, const SchemeParams &sp, QF Q, UF U)
{
    double uu = U(i + 1), xx = Q(i), xp = Q(i + 1);
    double ffip1 = xp - xx;
    double ffip2 = Q(i + 2) - xp;
    double ffim1 = xx - Q(i - 1);
    double phil, phir, fl, fr;
    if constexpr (NL) {
        const double cc = 1.0 / 4.0;
        phil = ffip1 - cc * (ffip1 - ffim1);
        phir = ffip1 + cc * (ffip2 - ffip1);
        fl = xx + 0.5 * phil;
        fr = xp - 0.5 * phir;
    } else {
        quick_limiter(sp.lim_tol, xx, ffip1, ffip2, ffim1, phil, phir);
        fl = xx + 0.5 * phil * ffim1;
        fr = xp - 0.5 * phir * ffip2;
    }
    return half_upwind(uu, fl, fr);
}

// bottom/core/top slopes of a z face (per metre), with the periodic-floor
// and periodic-lid substitutions of advsz_quick/advsz_lw
template <class QF>
__device__ __forceinline__ void z_slopes(int f, int nz, const ZMetrics &zm, double bc51, double bc61, QF Q,
                                         double ffip1, double &ffip2, double &ffim1)
{
    const double *dz1 = zm.dz1;
    if (f == 1) {
        ffip2 = (Q(3) - Q(2)) * dz1[1];
        ffim1 = bc51 * (Q(nz - 4) - Q(nz - 5)) * dz1[nz - 6] + (1.0 - bc51) * ffip1;
    } else if (f == nz - 1) {
        ffim1 = (Q(f) - Q(f - 1)) * dz1[f - 2];
        ffip2 = bc61 * (Q(6) - Q(5)) * dz1[4] + (1.0 - bc61) * ffip1;
    } else {
        ffip2 = (Q(f + 2) - Q(f + 1)) * dz1[f];
        ffim1 = (Q(f) - Q(f - 1)) * dz1[f - 2];
    }
}

// z face flux k+1/2 (advq_quick.f90:233-284, advq_quick_nolim.f90:239-286)
template <bool NL, class QF, class WF>
__device__ __forceinline__ double quick_zface(int f, int nz, const ZMetrics &zm, const SchemeParams &sp,
                                              double bc51, double bc61, QF Q, WF W)
{
    double xx = Q(f), ww = W(f + 1), xp = Q(f + 1);
    double ffip1 = (xp - xx) * zm.dz1[f - 1];
    double ffip2, ffim1;
    z_slopes(f, nz, zm, bc51, bc61, Q, ffip1, ffip2, ffim1);
    const double m = zm.dz1w[f - 1];
    double phil, phir, fl, fr;
    if constexpr (NL) {
        const double cc = 1.0 / 4.0;
        phil = ffip1 - cc * (ffip1 - ffim1);
        phir = ffip1 + cc * (ffip2 - ffip1);
        fl = xx + 0.5 * phil / m;
        fr = xp - 0.5 * phir / m;
    } else {
        quick_limiter(sp.lim_tol, xx / zm.dz, ffip1, ffip2, ffim1, phil, phir);
        fl = xx + 0.5 * phil * ffim1 / m;
        fr = xp - 0.5 * phir * ffip2 / m;
    }
    return half_upwind(ww, fl, fr);
}

// ---------------------------------------------------------------------------
// Lax-Wendroff (advq_lw.f90, limiter 'MC': the module's lwlimiter, which
// nothing in the model changes) and its unlimited form (advq_lw_nolim.f90)
// ---------------------------------------------------------------------------

// limiter_lw, 'MC' (advq_lw.f90:349-373,389-391)
__device__ __forceinline__ void lw_limiter(double lt, double xx, double ffip1, double ffip2, double ffim1,
                                           double &phir, double &phil)
{
    const double eps = 1.e-18;
    double rp1 = 1.0, rp2 = 1.0;
    if (fabs(ffim1) > lt * fabs(xx) + eps) rp1 = ffip1 / ffim1;
    if (fabs(ffip2) > lt * fabs(xx) + eps) rp2 = ffip1 / ffip2;
    phil = dmaxd(dmind(dmind(2.0 * rp1, 0.5 * (rp1 + 1.0)), 2.0), 0.0);
    phir = dmaxd(dmind(dmind(2.0 * rp2, 0.5 * (rp2 + 1.0)), 2.0), 0.0);
}

__device__ __forceinline__ double lw_flux(double uu, double fl, double fr)
{
    return 0.5 * uu * (fl * (sgn1(uu) + 1.0) - fr * (sgn1(uu) - 1.0));
}

// face flux i+1/2, x/y (advq_lw.f90:81-98, advq_lw_nolim.f90:79-88);
// Courant number from the full step dt0
template <bool NL, class QF, class UF>
__device__ __forceinline__ double lw_hface(int i, double d1, const SchemeParams &sp, QF Q, UF U)
{
    const double eps = 1.e-18;
    double uu = U(i + 1), xx = Q(i), xp = Q(i + 1);
    double ffip1 = xp - xx;
    double mu = fabs(uu) * sp.dt0 * d1;
    double fl, fr;
    if constexpr (NL) {
        fl = xx + 0.5 * ffip1 * (1.0 - mu);
        fr = xp - 0.5 * ffip1 * (1.0 - mu);
        return half_upwind(uu, fl, fr);
    }
    double ffim1 = xx - Q(i - 1);
    double ffip2 = Q(i + 2) - xp;
    double lt = sp.lim_tol;
    if (fabs(ffim1) > lt * fabs(xx) + eps || fabs(ffip2) > lt * fabs(xx) + eps) {
        double phir, phil;
        lw_limiter(lt, xx, ffip1, ffip2, ffim1, phir, phil);
        fl = xx + 0.5 * phil * ffim1 * (1.0 - mu);
        fr = xp - 0.5 * phir * ffip2 * (1.0 - mu);
    } else {
        fl = xx + 0.5 * ffip1 * (1.0 - mu);
        fr = xp - 0.5 * ffip1 * (1.0 - mu);
    }
    return lw_flux(uu, fl, fr);
}

// z face flux k+1/2 (advq_lw.f90:245-316, advq_lw_nolim.f90:226-235)
template <bool NL, class QF, class WF>
__device__ __forceinline__ double lw_zface(int f, int nz, const ZMetrics &zm, const SchemeParams &sp, double bc51,
                                           double bc61, QF Q, WF W)
{
    const double eps = 1.e-18;
    double xx = Q(f), ww = W(f + 1), xp = Q(f + 1);
    double mu = fabs(ww) * sp.dt0 * zm.dz1[f - 1];
    double ffip1 = (xp - xx) * zm.dz1[f - 1];
    const double m = zm.dz1w[f - 1];
    double fl, fr;
    if constexpr (NL) {
        fl = xx + 0.5 * ffip1 * (1.0 - mu) / m;
        fr = xp - 0.5 * ffip1 * (1.0 - mu) / m;
        return half_upwind(ww, fl, fr);
    }
    double ffip2, ffim1;
    z_slopes(f, nz, zm, bc51, bc61, Q, ffip1, ffip2, ffim1);
    double lt = sp.lim_tol, dz = zm.dz;
    if (dz * fabs(ffim1) > lt * fabs(xx) + eps || dz * fabs(ffip2) > lt * fabs(xx) + eps) {
        double phir, phil;
        lw_limiter(lt, xx, ffip1, ffip2, ffim1, phir, phil);
        fl = xx + 0.5 * phil * ffim1 * (1.0 - mu) / m;
        fr = xp - 0.5 * phir * ffip2 * (1.0 - mu) / m;
    } else {
        fl = xx + 0.5 * ffip1 * (1.0 - mu) / m;
        fr = xp - 0.5 * ffip1 * (1.0 - mu) / m;
    }
    return lw_flux(ww, fl, fr);
}

// ---------------------------------------------------------------------------
// PPM (advq_ppm.f90)
// ---------------------------------------------------------------------------

// sqrt(spacing(x)) of Fortran for REAL(8): spacing = 2**max(ilogb(x)-52,-1022)
// (2**-1022 for x == 0); the square root of a power of two is exact or
// sqrt(2) times one.
__device__ __forceinline__ double sqrt_spacing(double x)
{
    int e = (x == 0.0) ? -1022 : ilogb(x) - 52;
    if (e < -1022) e = -1022;
    if ((e & 1) == 0) return ldexp(1.0, e / 2);
    return ldexp(1.4142135623730951, (e - 1) / 2);
}

struct PPMRec {
    double fr, fl;
};

// limiter_ppm (advq_ppm.f90:449-530) followed by the CW84 face averages
// Flux_r/Flux_l (advq_ppm.f90:93-94)
__device__ __forceinline__ PPMRec ppm_cell(int use_lim, double xx, double xp, double xm, double xplus,
                                           double xmin, double ffip1, double ffip2, double ffim1,
                                           double ffim2, double mup, double muc)
{
    const double eps = 1.e-15, thres = 10.0;
    double rp1 = 1.0;
    double phir = xplus, phil = xmin;
    double g1a = ffip1 - ffim1, g1b = ffip1 + ffim1;
    double g2a = ffip2 - ffip1, g2b = ffip2 + ffip1;
    double g3a = ffim1 - ffim2, g3b = ffim1 + ffim2;
    double gamma1 = 0.5 * (g1a * g1a + g1b * g1b);
    double gamma2 = 0.5 * (g2a * g2a + g2b * g2b);
    double gamma3 = 0.5 * (g3a * g3a + g3b * g3b);
    double lambda = dmaxd(dmaxd(gamma1, gamma2), gamma3) / (dmind(dmind(gamma1, gamma2), gamma3) + eps);
    double phi1 = phir - phil;
    double phi2 = 3.0 * (phil - 2.0 * xx + phir);
    if (use_lim && lambda > thres) {
        if (phir < 0.0 || ((phir - xx) * (xp - phir) < 0.0 &&
                           (ffim1 * ffip2 >= 0.0 || ffim1 * ffip1 <= 0.0 || ffip2 * ffip1 <= 0.0))) {
            phir = (fabs(phir - xx) < fabs(phir - xp)) ? xx : xp;
        }
        if (phil < 0.0 || ((xx - phil) * (phil - xm) < 0.0 &&
                           (ffim2 * ffip1 >= 0.0 || ffim2 * ffim1 <= 0.0 || ffip1 * ffim1 <= 0.0))) {
            phil = (fabs(phil - xx) < fabs(phil - xm)) ? xx : xm;
        }
        phi1 = phir - phil;
        phi2 = 3.0 * (phil - 2.0 * xx + phir);
        double slope = -0.5 * phi1 / (sgn1(phi2) * dmaxd(fabs(phi2), eps));
        if (fabs(ffim1) > sqrt_spacing(ffim1)) rp1 = ffip1 / ffim1;
        // (slope > -1/2 .or. slope < 1/2) holds for every non-NaN slope
        if (fabs(phi2) > eps && (slope > -0.5 || slope < 0.5)) {
            double lim = dmaxd(0.0, dmind(dmind(2.0, 0.5 * (rp1 + 1.0)), 2.0 * rp1));
            phir = xx + 0.5 * ffim1 * lim;
            phil = xx - 0.5 * ffim1 * lim;
            if ((xx - phil) * (phir - xx) < 0.0) {
                phil = xx;
                phir = xx;
            }
            phi1 = phir - phil;
            phi2 = 0.0;
        }
    }
    PPMRec r;
    r.fr = phir - 0.5 * mup * (phi1 + (1.0 - (2.0 / 3.0) * mup) * phi2);
    r.fl = phil + 0.5 * muc * (phi1 - (1.0 - (2.0 / 3.0) * muc) * phi2);
    return r;
}

// right-face value xplus of cell c on an x/y line (left boundary c = 3,
// core c = 4..n-2), advq_ppm.f90:89,110
template <class QF>
__device__ __forceinline__ double ppm_hxplus(int c, QF Q)
{
    if (c == 3) return Q(3) + 0.5 * (Q(4) - Q(3));
    double ffip1 = Q(c + 1) - Q(c), ffip2 = Q(c + 2) - Q(c + 1), ffim1 = Q(c) - Q(c - 1);
    double sl = 0.5 * (ffip1 + ffim1);
    double slp1 = 0.5 * (ffip2 + ffip1);
    return Q(c) + 0.5 * ffip1 - (1.0 / 6.0) * (slp1 - sl);
}

// Flux_r/Flux_l of cell c in 3..n-1 on an x/y line (advq_ppm.f90:79-134)
template <class QF, class UF>
__device__ __forceinline__ PPMRec ppm_hcell(int c, int n, double d1, const SchemeParams &sp, QF Q, UF U)
{
    double dt0 = sp.dt0;
    double mup = fabs(U(c + 1)) * dt0 * d1;
    double muc = fabs(U(c)) * dt0 * d1;
    double xx = Q(c), xp = Q(c + 1), xm = Q(c - 1);
    if (c == 3) {
        double ffip1 = xp - xx, ffip2 = Q(5) - xp, ffim1 = xx - xm;
        double xmin = 0.5 * (xp + xx);
        double xplus = xx + 0.5 * ffip1;
        return ppm_cell(sp.ppm_limiter, xx, xp, xm, xplus, xmin, ffip1, ffip2, ffim1, ffim1, mup, muc);
    }
    double xmin = ppm_hxplus(c - 1, Q);
    if (c == n - 1) {
        double ffip1 = xp - xx, ffim1 = xx - xm, ffim2 = xm - Q(c - 2);
        double xplus = xx + 0.5 * ffip1;
        return ppm_cell(sp.ppm_limiter, xx, xp, xm, xplus, xmin, ffip1, ffip1, ffim1, ffim2, mup, muc);
    }
    double ffip1 = xp - xx, ffip2 = Q(c + 2) - xp, ffim1 = xx - xm, ffim2 = xm - Q(c - 2);
    double sl = 0.5 * (ffip1 + ffim1);
    double slp1 = 0.5 * (ffip2 + ffip1);
    double xplus = xx + 0.5 * ffip1 - (1.0 / 6.0) * (slp1 - sl);
    return ppm_cell(sp.ppm_limiter, xx, xp, xm, xplus, xmin, ffip1, ffip2, ffim1, ffim2, mup, muc);
}

__device__ __forceinline__ double ppm_faceflux(double fr, double fl, double uf)
{
    return 0.5 * fr * (uf + fabs(uf)) + 0.5 * fl * (uf - fabs(uf));
}

// x/y contribution of cell i in 4..n-2 (advq_ppm.f90:136-163); bc_hi is the
// east/north boundary code (non-periodic edge fix-up :143-146)
template <class QF, class UF>
__device__ __forceinline__ double ppm_h(int i, int n, double d1, bool adv, int bc_hi, const SchemeParams &sp,
                                        QF Q, UF U)
{
    PPMRec cm = ppm_hcell(i - 1, n, d1, sp, Q, U);
    PPMRec c0 = ppm_hcell(i, n, d1, sp, Q, U);
    PPMRec cp = ppm_hcell(i + 1, n, d1, sp, Q, U);
    double fr_i = c0.fr, fl_ip = cp.fl;
    if (bc_hi != 0 && i == n - 2) {
        fr_i = c0.fl;      // Flux_r(ii2) = Flux_l(ii2)
        fl_ip = cm.fr;     // Flux_l(ii2+1) = Flux_r(ii2-1)
    }
    double flux_old = ppm_faceflux(cm.fr, c0.fl, U(i));
    double flux = ppm_faceflux(fr_i, fl_ip, U(i + 1));
    double t = (flux - flux_old) * d1;
    if (adv) t = t - Q(i) * (U(i + 1) - U(i)) * d1;
    return t;
}

// Per-level PPM coefficients (metric-only subexpressions of the non-uniform
// face formula, advq_ppm.f90:379-383, evaluated in the same order):
//  [0] delta/(dm1+delta+dp1)   [1] (2dm1+delta)/(dp1+delta)  [2] (delta+2dp1)/(dm1+delta)
//  [3] dp1/(delta+dp1+dp2)     [4] (2delta+dp1)/(dp2+dp1)    [5] (dp1+2dp2)/(delta+dp1)
//  [6] delta/(delta+dp1)       [7] 1/(dm1+delta+dp1+dp2)
//  [8] 2dp1*delta/(delta+dp1)*((dm1+delta)/(2delta+dp1)-(dp2+dp1)/(2dp1+delta))
//  [9] delta*(dm1+delta)/(2delta+dp1)   [10] dp1*(dp1+dp2)/(2dp1+delta)
template <class QF>
__device__ __forceinline__ double ppm_zxplus(int c, const ZMetrics &zm, QF Q)
{
    if (c <= 2) return Q(c) + 0.5 * (Q(c + 1) - Q(c));
    const double *p = zm.ppmc + 12 * (c - 1);
    double ffip1 = Q(c + 1) - Q(c), ffip2 = Q(c + 2) - Q(c + 1), ffim1 = Q(c) - Q(c - 1);
    double sl = p[0] * (p[1] * ffip1 + p[2] * ffim1);
    double slp1 = p[3] * (p[4] * ffip2 + p[5] * ffip1);
    return Q(c) + p[6] * ffip1 + p[7] * (p[8] * ffip1 - p[9] * slp1 + p[10] * sl);
}

// Flux_r/Flux_l of level c in 1..nz-1, advsz_ppm (advq_ppm.f90:326-409)
template <class QF, class WF>
__device__ __forceinline__ PPMRec ppm_zcell(int c, int nz, const ZMetrics &zm, const SchemeParams &sp,
                                            double bc51, double bc61, QF Q, WF W)
{
    double dt0 = sp.dt0;
    double dz1c = zm.dz1[c - 1];
    double mup = fabs(W(c + 1)) * dt0 * dz1c;
    double muc = fabs(W(c)) * dt0 * dz1c;
    double xx = Q(c), xp = Q(c + 1);
    double ffip1 = xp - xx;
    if (c == 1) {
        double ffip2 = Q(3) - xp;
        double ffim1 = bc51 * (Q(nz - 4) - Q(nz - 5));
        double ffim2 = bc51 * (Q(nz - 5) - Q(nz - 6));
        double xmin = xx - 0.5 * ffip1;
        double xplus = xx + 0.5 * ffip1;
        return ppm_cell(sp.ppm_limiter, xx, xp, xp, xplus, xmin, ffip1, ffip2, ffim1, ffim2, mup, muc);
    }
    double xm = Q(c - 1);
    double xmin = ppm_zxplus(c - 1, zm, Q);
    if (c == 2) {
        double ffip2 = Q(4) - xp;
        double ffim1 = xx - xm;
        double ffim2 = bc51 * (Q(nz - 4) - Q(nz - 5));
        double xplus = xx + 0.5 * ffip1;
        return ppm_cell(sp.ppm_limiter, xx, xp, xm, xplus, xmin, ffip1, ffip2, ffim1, ffim2, mup, muc);
    }
    double ffim1 = xx - xm, ffim2 = xm - Q(c - 2);
    if (c == nz - 1) {
        double ffip2 = bc61 * (Q(6) - Q(5)) + (1.0 - bc61) * ffip1;
        double xplus = xx + 0.5 * ffip1;
        return ppm_cell(sp.ppm_limiter, xx, xp, xm, xplus, xmin, ffip1, ffip2, ffim1, ffim2, mup, muc);
    }
    double ffip2 = Q(c + 2) - xp;
    double xplus = ppm_zxplus(c, zm, Q);
    return ppm_cell(sp.ppm_limiter, xx, xp, xm, xplus, xmin, ffip1, ffip2, ffim1, ffim2, mup, muc);
}

// Flux_l of the top level (advq_ppm.f90:411-414)
template <class QF, class WF>
__device__ __forceinline__ double ppm_ztop_fl(int nz, const ZMetrics &zm, const SchemeParams &sp, QF Q, WF W)
{
    double muc = fabs(W(nz)) * sp.dt0 * zm.dz1[nz - 1];
    double ffim1 = Q(nz) - Q(nz - 1);
    return Q(nz) + 0.5 * ffim1 * (1.0 - muc);
}

// ---------------------------------------------------------------------------
// z contributions: the tendency loop of advs z_* (face differences, adv
// term, nnf closures), same for all three schemes
// ---------------------------------------------------------------------------

template <class FF, class QF, class WF>
__device__ __forceinline__ double z_assemble(int k, int nz, bool adv, int bc5, int bc6, const ZMetrics &zm,
                                             FF face, QF Q, WF W)
{
    const double *dz1w = zm.dz1w;
    if (k >= 2 && k <= nz - 1) {
        double t = (face(k) - face(k - 1)) * dz1w[k - 1];
        if (adv) t = t - Q(k) * (W(k + 1) - W(k)) * dz1w[k - 1];
        return t;
    }
    double t = 0.0;
    if (k == 1 && bc5 == 1) {
        t = t + (face(1) - 0.0) * dz1w[0];
        if (adv) t = t - Q(1) * (W(2) - 0.0) * dz1w[0];
    }
    if (k == nz && bc6 == 1) {
        t = t + (0.0 - face(nz - 1)) * dz1w[nz - 1];
        if (adv) t = t - Q(nz) * (0.0 - W(nz)) * dz1w[nz - 1];
    }
    return t;
}

}  // namespace mimadv
