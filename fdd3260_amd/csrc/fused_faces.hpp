// fused_faces.hpp -- per-face formulas of the batched kernels (FP32/FP64):
// the reference WENO face (advq_weno.f90) and MUSCL face (advq_muscl.f90)
// in operation-reduced form, their vertical closures and the periodic
// index wrap.  Shared by the single-grid and multi-domain batched kernels.
#pragma once
#include "fused.hpp"

namespace mimadv {
namespace faces {

template <class T>
__device__ __forceinline__ T rcp(T x);
template <>
__device__ __forceinline__ float rcp<float>(float x)
{
    return __builtin_amdgcn_rcpf(x);
}
template <>
__device__ __forceinline__ double rcp<double>(double x)
{
    // hardware estimate refined by one Newton step, which squares its
    // relative error; the FP64 members' face values stay at FP64 round-off
    // of the reference formulas (a second step left the errors against the
    // FP64 oracle unchanged)
    double y = __builtin_amdgcn_rcp(x);
    const double e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

// NaN-free min/max: single v_min/v_max instructions (the ternary forms
// compile to a compare and a select).
__device__ __forceinline__ float tmin(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ double tmin(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ float tmax(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ double tmax(double a, double b) { return __builtin_fmax(a, b); }

// Reference WENO face flux (advq_weno.f90:92-110 with weight{x,y,z}
// :329-350 and limiter_weno :462-491), rearranged so that FP32 resolves
// the limiter's eps scale.  a..d = q(f-1..f+2), uu the face velocity,
// nonlin: whether the smoothness switch is active at this face (always in
// x/y, k = 3..nz-3 in z).  Candidates ffim1 = b + (b-a)/2, ffip1 =
// b + (c-b)/2, ffip2 = c - (d-c)/2; only the upwind outer candidate fo and
// its smoothness enter the weights (wo outer, wi inner, wo + wi = 1).
//
// limiter_weno rescales every candidate about b by phi = min(|b/D|, 1),
// D = minp - b + eps, so the face is F = b + phi (Fc - b) with Fc =
// wi ffip1 + wo fo.  Where a candidate undershoots zero, phi pulls minp to
// -eps b/|D|: cells next to exact zeros receive values of order 1e-15,
// and at the next stage those values (comparable to eps) decide phi of
// their own faces.  b + phi (Fc - b) cannot resolve 1e-15 against
// b ~ 1e-5 in FP32, so the face is evaluated in the equivalent form
//   e  = Fc - minp = wi (ffip1 - minp) + wo (fo - minp)   (terms >= 0)
//   m  = minp + eps,  D = m - b,  t = b (eps - e) / D
//   F  = t                  if m and b have opposite signs (phi < 1)
//      = 2b - t             if |m| > 2|b| with the same sign (phi < 1)
//      = minp + e           otherwise (phi = 1),
// whose branch tests compare FP32 values exactly and whose limited value
// carries no cancellation of b.  In exact arithmetic it equals the
// reference face.
template <class T>
__device__ __forceinline__ T weno_face(T a, T b, T c, T d, T uu, bool nonlin)
{
    const T eps = T(1e-15), thr = T(20);
    const T dm = b - a, dc = c - b, dp = d - c;
    const T ffim1 = b + T(0.5) * dm;
    const T ffip1 = b + T(0.5) * dc;
    const T ffip2 = c - T(0.5) * dp;
    const T minp = tmin(tmin(ffim1, ffip1), ffip2);
    const bool pos = uu > T(0);
    const T dout = pos ? dm : dp;
    const T fo = pos ? ffim1 : ffip2;
    // smoothness eta + eps of the inner and the upwind outer pair; the
    // switch max(eta) > 20 (min(eta) + eps) in these terms is
    // max(e) > 20 min(e) + eps
    const T ei = dc * dc + eps, eo = dout * dout + eps;
    T wo = T(1.0 / 3.0), wi = T(2.0 / 3.0);
    if (nonlin) {
        const T mx = tmax(eo, ei), mn = tmin(eo, ei);
        if (mx > thr * mn + eps) {
            // (g1/so) / (g1/so + g2/si) with g2 = 2 g1, and its complement
            const T so = eo * eo, si = ei * ei;
            const T r = rcp(si + T(2) * so);
            wo = si * r;
            wi = T(2) * so * r;
        }
    }
    const T e = wi * (ffip1 - minp) + wo * (fo - minp);
    const T m = minp + eps;
    const T t = b * (eps - e) * rcp(m - b);
    T F = minp + e;
    if ((m < T(0)) != (b < T(0))) F = t;
    else if (fabs(m) > T(2) * fabs(b)) F = T(2) * b - t;
    return uu * F;
}

// WENO5-JS (Jiang & Shu 1996) value at the right face of c from a..e =
// q(i-2..i+2): candidates (2,-7,11)/6, (-1,5,2)/6, (2,5,-1)/6, linear weights
// 0.1, 0.6, 0.3, smoothness indicators beta_r, eps = 1e-6, p = 2; the
// divisions are reciprocals.
template <class T>
__device__ __forceinline__ T weno5_left(T a, T b, T c, T d, T e)
{
    const T eps = T(1e-6), k13 = T(13.0 / 12.0), sixth = T(1.0 / 6.0);
    const T q0 = (T(2) * a - T(7) * b + T(11) * c) * sixth;
    const T q1 = (-b + T(5) * c + T(2) * d) * sixth;
    const T q2 = (T(2) * c + T(5) * d - e) * sixth;
    const T s0 = a - T(2) * b + c, t0 = a - T(4) * b + T(3) * c;
    const T s1 = b - T(2) * c + d, t1 = b - d;
    const T s2 = c - T(2) * d + e, t2 = T(3) * c - T(4) * d + e;
    T b0 = eps + (k13 * s0 * s0 + T(0.25) * t0 * t0);
    T b1 = eps + (k13 * s1 * s1 + T(0.25) * t1 * t1);
    T b2 = eps + (k13 * s2 * s2 + T(0.25) * t2 * t2);
    const T a0 = T(0.1) * rcp(b0 * b0), a1 = T(0.6) * rcp(b1 * b1), a2 = T(0.3) * rcp(b2 * b2);
    return (a0 * q0 + a1 * q1 + a2 * q2) * rcp(a0 + a1 + a2);
}

// WENO5 face flux from q(f-2..f+3), upwind side only
template <class T>
__device__ __forceinline__ T weno5_face(T a, T b, T c, T d, T e, T f, T uu)
{
    return uu > T(0) ? uu * weno5_left(a, b, c, d, e) : uu * weno5_left(f, e, d, c, b);
}

// top closure k = nz-1 of advsz_weno (no limiter, linear weights)
template <class T>
__device__ __forceinline__ T weno_top_face(T a, T b, T c, T uu)
{
    T ffim1 = T(1.5) * b - T(0.5) * a;
    T ffip1 = T(0.5) * (b + c);
    T F = uu > T(0) ? T(1.0 / 3.0) * ffim1 + T(2.0 / 3.0) * ffip1 : ffip1;
    return uu * F;
}

// MUSCL face (advq_muscl.f90), x/y form with slopes in index units.
template <class T>
__device__ __forceinline__ T muscl_phi(int lim, T r)
{
    switch (lim) {
    case 1: return tmax(T(0), tmin(tmin(T(2) * r, T(0.5) * (r + T(1))), T(2)));
    case 2: return (r + fabs(r)) * rcp(T(1) + fabs(r));
    case 0: return tmax(T(0), tmin(T(1), r));
    default: return T(0.5) * (r + T(1));
    }
}

// Only the upwind side's face value is formed: the reference forms both
// (fl from the slope ratio sp/sm, fr from sp/sp2) and keeps one by the sign
// of u; the limited/centred switch still looks at both outer slopes.
// rlt = 1/lim_tol.
template <class T>
__device__ __forceinline__ T muscl_face(T a, T b, T c, T d, T uu, int lim, T lt, T rlt, T d1)
{
    const T eps = T(1e-18);
    const bool pos = uu > T(0);
    const T sp = c - b, sm = b - a, sp2 = d - c;  // slopes times dx
    const T so = pos ? sm : sp2;                  // upwind outer slope
    const T base = pos ? b : c, h = pos ? T(0.5) : T(-0.5);
    const T tol = lt * fabs(b) + eps;
    T F;
    if (fabs(sm) > tol || fabs(sp2) > tol) {
        // limiter_muscl compares per-metre slopes with lim_tol*|x|
        // (advq_muscl.f90:361-362)
        T r = T(1);
        if (fabs(so) * d1 > lt * fabs(b) && fabs(sp) < fabs(so) * rlt) r = sp * rcp(so);
        F = base + h * muscl_phi(lim, r) * so;
    } else {
        F = base + h * sp;
    }
    return uu * F;
}

// MUSCL z face f (1..nz-1) with the reference's stretched metrics, upwind
// side only
template <class T>
__device__ __forceinline__ T muscl_zface(int f, int nz, T a, T b, T c, T d, T ww, const FusedArgs &A)
{
    const T eps = T(1e-18);
    const int lim = A.muscl_limiter;
    const T lt = T(A.lim_tol), rlt = T(A.lim_tol_inv), dz = T(A.dz);
    const double *dz1 = A.dz1, *fdz0 = A.fdz0;
    const T m = T(A.dzw[f - 1]);  // 1/dz1w(f)
    const T sp = (c - b) * T(dz1[f - 1]);
    T sm, sp2;
    bool limited;
    if (f == 1) {
        sm = sp;  // no-flux bottom: bc51 = 0
        sp2 = (d - c) * T(dz1[1]);
        limited = T(fdz0[1]) * dz * fabs(sp2) > lt * fabs(b) + eps;
    } else if (f == nz - 1) {
        sm = (b - a) * T(dz1[f - 2]);
        sp2 = sp;  // no-flux top: bc61 = 0
        limited = T(fdz0[f - 2]) * dz * fabs(sm) > lt * fabs(b) + eps;
    } else {
        sm = (b - a) * T(dz1[f - 2]);
        sp2 = (d - c) * T(dz1[f]);
        limited = T(fdz0[f - 2]) * dz * fabs(sm) > lt * fabs(b) + eps ||
                  T(fdz0[f]) * dz * fabs(sp2) > lt * fabs(b) + eps;
    }
    const bool pos = ww > T(0);
    const T so = pos ? sm : sp2;
    const T base = pos ? b : c, h = pos ? T(0.5) : T(-0.5);
    T F;
    if (limited) {
        T r = T(1);
        if (fabs(so) > lt * fabs(b) && fabs(sp) < fabs(so) * rlt) r = sp * rcp(so);
        F = base + h * muscl_phi(lim, r) * so * m;
    } else {
        F = base + h * sp * m;
    }
    return ww * F;
}

// Stencil half-width of a scheme: the x/y halo of the LDS tile and the
// offset of q(f) in a face's stencil (WENO5 reads q(f-2..f+3), the others
// q(f-1..f+2)).
template <int SCHEME>
constexpr int stencil_h()
{
    return SCHEME == 3 ? 3 : 2;
}

// x/y face from the stencil s[0..2H-1] = q(f-H+1..f+H)
template <int SCHEME, class T>
__device__ __forceinline__ T hface(const T *s, T uu, const FusedArgs &A, double d1)
{
    if constexpr (SCHEME == 0) return weno_face(s[0], s[1], s[2], s[3], uu, true);
    else if constexpr (SCHEME == 3) return weno5_face(s[0], s[1], s[2], s[3], s[4], s[5], uu);
    else return muscl_face(s[0], s[1], s[2], s[3], uu, A.muscl_limiter, T(A.lim_tol), T(A.lim_tol_inv), T(d1));
}

template <int SCHEME, class T>
__device__ __forceinline__ T zface(int f, int nz, const T *zw, T ww, const FusedArgs &A)
{
    // zw = q(f-H+1..f+H) (H = stencil_h), levels below 1 read as q(1)
    if (f >= nz) return T(0);
    if constexpr (SCHEME == 3) {
        // WENO5 where the 5-cell upwind stencil lies in the column, the
        // reference face elsewhere
        if (ww > T(0) && f >= 3 && f <= nz - 2) return ww * weno5_left(zw[0], zw[1], zw[2], zw[3], zw[4]);
        if (ww < T(0) && f >= 2 && f <= nz - 3) return ww * weno5_left(zw[5], zw[4], zw[3], zw[2], zw[1]);
        return zface<0, T>(f, nz, zw + 1, ww, A);
    } else if constexpr (SCHEME == 0) {
        if (f == nz - 1) return weno_top_face(zw[0], zw[1], zw[2], ww);
        if (f == 1 && A.weno_bottom == 1) {
            T ffip1 = T(0.5) * (zw[1] + zw[2]);
            T ffip2 = T(1.5) * zw[2] - T(0.5) * zw[3];
            T F = ww > T(0) ? ffip1 : T(2.0 / 3.0) * ffip1 + T(1.0 / 3.0) * ffip2;
            return ww * F;
        }
        return weno_face(zw[0], zw[1], zw[2], zw[3], ww, f >= 3 && f <= nz - 3);
    } else {
        return muscl_zface(f, nz, zw[0], zw[1], zw[2], zw[3], ww, A);
    }
}

// Element at a 32-bit byte offset from a wave-uniform base pointer: the
// global load/store takes the base from scalar registers and the offset
// from one vector register, with no 64-bit vector address arithmetic.
template <class T>
__device__ __forceinline__ T load_at(const T *base, uint32_t off)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
}
template <class T>
__device__ __forceinline__ void store_at(T *base, uint32_t off, T v)
{
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + off) = v;
}

__device__ __forceinline__ int wrap(int i, int n, int m)
{
    // 1-based index i mapped into [1, n]; outside positions take their
    // periodic interior image (period m, interior starts at 4)
    if (i >= 1 && i <= n) return i;
    int r = (i - 4) % m;
    return (r < 0 ? r + m : r) + 4;
}

}  // namespace faces
}  // namespace mimadv
