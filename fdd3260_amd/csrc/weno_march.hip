// weno_march.hip -- plain per-field FP64 kernels, column-marching form: the
// reference 'weno' scheme, the WENO5-JS extension, and the line
// schemes MUSCL, QUICK and Lax-Wendroff with their unlimited forms (face
// formulas of schemes_fp64.hpp).
//
// Same results as the CPU restatement, bit for bit: every face value is
// computed with the reference operations (for 'weno' advq_weno.f90:92-110,
// weight{x,y,z} :329-350, limiter_weno :462-491) in the reference order,
// only
//   * once per face instead of once per adjacent cell: the x and y faces of
//     a level are evaluated by a flattened loop over the workgroup and
//     exchanged through LDS, the z face k+1/2 is carried to level k+1;
//   * for 'weno' and WENO5, for the upwind side only.  The reference forms
//     F = 0.5*fl*(u+|u|) + 0.5*fr*(u-|u|), where the off-side term is an
//     exact zero, and the two sides' weights and candidates are the same
//     expressions with outer and inner swapped (fl = w1*f1 + w2*f2,
//     fr = w3*f2 + w4*f4 with w1, w4 the outer weights), so the upwind
//     side alone gives the same value up to the sign of an exact zero.
//
// Work mapping: a workgroup owns a TX x TY tile of columns and a chunk of
// levels.  Per level it stages the tile plus a 2-cell (WENO5: 3-cell) halo
// of q and the face velocities in LDS (ghost columns read as the caller
// filled them; everything is requested one level ahead), evaluates the
// TY*(TX+1) x faces and (TY+1)*TX y faces of the tile, and each thread
// finishes its cell with the z faces from a register window q(k-1..k+2).  A chunk
// starting above the floor first evaluates its bottom face.
//
// Compiled with -ffp-contract=off.
#include <algorithm>

#include "mimadv_internal.hpp"
#include "weno_ratio.hpp"

namespace mimadv {

namespace {

// FP64 division in the steps the compiler emits for n / d on gfx950
// (reciprocal estimate, two Newton steps, quotient, residual, corrected
// quotient) without its operand scaling and special-value fix-up.  Those
// two steps leave operands and result unchanged when n, d and n/d are
// normal numbers well inside the exponent range, so within that range the
// result is the correctly rounded quotient, bit for bit the same as n / d
// (checked on 3.4e10 random operand pairs in [2^-400, 2^400]).  A
// reciprocal serves every quotient with the same denominator.
__device__ __forceinline__ double refined_rcp(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    return fma(r, e, r);
}

__device__ __forceinline__ double quot(double n, double d, double r)
{
    const double q = n * r;
    return fma(fma(-d, q, n), r, q);
}

// n / 6, correctly rounded, from the correctly rounded reciprocal y of 6:
// q = n y is within an ulp of n/6 and q + (n - 6 q) y rounds to n / 6
// (Markstein's corrected quotient; the residual is exact in an fma) when
// n, q and the residual are normal.  Zero keeps its sign; operands outside
// [2^-960, 2^960] and non-finite ones take the division.
__device__ __forceinline__ double div6(double n)
{
    constexpr double y = 1.0 / 6.0;
    const double q = n * y;
    const double an = fabs(n);
    if (an >= 0x1p-960 && an <= 0x1p960) return fma(fma(-6.0, q, n), y, q);
    if (n == 0.0) return q;
    return n / 6.0;
}

// n / d, d > 0, with the refined reciprocal inside the range the steps are
// exact for (see refined_rcp); zero numerators keep their sign.
__device__ __forceinline__ double div_pos(double n, double d)
{
    const double an = fabs(n);
    if (d >= 0x1p-400 && d <= 0x1p400 && an >= 0x1p-400 && an <= 0x1p400) return quot(n, d, refined_rcp(d));
    if (n == 0.0 && d >= 0x1p-400 && d <= 0x1p400) return n * refined_rcp(d);
    return n / d;
}

// WENO5-JS reconstruction at the right face of c from a..e (the oracle's
// weno5_left), bit for bit: the same operation order, with the divisions as
// div6 / div_pos.
__device__ __forceinline__ double weno5_left_x(double a, double b, double c, double d, double e)
{
    const double eps = 1e-6;
    const double q0 = div6(2.0 * a - 7.0 * b + 11.0 * c);
    const double q1 = div6(-b + 5.0 * c + 2.0 * d);
    const double q2 = div6(2.0 * c + 5.0 * d - e);
    const double b0 = 13.0 / 12.0 * (a - 2.0 * b + c) * (a - 2.0 * b + c) + 0.25 * (a - 4.0 * b + 3.0 * c) * (a - 4.0 * b + 3.0 * c);
    const double b1 = 13.0 / 12.0 * (b - 2.0 * c + d) * (b - 2.0 * c + d) + 0.25 * (b - d) * (b - d);
    const double b2 = 13.0 / 12.0 * (c - 2.0 * d + e) * (c - 2.0 * d + e) + 0.25 * (3.0 * c - 4.0 * d + e) * (3.0 * c - 4.0 * d + e);
    const double a0 = div_pos(0.1, (eps + b0) * (eps + b0));
    const double a1 = div_pos(0.6, (eps + b1) * (eps + b1));
    const double a2 = div_pos(0.3, (eps + b2) * (eps + b2));
    return div_pos(a0 * q0 + a1 * q1 + a2 * q2, a0 + a1 + a2);
}

// Reference face flux for the upwind side only.  nonlin: weights may
// switch (x/y always, z for faces 3..nz-3).
__device__ __forceinline__ double weno_face_up(double a, double b, double c, double d, double uu, bool nonlin)
{
    const double g1 = 1.0 / 3.0, g2 = 2.0 / 3.0, eps = 1.e-15;
    const bool left = uu > 0.0;
    double wo = g1, wi = g2;  // outer, inner
    if (nonlin) {
        const double eta_i = (c - b) * (c - b);
        const double eta_o = left ? (b - a) * (b - a) : (d - c) * (d - c);
        // left pair (eta1, eta2) = (outer, inner), right pair (eta2, eta3) =
        // (inner, outer); max and min do not depend on the order
        if (ratio_above_20(dmaxd(eta_o, eta_i), dmind(eta_o, eta_i) + eps)) {
            const double so = (eps + eta_o) * (eps + eta_o);
            const double si = (eps + eta_i) * (eps + eta_i);
            // so, si >= 1e-30; below 2^400 every operand and quotient here
            // lies in [2^-502, 2^400] (NaN and inf take the plain divisions)
            if (so <= 0x1p400 && si <= 0x1p400) {
                const double w_o = quot(g1, so, refined_rcp(so)), w_i = quot(g2, si, refined_rcp(si));
                const double sumw = w_o + w_i, rs = refined_rcp(sumw);
                wo = quot(w_o, sumw, rs);
                wi = quot(w_i, sumw, rs);
            } else {
                const double w_o = g1 / so, w_i = g2 / si;
                const double sumw = w_o + w_i;
                wo = w_o / sumw;
                wi = w_i / sumw;
            }
        }
    }
    double ffim1 = -(0.5 * a) + 1.5 * b;
    double ffip1 = 0.5 * b + 0.5 * c;
    double ffip2 = 1.5 * c - 0.5 * d;
    const double minp = dmind(dmind(dmind(ffim1, ffip1), ffip1), ffip2);
    // phi = min(|(0-b)/den|, 1): when |b| >= |den| the correctly rounded
    // quotient is at least 1 in magnitude (or inf, or NaN for 0/0, which
    // fmin maps to 1), so phi = 1 without the division
    const double den = minp - b + eps;
    const double phi = fabs(b) >= fabs(den) ? 1.0 : fmin(fabs((0.0 - b) / den), 1.0);
    const double fo = phi * ((left ? ffim1 : ffip2) - b) + b;
    const double fi = phi * (ffip1 - b) + b;
    const double f = wo * fo + wi * fi;
    return left ? 0.5 * f * (uu + fabs(uu)) : 0.5 * f * (uu - fabs(uu));
}

// top closure k = nz-1 (advq_weno.f90:262-271): no limiter, linear weights
__device__ __forceinline__ double weno_face_top(double a, double b, double c, double ww)
{
    const double g1 = 1.0 / 3.0, g2 = 2.0 / 3.0;
    const double ffip1 = 0.5 * b + 0.5 * c;
    if (ww > 0.0) {
        const double fl = g1 * (-(0.5 * a) + 1.5 * b) + g2 * ffip1;
        return 0.5 * fl * (ww + fabs(ww));
    }
    return 0.5 * ffip1 * (ww - fabs(ww));
}

// bottom closure kept (weno_bottom = 1): fl = (x1+x2)/2, fr = 2/3 fl + 1/3 f4
__device__ __forceinline__ double weno_face_bottom(double b, double c, double d, double ww)
{
    const double g1 = 1.0 / 3.0, g2 = 2.0 / 3.0;
    const double ffip1 = 0.5 * b + 0.5 * c;
    if (ww > 0.0) return 0.5 * ffip1 * (ww + fabs(ww));
    const double fr = g2 * ffip1 + g1 * (1.5 * c - 0.5 * d);
    return 0.5 * fr * (ww - fabs(ww));
}

__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

// z face f + 1/2 of a column; q(f-1..f+2) given (q(0) := q(1) by the caller)
__device__ __forceinline__ double zface(int f, int nz, int bottom, double a, double b, double c, double d, double ww)
{
    if (f == nz - 1) return weno_face_top(a, b, c, ww);
    if (f == 1 && bottom == 1) return weno_face_bottom(b, c, d, ww);
    return weno_face_up(a, b, c, d, ww, f >= 3 && f <= nz - 3);
}

// x/y face flux from the LDS stencil starting at b (step st): the reference
// 'weno' and the line schemes read q(f-1..f+2) (4 values), WENO5 reads
// q(f-2..f+3) (6 values).  'weno' and WENO5 evaluate the upwind side only:
// in the reference-form sum 0.5*fl*(u+|u|) + 0.5*fr*(u-|u|) the off-side
// term is an exact zero.  The line schemes (MUSCL, QUICK, Lax-Wendroff and
// the unlimited forms) evaluate the reference face formula as it stands;
// they have no position-dependent closures in x and y, so the stencil is
// passed as cells 0..3 of face 1+1/2.  d, d1: grid spacing and its inverse.
template <int S>
__device__ __forceinline__ double xy_face(const double *t, int b, int st, double uu, double d, double d1,
                                          const SchemeParams &sp)
{
    if constexpr (S == 0) {
        return weno_face_up(t[b], t[b + st], t[b + 2 * st], t[b + 3 * st], uu, true);
    } else if constexpr (S != 3) {
        auto Q = [&](int m) { return t[b + m * st]; };
        auto U = [&](int) { return uu; };
        if constexpr (S == 2) return muscl_hface(1, d, d1, sp, Q, U);
        else if constexpr (S == 4) return quick_hface<false>(1, sp, Q, U);
        else if constexpr (S == 6) return quick_hface<true>(1, sp, Q, U);
        else if constexpr (S == 5) return lw_hface<false>(1, d1, sp, Q, U);
        else return lw_hface<true>(1, d1, sp, Q, U);
    } else {
        if (uu > 0.0) {
            const double fl = weno5_left_x(t[b], t[b + st], t[b + 2 * st], t[b + 3 * st], t[b + 4 * st]);
            return 0.5 * fl * (uu + fabs(uu));
        }
        const double fr = weno5_left_x(t[b + 5 * st], t[b + 4 * st], t[b + 3 * st], t[b + 2 * st], t[b + st]);
        return 0.5 * fr * (uu - fabs(uu));
    }
}

// z face f + 1/2 of a line scheme; Q(kk) any level of the column, ww = w(f+1)
template <int S, class QF>
__device__ __forceinline__ double line_zface(int f, int nz, const ZMetrics &zm, const SchemeParams &sp, double bc51,
                                             double bc61, QF Q, double ww)
{
    auto W = [&](int) { return ww; };
    if constexpr (S == 2) return muscl_zface(f, nz, zm, sp, bc51, bc61, Q, W);
    else if constexpr (S == 4) return quick_zface<false>(f, nz, zm, sp, bc51, bc61, Q, W);
    else if constexpr (S == 6) return quick_zface<true>(f, nz, zm, sp, bc51, bc61, Q, W);
    else if constexpr (S == 5) return lw_zface<false>(f, nz, zm, sp, bc51, bc61, Q, W);
    else return lw_zface<true>(f, nz, zm, sp, bc51, bc61, Q, W);
}

// S (scheme code): 0 the reference 'weno', 3 WENO5-JS (in z where the
// upwind stencil fits the column, the reference 'weno' face elsewhere),
// 2 MUSCL, 4 QUICK, 5 Lax-Wendroff, 6 and 7 their
// unlimited forms
template <int TX, int TY, int S>
__device__ __forceinline__ void march_body(const GridDev &g, const SchemeParams &sp, const ScalarArgs &a,
                                                int kchunk, int ntx, int nwork)
{
    constexpr bool THREE_D = TY > 1;
    constexpr int NT = TX * TY;
    constexpr int H = S == 3 ? 3 : 2;  // x/y halo of the LDS tile
    constexpr int LW = TX + 2 * H, LH = THREE_D ? TY + 2 * H : 1;
    constexpr int NXF = TY * (TX + 1), NYF = THREE_D ? (TY + 1) * TX : 0, NF = NXF + NYF;
    constexpr int FS = (NF + NT - 1) / NT, TS = (LH * LW + NT - 1) / NT;
    __shared__ double tile[LH * LW];
    __shared__ double face[FS * NT];
    __shared__ double fvel[FS * NT];  // face velocities of each thread's slots
    // per-thread slot constants of the current work item, kept in LDS rather
    // than registers: [t_off(TS) | f_off(FS) | f_slot(FS)] x NT
    __shared__ unsigned slots[(TS + 2 * FS) * NT];

    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const size_t plane = (size_t)nx * ny;
    const int mx = nx - 5, my = ny - 5;
    const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
    const int nchunk = (nz + kchunk - 1) / kchunk;
    const double dx1 = 1.0 / g.dx, dy1 = 1.0 / g.dy;
    const bool adv = a.adv != 0;
    const int bc5 = g.bcl[4], bc6 = g.bcl[5];
    const double bc51 = bc5 == 0 ? 1.0 : 0.0, bc61 = bc6 == 0 ? 1.0 : 0.0;  // periodic floor, lid
    const double *__restrict__ dz1w = g.zm.dz1w;
    const bool zdir = (a.dirmask & 4) != 0;
    const bool accum = a.epi == EPI_ACCUM, heun = a.epi == EPI_HEUN, so = a.epi == EPI_SO;
    const bool need_tend = heun && a.b_lm != 0.0, need_q0 = heun || so;

    // one work item per workgroup (the grid is nwork groups): values derived
    // from the thread index are not hoisted into a loop prologue and spilled
    if (const int wi = blockIdx.x; wi < nwork) {
        const int tile_id = wi / nchunk, kc = wi - tile_id * nchunk;
        const int i0 = 4 + (tile_id % ntx) * TX;
        const int j0 = THREE_D ? 4 + (tile_id / ntx) * TY : 1;
        const int k0 = 1 + kc * kchunk, k1 = min(nz, k0 + kchunk - 1);
        const int i = i0 + tx, j = j0 + ty;
        const bool valid = i <= nx - 2 && (!THREE_D || j <= ny - 2);
        // loads of threads past the edge use the last interior column
        const int il = min(i, nx - 2), jl = THREE_D ? min(j, ny - 2) : 1;
        const unsigned col = (unsigned)((il - 1) + nx * (jl - 1));
        auto Qz = [&](int kk) { return a.q[col + (size_t)(clampi(kk, 1, nz) - 1) * plane]; };
        // fixed slots of this thread: tile values (ghost columns read as the
        // caller filled them; positions past the array take a periodic image
        // that feeds no interior cell) and faces of the flattened list
        // [x faces row-major | y faces] with their velocity offsets
        // (plane offsets are 32-bit: a plane holds fewer than 2^31 points;
        // f_slot = LDS stencil start | kind << 16, kind 0 empty, 1 x, 2 y face)
        auto t_off = [&](int s) { return slots[s * NT + tid]; };
        auto f_off = [&](int s) { return slots[(TS + s) * NT + tid]; };
        auto f_slot = [&](int s) { return (int)slots[(TS + FS + s) * NT + tid]; };
#pragma unroll
        for (int s = 0; s < TS; ++s) {
            const int e = tid + s * NT;
            unsigned off = 0;
            if (e < LH * LW) {
                const int r = e / LW, c = e - r * LW;
                int ii = i0 - H + c;
                if (ii > nx) ii = (ii - 4) % mx + 4;
                int jj = 1;
                if (THREE_D) {
                    jj = j0 - H + r;
                    if (jj > ny) jj = (jj - 4) % my + 4;
                }
                off = (unsigned)((ii - 1) + nx * (jj - 1));
            }
            slots[s * NT + tid] = off;
        }
#pragma unroll
        for (int s = 0; s < FS; ++s) {
            const int e = tid + s * NT;
            unsigned slot = 0, off = 0;
            if (e < NXF) {
                // x face f + 1/2 (velocity u(f+1)), f = i0-1+fl, of tile row r
                const int r = e / (TX + 1), fl = e - r * (TX + 1);
                const int jr = THREE_D ? min(j0 + r, ny - 2) : 1;
                slot = ((THREE_D ? r + H : 0) * LW + fl) | 1 << 16;
                off = (unsigned)((min(i0 + fl, nx) - 1) + nx * (jr - 1));
            } else if (e < NF) {
                // y face g + 1/2 (velocity v(g+1)), g = j0-1+fr, of tile column c
                const int e2 = e - NXF, fr = e2 / TX, c = e2 - fr * TX;
                slot = (fr * LW + c + H) | 2 << 16;
                off = (unsigned)((min(i0 + c, nx - 2) - 1) + nx * (min(j0 + fr, ny) - 1));
            }
            slots[(TS + s) * NT + tid] = off;
            slots[(TS + FS + s) * NT + tid] = slot;
        }
        // A chunk above the floor first evaluates its bottom face k0-1/2 in a
        // z-only pass, so that every z face comes from the same code.
        const int kstart = (k0 > 1 && zdir) ? k0 - 1 : k0;
        double z0 = Qz(kstart - 1), z1 = Qz(kstart), z2 = Qz(kstart + 1), z3 = Qz(kstart + 2);
        // WENO5 widens the window to q(k-2..k+3)
        double zm1 = S == 3 ? Qz(kstart - 2) : 0.0, z4 = S == 3 ? Qz(kstart + 3) : 0.0;
        double wz = kstart < nz ? a.w[col + (size_t)kstart * plane] : 0.0;
        double fzb = 0.0;
        // operands of level k, requested one level ahead
        double tv[TS], vv[FS], wv = 0.0, zn = 0.0, q0v = 0.0, tdv = 0.0, xsv = 0.0;
        auto fetch = [&](int k) {
            const size_t kb = (size_t)(k - 1) * plane;
            const double *qk = a.q + kb, *uk = a.u + kb, *vk = a.v + kb;
#pragma unroll
            for (int s = 0; s < TS; ++s)
                if (tid + s * NT < LH * LW) tv[s] = qk[t_off(s)];
#pragma unroll
            for (int s = 0; s < FS; ++s)
            {
                const int kind = f_slot(s) >> 16;
                vv[s] = kind == 1 ? uk[f_off(s)] : (kind == 2 ? vk[f_off(s)] : 0.0);
            }
            wv = k < nz ? a.w[col + kb + plane] : 0.0;
            zn = Qz(S == 3 ? k + 3 : k + 2);
            if (valid) {
                if (need_q0) q0v = a.q0[col + kb];
                if (need_tend) tdv = a.tend[col + kb];
                if (accum) xsv = a.xs[col + kb];
            }
        };
        fetch(k0);
#pragma unroll 1
        for (int k = kstart; k <= k1; ++k) {
            if (k > kstart) {
                if constexpr (S == 3) {
                    zm1 = z0;
                    z0 = z1;
                    z1 = z2;
                    z2 = z3;
                    z3 = z4;
                    z4 = zn;
                } else {
                    z0 = z1;
                    z1 = z2;
                    z2 = z3;
                    z3 = zn;
                }
            }
            const double wc = k < k0 ? wz : wv;
            double fzt = 0.0;
            if (zdir && k <= nz - 1) {
                if constexpr (S == 3) {
                    // WENO5-JS where the 5-cell upwind stencil lies in the
                    // column, the reference face elsewhere
                    if (wc > 0.0 && k >= 3 && k <= nz - 2)
                        fzt = 0.5 * weno5_left_x(zm1, z0, z1, z2, z3) * (wc + fabs(wc));
                    else if (wc < 0.0 && k >= 2 && k <= nz - 3)
                        fzt = 0.5 * weno5_left_x(z4, z3, z2, z1, z0) * (wc - fabs(wc));
                    else
                        fzt = zface(k, nz, sp.weno_bottom, k == 1 ? z1 : z0, z1, z2, z3, wc);
                } else if constexpr (S == 0) {
                    fzt = zface(k, nz, sp.weno_bottom, k == 1 ? z1 : z0, z1, z2, z3, wc);
                } else {
                    // q(k-1..k+2) from the window, other levels (the periodic
                    // floor and lid closures) from memory
                    auto Qw = [&](int kk) {
                        const int m = kk - (k - 1);
                        return m == 0 ? z0 : m == 1 ? z1 : m == 2 ? z2 : m == 3 ? z3 : Qz(kk);
                    };
                    fzt = line_zface<S>(k, nz, g.zm, sp, bc51, bc61, Qw, wc);
                }
            }
            if (k < k0) {
                fzb = fzt;
                continue;
            }
            const size_t kb = (size_t)(k - 1) * plane;
#pragma unroll
            for (int s = 0; s < TS; ++s)
                if (tid + s * NT < LH * LW) tile[tid + s * NT] = tv[s];
#pragma unroll
            for (int s = 0; s < FS; ++s) fvel[tid + s * NT] = vv[s];
            const double q0c = q0v, tdc = tdv, xsc = xsv;
            __syncthreads();
            if (k < k1) fetch(k + 1);
            // one face at a time: the exact FP64 face is register-hungry
#pragma unroll 1
            for (int s = 0; s < FS; ++s) {
                const int fs = f_slot(s), kind = fs >> 16;
                if (kind == 0) continue;
                const int st = kind == 1 ? 1 : LW, b = fs & 0xffff;
                face[tid + s * NT] = xy_face<S>(tile, b, st, fvel[tid + s * NT], kind == 1 ? g.dx : g.dy,
                                                kind == 1 ? dx1 : dy1, sp);
            }
            __syncthreads();
            if (valid) {
                const size_t id = col + kb;
                double X = 0.0, Y = 0.0, Z = 0.0;
                if (a.dirmask & 1) {
                    const int e = ty * (TX + 1) + tx;
                    X = (face[e + 1] - face[e]) * dx1;
                    if (adv) {
                        const double *urow = a.u + kb + (size_t)nx * (j - 1);
                        X = X - z1 * (urow[i] - urow[i - 1]) * dx1;
                    }
                }
                if (THREE_D && (a.dirmask & 2)) {
                    const int e = NXF + ty * TX + tx;
                    Y = (face[e + TX] - face[e]) * dy1;
                    if (adv) {
                        const double *vcol = a.v + kb + (size_t)(i - 1);
                        Y = Y - z1 * (vcol[(size_t)nx * j] - vcol[(size_t)nx * (j - 1)]) * dy1;
                    }
                }
                if (zdir) {
                    // z_assemble (no-flux 'nnf' floor/lid, zero z tendency at other ends)
                    if (k >= 2 && k <= nz - 1) {
                        Z = (fzt - fzb) * dz1w[k - 1];
                        if (adv) Z = Z - z1 * (wc - a.w[col + kb]) * dz1w[k - 1];
                    } else {
                        double t = 0.0;
                        if (k == 1 && bc5 == 1) {
                            t = t + (fzt - 0.0) * dz1w[0];
                            if (adv) t = t - z1 * (wc - 0.0) * dz1w[0];
                        }
                        if (k == nz && bc6 == 1) {
                            t = t + (0.0 - fzb) * dz1w[nz - 1];
                            if (adv) t = t - z1 * (0.0 - a.w[col + kb]) * dz1w[nz - 1];
                        }
                        Z = t;
                    }
                }
                if (accum || a.epi == EPI_OVERWRITE) {
                    double r = accum ? xsc : 0.0;
                    if (a.dirmask & 1) r = r - X;
                    if (THREE_D && (a.dirmask & 2)) r = r - Y;
                    if (zdir) r = r - Z;
                    a.xs[id] = r;
                } else {
                    double A = 0.0;
                    if (a.dirmask & 1) A = A - X;
                    if (THREE_D && (a.dirmask & 2)) A = A - Y;
                    if (zdir) A = A - Z;
                    double qn;
                    if (heun) {
                        const double T = need_tend ? A + a.b_lm * tdc : A + 0.0;
                        a.tend[id] = T;
                        qn = q0c + a.adt * T;
                    } else {
                        qn = a.c0 * q0c + a.c1 * (z1 + a.dt * A);
                    }
                    a.qn[id] = qn;
                    if (a.fill) {
                        // periodic images (qnbc_p4, bc.f90:409-414, :431-436)
                        int ix = 0, jy = 0;
                        if (i >= nx - 4) ix = i - mx;
                        else if (i <= 5) ix = i + mx;
                        if (THREE_D) {
                            if (j >= ny - 4) jy = j - my;
                            else if (j <= 5) jy = j + my;
                        }
                        double *qo = a.qn + kb;
                        if (ix) qo[(size_t)(ix - 1) + (size_t)nx * (j - 1)] = qn;
                        if (jy) {
                            qo[(size_t)(i - 1) + (size_t)nx * (jy - 1)] = qn;
                            if (ix) qo[(size_t)(ix - 1) + (size_t)nx * (jy - 1)] = qn;
                        }
                    }
                }
            }
            fzb = fzt;
        }
    }
}

// All instances are capped at 128 VGPRs (4 waves per SIMD); the per-thread
// slot constants live in LDS to make room.  Uncapped, the reference-weno
// 32x8 instance took 151 VGPRs (3 waves) and 217 us per launch at
// 256x256x128; WENO5 took 170 (2 waves).
template <int TX, int TY>
__global__ __launch_bounds__(TX *TY) __attribute__((amdgpu_waves_per_eu(4, 8))) void weno_march_kernel(
    GridDev g, SchemeParams sp, ScalarArgs a, int kchunk, int ntx, int nwork)
{
    march_body<TX, TY, 0>(g, sp, a, kchunk, ntx, nwork);
}

template <int TX, int TY>
__global__ __launch_bounds__(TX *TY) __attribute__((amdgpu_waves_per_eu(4, 8))) void weno5_march_kernel(
    GridDev g, SchemeParams sp, ScalarArgs a, int kchunk, int ntx, int nwork)
{
    march_body<TX, TY, 3>(g, sp, a, kchunk, ntx, nwork);
}

template <int S, int TX, int TY>
__global__ __launch_bounds__(TX *TY) __attribute__((amdgpu_waves_per_eu(4, 8))) void line_march_kernel(
    GridDev g, SchemeParams sp, ScalarArgs a, int kchunk, int ntx, int nwork)
{
    march_body<TX, TY, S>(g, sp, a, kchunk, ntx, nwork);
}

template <int TX, int TY, int S, int KC>
hipError_t launch_tile(const GridDev &g, const SchemeParams &sp, const ScalarArgs &a, int kchunk, hipStream_t s)
{
    const int ntx = (g.nx - 5 + TX - 1) / TX;
    const int nty = TY > 1 ? (g.ny - 5 + TY - 1) / TY : 1;
    if (kchunk <= 0) kchunk = KC;
    if (kchunk > g.nz) kchunk = g.nz;
    const int nchunk = (g.nz + kchunk - 1) / kchunk;
    const int nwork = ntx * nty * nchunk;
    if constexpr (S == 3)
        hipLaunchKernelGGL((weno5_march_kernel<TX, TY>), dim3(nwork), dim3(TX * TY), 0, s, g, sp, a, kchunk, ntx,
                           nwork);
    else if constexpr (S == 0)
        hipLaunchKernelGGL((weno_march_kernel<TX, TY>), dim3(nwork), dim3(TX * TY), 0, s, g, sp, a, kchunk, ntx,
                           nwork);
    else
        hipLaunchKernelGGL((line_march_kernel<S, TX, TY>), dim3(nwork), dim3(TX * TY), 0, s, g, sp, a, kchunk,
                           ntx, nwork);
    return hipGetLastError();
}

// Tiles, measured at 256x256x128 before the VGPR cap of the reference-weno
// instance: the reference 'weno' on 32x8 columns with 8-level chunks (217 us
// per launch; 64x4 took 228: 2.27 x/y faces per cell instead of 2.16 and a
// wider halo), WENO5 on 16x16 with 16-level chunks (285 us; 32x8 takes 308,
// 64x4 323).  Tiles of 512 threads are slower (fewer waves per SIMD).
// The line schemes use the reference-weno tiles.
template <int S, int TX3, int TY3, int KC3>
hipError_t launch_dims(const GridDev &g, const SchemeParams &sp, const ScalarArgs &a, int kchunk, hipStream_t s)
{
    if (g.ny > 1) return launch_tile<TX3, TY3, S, KC3>(g, sp, a, kchunk, s);
    return launch_tile<64, 1, S, 16>(g, sp, a, kchunk, s);
}

}  // namespace

// ---- positivity-preserving Shu-Osher stage (MIMADV_POSITIVE; no
// reference counterpart).  Three passes over the grid, one thread per
// cell: the 'weno' face fluxes (the same face functions as the march
// kernel, so the same bits), the per-cell outflow factor theta, and the
// stage from the donor-scaled fluxes.  The arithmetic follows the oracle's
// orc_positive_stage operation for operation.
namespace {

__device__ __forceinline__ int wrapi(int i, int n, int m) { return i < 4 ? i + m : (i > n - 2 ? i - m : i); }

__global__ __launch_bounds__(256) void pos_flux_kernel(GridDev g, SchemeParams sp, PosArgs a)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const bool three_d = ny > 1;
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny, n3 = sz * nz;
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < n3; id += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(id % sy) + 1, j = (int)((id / sy) % ny) + 1, k = (int)(id / sz) + 1;
        const bool irow = three_d ? (j >= 4 && j <= ny - 2) : true;
        const bool icol = i >= 4 && i <= nx - 2;
        const double *q = a.q;
        // x face i+1/2 of the interior rows
        double fx = 0.0;
        if (irow && i >= 3 && i <= nx - 2) fx = weno_face_up(q[id - 1], q[id], q[id + 1], q[id + 2], a.u[id + 1], true);
        a.fx[id] = fx;
        // y face j+1/2 of the interior columns
        double fy = 0.0;
        if (three_d && icol && j >= 3 && j <= ny - 2)
            fy = weno_face_up(q[id - sy], q[id], q[id + sy], q[id + 2 * sy], a.v[id + sy], true);
        a.fy[id] = fy;
        // z face k+1/2 of the interior columns, q(0) = q(1)
        double fz = 0.0;
        if (irow && icol && k <= nz - 1) {
            const double z0 = k == 1 ? q[id] : q[id - sz];
            const double z3 = k + 2 <= nz ? q[id + 2 * sz] : 0.0;
            fz = zface(k, nz, sp.weno_bottom, z0, q[id], q[id + sz], z3, a.w[id + sz]);
        }
        a.fz[id] = fz;
    }
}

__device__ __forceinline__ double pos_part(double f) { return f > 0.0 ? f : 0.0; }
__device__ __forceinline__ double neg_part(double f) { return f < 0.0 ? -f : 0.0; }

__global__ __launch_bounds__(256) void pos_theta_kernel(GridDev g, PosArgs a)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const bool three_d = ny > 1;
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny, n3 = sz * nz;
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < n3; id += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(id % sy) + 1, j = (int)((id / sy) % ny) + 1, k = (int)(id / sz) + 1;
        if (i < 4 || i > nx - 2 || (three_d && (j < 4 || j > ny - 2))) continue;
        const double ox = (pos_part(a.fx[id]) + neg_part(a.fx[id - 1])) * a.dx1;
        double o = ox;
        if (three_d) o = o + (pos_part(a.fy[id]) + neg_part(a.fy[id - sy])) * a.dy1;
        const double fzt = k < nz ? a.fz[id] : 0.0, fzb = k > 1 ? a.fz[id - sz] : 0.0;
        o = o + (pos_part(fzt) + neg_part(fzb)) * g.zm.dz1w[k - 1];
        o = a.dt * o;
        const double qp = a.q[id] > 0.0 ? a.q[id] : 0.0;
        a.th[id] = o > qp ? (qp / o) * (1.0 - 0x1p-40) : 1.0;
    }
}

__global__ __launch_bounds__(256) void pos_update_kernel(GridDev g, PosArgs a)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz, mx = nx - 5, my = ny - 5;
    const bool three_d = ny > 1;
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny, n3 = sz * nz;
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < n3; id += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(id % sy) + 1, j = (int)((id / sy) % ny) + 1, k = (int)(id / sz) + 1;
        if (i < 4 || i > nx - 2 || (three_d && (j < 4 || j > ny - 2))) continue;
        const size_t kb = sz * (size_t)(k - 1);
        const size_t rowb = kb + sy * (size_t)(j - 1);
        // flux of x face f+1/2 of this row, scaled by its donor's theta
        auto fxs = [&](int f) {
            const double F = a.fx[rowb + (size_t)(f - 1)];
            return F * (F > 0.0 ? a.th[rowb + (size_t)(wrapi(f, nx, mx) - 1)]
                                : a.th[rowb + (size_t)(wrapi(f + 1, nx, mx) - 1)]);
        };
        auto fys = [&](int f) {
            const size_t cb = kb + (size_t)(i - 1);
            const double F = a.fy[cb + sy * (size_t)(f - 1)];
            return F * (F > 0.0 ? a.th[cb + sy * (size_t)(wrapi(f, ny, my) - 1)]
                                : a.th[cb + sy * (size_t)(wrapi(f + 1, ny, my) - 1)]);
        };
        auto fzs = [&](int f) {
            const size_t cb = (size_t)(i - 1) + sy * (size_t)(j - 1);
            const double F = a.fz[cb + sz * (size_t)(f - 1)];
            return F * (F > 0.0 ? a.th[cb + sz * (size_t)(f - 1)] : a.th[cb + sz * (size_t)f]);
        };
        const double X = (fxs(i) - fxs(i - 1)) * a.dx1;
        const double Y = three_d ? (fys(j) - fys(j - 1)) * a.dy1 : 0.0;
        double Z;
        if (k == 1) Z = (fzs(1) - 0.0) * g.zm.dz1w[0];
        else if (k == nz) Z = (0.0 - fzs(nz - 1)) * g.zm.dz1w[nz - 1];
        else Z = (fzs(k) - fzs(k - 1)) * g.zm.dz1w[k - 1];
        double A = 0.0;
        A = A - X;
        if (three_d) A = A - Y;
        A = A - Z;
        const double qn = a.c0 * a.q0[id] + a.c1 * (a.q[id] + a.dt * A);
        a.qn[id] = qn;
        // periodic images (qnbc_p4, bc.f90:409-414, :431-436)
        int ix = 0, jy = 0;
        if (i >= nx - 4) ix = i - mx;
        else if (i <= 5) ix = i + mx;
        if (three_d) {
            if (j >= ny - 4) jy = j - my;
            else if (j <= 5) jy = j + my;
        }
        double *qo = a.qn + kb;
        if (ix) qo[(size_t)(ix - 1) + sy * (j - 1)] = qn;
        if (jy) {
            qo[(size_t)(i - 1) + sy * (jy - 1)] = qn;
            if (ix) qo[(size_t)(ix - 1) + sy * (jy - 1)] = qn;
        }
    }
}

}  // namespace

hipError_t launch_positive_stage(const GridDev &g, const SchemeParams &sp, const PosArgs &a, hipStream_t s)
{
    const size_t n3 = (size_t)g.nx * g.ny * g.nz;
    const int blocks = (int)std::min<size_t>((n3 + 255) / 256, (size_t)256 * 64);
    hipLaunchKernelGGL(pos_flux_kernel, dim3(blocks), dim3(256), 0, s, g, sp, a);
    hipLaunchKernelGGL(pos_theta_kernel, dim3(blocks), dim3(256), 0, s, g, a);
    hipLaunchKernelGGL(pos_update_kernel, dim3(blocks), dim3(256), 0, s, g, a);
    return hipGetLastError();
}

hipError_t launch_march(const GridDev &g, const SchemeParams &sp, const ScalarArgs &a, int kchunk, hipStream_t s)
{
    switch (sp.scheme) {
    case 0: return launch_dims<0, 32, 8, 8>(g, sp, a, kchunk, s);
    case 3: return launch_dims<3, 16, 16, 16>(g, sp, a, kchunk, s);
    case 2: return launch_dims<2, 32, 8, 8>(g, sp, a, kchunk, s);
    case 4: return launch_dims<4, 32, 8, 8>(g, sp, a, kchunk, s);
    case 5: return launch_dims<5, 32, 8, 8>(g, sp, a, kchunk, s);
    case 6: return launch_dims<6, 32, 8, 8>(g, sp, a, kchunk, s);
    case 7: return launch_dims<7, 32, 8, 8>(g, sp, a, kchunk, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace mimadv
