// adjoint.hip -- hand-written adjoint of the reference WENO advection step.
//
// Forward (one Shu-Osher stage of the plain FP64 path, flux form, periodic
// sides, no-flux floor and lid):
//   qn = c0*q0 + c1*(q + dt*L(q; u, v, w)),   L = -div F,
//   F_f = u_f * R_f(q_{f-1}, q_f, q_{f+1}, q_{f+2})
// with R the reference 'weno' face value (advq_weno.f90: candidates,
// limiter_weno rescale about the upwind-left cell, ratio-20 nonlinear
// weights, vertical closures).  Inputs are taken with periodic ghost
// images; derivatives are with respect to the interior values.
//
// Adjoint of the stage, given qbn = dJ/dqn and lam = ls*qbn:
//   out   (=|+=)  ic*qbn + (dL/dq)^T lam
//   acc0  (=|+=)  c0*qbn
//   ub/vb/wb  +=  (dL/d(u,v,w))^T lam
// (ic = c1, ls = c1*dt for a stage; ic = 0, ls = 1 for the plain
// vector-Jacobian product of L.)  Per face the adjoint of the flux is
// Fb_f = lam_{f+1}/h_{f+1} - lam_f/h_f (1/h the divergence metric), which
// gives Fb_f*u_f*dR_f/dq to the four stencil cells and Fb_f*R_f to u_f.
//
// The face derivative is written in reverse mode: one value sweep, one
// adjoint sweep, the nonlinear weights differentiated exactly through
// g/(eps+eta)^2 and the normalisation.  One-sided choices at the kinks:
//   * min() of the limiter candidates: the candidate the value selects
//     (ties: the later argument);
//   * phi = min(|r|, 1): zero derivative on the clamped side (|r| >= 1),
//     sign(0) = +1 in |r|;
//   * the ratio-20 switch: derivative of the branch taken;
//   * u_f = 0: the u >= 0 side.
//
// Work mapping: a workgroup owns a TX x TY column tile and a chunk of
// levels and marches up it.  Per level the x and y faces that touch the
// tile (TX+3 per row, TY+3 per column) are evaluated once each and their
// four stencil contributions are gathered through LDS; the z face k+1/2 is
// evaluated by its column's thread and scattered into a register window of
// the partial adjoints of levels k-1..k+2, level k-1 being complete after
// it.  A chunk starts two levels early and ends one level late (z faces
// only) so that its levels receive every z face.  Each velocity face is
// accumulated by the one workgroup that owns it: no atomics, and the
// result does not depend on scheduling.
#include <algorithm>

#include "mimadv_internal.hpp"
#include "weno_ratio.hpp"

namespace mimadv {

namespace {

// 1/x from the hardware estimate and two Newton steps (within an ulp or
// two; 0 and infinities give NaN, which the callers treat like the
// reference's infinite or undefined quotient).
__device__ __forceinline__ double rcp_nr(double x)
{
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

// Face value R and Rb * dR/d(a, b, c, d).  kind: 0 loop body (limiter and,
// when nonlin, the ratio-20 weights), 2 top closure (no limiter).
__device__ __forceinline__ double face_vjp(double a, double b, double c, double d, bool upos, int kind, bool nonlin,
                                           double Rb, double &ga, double &gb, double &gc, double &gd)
{
    const double eps = 1.e-15, g1 = 1.0 / 3.0, g2 = 2.0 / 3.0;
    const double f1 = -0.5 * a + 1.5 * b, f2 = 0.5 * b + 0.5 * c, f4 = 1.5 * c - 0.5 * d;
    double f1b = 0.0, f2b = 0.0, f4b = 0.0;
    double ab = 0.0, bb = 0.0, cb = 0.0, db = 0.0;
    double R;
    if (kind == 2) {
        if (upos) {
            R = g1 * f1 + g2 * f2;
            f1b = g1 * Rb;
            f2b = g2 * Rb;
        } else {
            R = f2;
            f2b = Rb;
        }
    } else {
        // limiter_weno: phi = min(|-b / (min(f1, f2, f2, f4) - b + eps)|, 1)
        int sel = f1 < f2 ? 1 : 2;
        double m = sel == 1 ? f1 : f2;
        if (!(m < f2)) {
            sel = 2;
            m = f2;
        }
        if (!(m < f4)) {
            sel = 4;
            m = f4;
        }
        const double rden = rcp_nr(m - b + eps);
        const double r = -b * rden;
        const double ar = r < 0.0 ? -r : r;
        const bool clamped = !(ar < 1.0);
        const double phi = clamped ? 1.0 : ar;
        const double fo = upos ? f1 : f4;
        const double fo_r = phi * (fo - b) + b, f2_r = phi * (f2 - b) + b;
        double wo = g1, wi = g2;
        if (nonlin) {
            const double to = upos ? b - a : d - c, ti = c - b;
            const double eo = to * to, ei = ti * ti;
            const double mx = eo > ei ? eo : ei, mn = eo < ei ? eo : ei;
            if (ratio_above_20(mx, mn + eps)) {
                // wo = (g1/Q)/(g1/Q + g2/P) = P/(P + 2Q), wi = 2Q/(P + 2Q)
                // with Q = (eps + eo)^2, P = (eps + ei)^2
                const double po = eps + eo, pi = eps + ei;
                const double Q = po * po, P = pi * pi;
                const double rS = rcp_nr(P + 2.0 * Q);
                wo = P * rS;
                wi = 2.0 * Q * rS;
                // dR/dwo = fo_r, dR/dwi = f2_r; dwo/dP = -dwi/dP = 2Q/S^2,
                // dwo/dQ = -dwi/dQ = -2P/S^2; dP/dti = 4 pi ti, dQ/dto = 4 po to
                const double t = 2.0 * Rb * (fo_r - f2_r) * rS * rS;
                const double tob = -4.0 * t * P * po * to;
                const double tib = 4.0 * t * Q * pi * ti;
                if (upos) {
                    bb += tob;
                    ab -= tob;
                } else {
                    db += tob;
                    cb -= tob;
                }
                cb += tib;
                bb -= tib;
            }
        }
        R = wo * fo_r + wi * f2_r;
        const double fo_rb = Rb * wo, f2_rb = Rb * wi;
        if (upos) f1b += fo_rb * phi;
        else f4b += fo_rb * phi;
        f2b += f2_rb * phi;
        bb += (fo_rb + f2_rb) * (1.0 - phi);
        if (!clamped) {
            const double phib = fo_rb * (fo - b) + f2_rb * (f2 - b);
            const double rb = r < 0.0 ? -phib : phib;
            // r = -b/den, den = m - b + eps
            const double denb = -rb * r * rden;
            bb += -rb * rden - denb;
            if (sel == 1) f1b += denb;
            else if (sel == 2) f2b += denb;
            else f4b += denb;
        }
    }
    ga = ab - 0.5 * f1b;
    gb = bb + 1.5 * f1b + 0.5 * f2b;
    gc = cb + 0.5 * f2b + 1.5 * f4b;
    gd = db - 0.5 * f4b;
    return R;
}

// WENO5-JS value at the right face of c from a..e (Jiang & Shu 1996:
// candidates (2,-7,11)/6, (-1,5,2)/6, (2,5,-1)/6, linear weights 0.1, 0.6,
// 0.3, eps = 1e-6, p = 2) and Rb * dR/d(a..e) added to g[0..4].  With
// alpha_r = gamma_r/(eps + beta_r)^2 and S = sum alpha: dR/dq_r = alpha_r/S,
// dR/dalpha_r = (q_r - R)/S, dalpha_r/dbeta_r = -2 alpha_r/(eps + beta_r).
__device__ __forceinline__ double weno5_vjp(double a, double b, double c, double d, double e, double Rb, double *g)
{
    const double eps = 1e-6, sixth = 1.0 / 6.0, k13 = 13.0 / 12.0;
    const double q0 = (2.0 * a - 7.0 * b + 11.0 * c) * sixth;
    const double q1 = (-b + 5.0 * c + 2.0 * d) * sixth;
    const double q2 = (2.0 * c + 5.0 * d - e) * sixth;
    const double s0 = a - 2.0 * b + c, t0 = a - 4.0 * b + 3.0 * c;
    const double s1 = b - 2.0 * c + d, t1 = b - d;
    const double s2 = c - 2.0 * d + e, t2 = 3.0 * c - 4.0 * d + e;
    const double p0 = eps + (k13 * s0 * s0 + 0.25 * t0 * t0);
    const double p1 = eps + (k13 * s1 * s1 + 0.25 * t1 * t1);
    const double p2 = eps + (k13 * s2 * s2 + 0.25 * t2 * t2);
    const double r0 = rcp_nr(p0), r1 = rcp_nr(p1), r2 = rcp_nr(p2);
    const double a0 = 0.1 * r0 * r0, a1 = 0.6 * r1 * r1, a2 = 0.3 * r2 * r2;
    const double rS = rcp_nr(a0 + a1 + a2);
    const double R = (a0 * q0 + a1 * q1 + a2 * q2) * rS;
    const double qb0 = Rb * a0 * rS, qb1 = Rb * a1 * rS, qb2 = Rb * a2 * rS;
    // beta adjoints: Rb (q_r - R)/S * (-2 alpha_r/(eps + beta_r))
    const double bb0 = -2.0 * Rb * (q0 - R) * rS * a0 * r0;
    const double bb1 = -2.0 * Rb * (q1 - R) * rS * a1 * r1;
    const double bb2 = -2.0 * Rb * (q2 - R) * rS * a2 * r2;
    const double sb0 = bb0 * (2.0 * k13) * s0, tb0 = bb0 * 0.5 * t0;
    const double sb1 = bb1 * (2.0 * k13) * s1, tb1 = bb1 * 0.5 * t1;
    const double sb2 = bb2 * (2.0 * k13) * s2, tb2 = bb2 * 0.5 * t2;
    g[0] += (2.0 * sixth) * qb0 + sb0 + tb0;
    g[1] += (-7.0 * sixth) * qb0 - sixth * qb1 - 2.0 * sb0 - 4.0 * tb0 + sb1 + tb1;
    g[2] += (11.0 * sixth) * qb0 + (5.0 * sixth) * qb1 + (2.0 * sixth) * qb2 + sb0 + 3.0 * tb0 - 2.0 * sb1 + sb2 +
            3.0 * tb2;
    g[3] += (2.0 * sixth) * qb1 + (5.0 * sixth) * qb2 + sb1 - tb1 - 2.0 * sb2 - 4.0 * tb2;
    g[4] += -sixth * qb2 + sb2 + tb2;
    return R;
}

// WENO5 face of the 6-cell stencil s[0..5] = q(f-2..f+3): the upwind
// reconstruction (left from s[0..4] for u >= 0, right from s[5..1]) and
// Rb * dR/ds added to g[0..5].
__device__ __forceinline__ double face5_vjp(const double *s, bool upos, double Rb, double *g)
{
    if (upos) return weno5_vjp(s[0], s[1], s[2], s[3], s[4], Rb, g);
    double h[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const double R = weno5_vjp(s[5], s[4], s[3], s[2], s[1], Rb, h);
    g[5] += h[0];
    g[4] += h[1];
    g[3] += h[2];
    g[2] += h[3];
    g[1] += h[4];
    return R;
}

// periodic interior image of a 1-based index (period m, interior from 4)
__device__ __forceinline__ int perio(int i, int n, int m)
{
    if (i >= 4 && i <= n - 2) return i;
    int r = (i - 4) % m;
    return (r < 0 ? r + m : r) + 4;
}

// S5: the WENO5-JS extension (6-cell face stencils q(f-2..f+3) in x and y,
// in z where the upwind stencil lies in the column); otherwise the
// reference 'weno' (4-cell stencils q(f-1..f+2)).
template <int TX, int TY, bool S5>
__global__ __launch_bounds__(TX *TY) __attribute__((amdgpu_waves_per_eu(S5 ? 2 : 4, 8))) void adjoint_kernel(
    AdjArgs A)
{
    constexpr bool THREE_D = TY > 1;
    constexpr int NT = TX * TY;
    constexpr int NS = S5 ? 6 : 4;          // stencil cells of a face
    constexpr int QHL = S5 ? 5 : 3;         // q tile halo
    constexpr int LHL = S5 ? 3 : 2;         // lam tile halo
    constexpr int FO = S5 ? 3 : 2;          // first face of a row: i0 - FO (+1/2)
    constexpr int QW = TX + 2 * QHL, QH = THREE_D ? TY + 2 * QHL : 1;
    constexpr int LW = TX + 2 * LHL, LH = THREE_D ? TY + 2 * LHL : 1;
    constexpr int NFX = TX + NS - 1, NFY = THREE_D ? TY + NS - 1 : 0;  // faces per row / column
    constexpr int NXF = TY * NFX, NYF = NFY * TX, NF = NXF + NYF;
    constexpr int QS = (QH * QW + NT - 1) / NT, LS = (LH * LW + NT - 1) / NT, FS = (NF + NT - 1) / NT;
    __shared__ double qt[QH][QW];
    __shared__ double lt[LH][LW];
    __shared__ double cf[NS][NF];  // stencil contributions per face, slot-major
    // plane offsets of this thread's tile values and face velocities, fixed
    // for the work item: [q tile (QS) | lam tile (LS) | faces (FS)] x NT
    __shared__ unsigned slots[(QS + LS + FS) * NT];

    const int nx = A.nx, ny = A.ny, nz = A.nz, mx = nx - 5, my = ny - 5;
    const size_t plane = (size_t)nx * ny;
    const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
    const int nchunk = (nz + A.kchunk - 1) / A.kchunk;

    // one work item per workgroup (the grid is nwork groups)
    if (const int wi = blockIdx.x; wi < A.nwork) {
        const int tile = wi / nchunk, kc = wi - tile * nchunk;
        const int bx = tile % A.ntx, by = tile / A.ntx;
        const int i0 = 4 + bx * TX, j0 = THREE_D ? 4 + by * TY : 1;
        const int k0 = 1 + kc * A.kchunk, k1 = min(nz, k0 + A.kchunk - 1);
        const int i = i0 + tx, j = j0 + ty;
        const bool valid = i <= nx - 2 && (!THREE_D || j <= ny - 2);
        const int ip = perio(i, nx, mx), jp = THREE_D ? perio(j, ny, my) : 1;
        const size_t col = (size_t)(ip - 1) + (size_t)nx * (jp - 1);
#pragma unroll
        for (int s = 0; s < QS; ++s) {
            const int t = tid + s * NT;
            unsigned off = 0;
            if (t < QH * QW) {
                const int r = t / QW, c = t - r * QW;
                const int ii = perio(i0 - QHL + c, nx, mx), jj = THREE_D ? perio(j0 - QHL + r, ny, my) : 1;
                off = (unsigned)((ii - 1) + nx * (jj - 1));
            }
            slots[s * NT + tid] = off;
        }
#pragma unroll
        for (int s = 0; s < LS; ++s) {
            const int t = tid + s * NT;
            unsigned off = 0;
            if (t < LH * LW) {
                const int r = t / LW, c = t - r * LW;
                const int ii = perio(i0 - LHL + c, nx, mx), jj = THREE_D ? perio(j0 - LHL + r, ny, my) : 1;
                off = (unsigned)((ii - 1) + nx * (jj - 1));
            }
            slots[(QS + s) * NT + tid] = off;
        }
#pragma unroll
        for (int s = 0; s < FS; ++s) {
            const int e = tid + s * NT;
            unsigned off = 0;
            if (!THREE_D || e < NXF) {
                if (e < NXF) {
                    // x face f + 1/2 (velocity u(f+1)), f = i0-FO+fl, of tile row r
                    const int r = e / NFX, fl = e - r * NFX, f = i0 - FO + fl;
                    const int jr = THREE_D ? perio(j0 + r, ny, my) : 1;
                    off = (unsigned)((perio(f + 1, nx, mx) - 1) + nx * (jr - 1));
                }
            } else if (e < NF) {
                // y face g + 1/2 (velocity v(g+1)), g = j0-FO+fr, of tile column c
                const int e2 = e - NXF, fr = e2 / TX, c = e2 - fr * TX, g = j0 - FO + fr;
                off = (unsigned)((perio(i0 + c, nx, mx) - 1) + nx * (perio(g + 1, ny, my) - 1));
            }
            slots[(QS + LS + s) * NT + tid] = off;
        }
        auto q_off = [&](int s) { return slots[s * NT + tid]; };
        auto l_off = [&](int s) { return slots[(QS + s) * NT + tid]; };
        auto f_off = [&](int s) { return slots[(QS + LS + s) * NT + tid]; };

        // partial adjoints of levels k-ZL .. k-ZL+NS-1 (the levels the z
        // face k+1/2 touches)
        constexpr int ZL = S5 ? 2 : 1;
        double acc[NS];
#pragma unroll
        for (int m = 0; m < NS; ++m) acc[m] = 0.0;
        auto finalize = [&](int m, double val, double qb) {
            if (!valid || m < k0 || m > k1) return;
            const size_t id = col + (size_t)(m - 1) * plane;
            const double o = A.ic * qb + val;
            if (A.accumulate) A.out[id] += o;
            else A.out[id] = o;
            if (A.acc0_mode == 1) A.acc0[id] = A.c0 * qb;
            else if (A.acc0_mode == 2) A.acc0[id] += A.c0 * qb;
        };
        // z faces f touching the chunk's levels: k0-NS/2 <= f <= k1+NS/2-1
        const int ks = max(1, k0 - NS / 2), ke = min(nz, k1 + NS / 2 - 1);
        // column values at level kk (clamped to 1..nz; q(0) := q(1) is the
        // floor convention of the first face)
        auto Qz = [&](int kk) { return A.q[col + (size_t)(min(max(kk, 1), nz) - 1) * plane]; };
        auto QBz = [&](int kk) { return A.qbn[col + (size_t)(min(max(kk, 1), nz) - 1) * plane]; };
        // register windows of iteration k: q(k-ZL..k-ZL+NS-1), qbn(k-ZL..k+1), w(k+1)
        double zq[NS], qbw[ZL + 2];
#pragma unroll
        for (int m = 0; m < NS; ++m) zq[m] = Qz(ks - ZL + m);
#pragma unroll
        for (int m = 0; m < ZL + 2; ++m) qbw[m] = QBz(ks - ZL + m);
        double wz = ks <= nz - 1 ? A.w[col + (size_t)ks * plane] : 0.0;
        // tile values and face velocities of a chunk level, requested one
        // level ahead
        double qv[QS], lv[LS], fv[FS];
        auto fetch = [&](int k) {
            const size_t kb = (size_t)(k - 1) * plane;
            const double *qk = A.q + kb, *lk = A.qbn + kb, *uk = A.u + kb, *vk = A.v + kb;
#pragma unroll
            for (int s = 0; s < QS; ++s)
                if (tid + s * NT < QH * QW) qv[s] = qk[q_off(s)];
#pragma unroll
            for (int s = 0; s < LS; ++s)
                if (tid + s * NT < LH * LW) lv[s] = lk[l_off(s)];
#pragma unroll
            for (int s = 0; s < FS; ++s) {
                const int e = tid + s * NT;
                fv[s] = (!THREE_D || e < NXF) ? (e < NXF ? uk[f_off(s)] : 0.0) : (e < NF ? vk[f_off(s)] : 0.0);
            }
        };
        fetch(k0);
#pragma unroll 1
        for (int k = ks; k <= ke; ++k) {
            const size_t kb = (size_t)(k - 1) * plane;
            // the next iteration's column values, requested now
            double zn = 0.0, qbn1 = 0.0, wn = 0.0;
            if (k < ke) {
                zn = Qz(k + NS - ZL);
                qbn1 = QBz(k + 2);
                wn = k + 1 <= nz - 1 ? A.w[col + (size_t)(k + 1) * plane] : 0.0;
            }
            if (k >= k0 && k <= k1) {
#pragma unroll
                for (int s = 0; s < QS; ++s)
                    if (tid + s * NT < QH * QW) (&qt[0][0])[tid + s * NT] = qv[s];
#pragma unroll
                for (int s = 0; s < LS; ++s)
                    if (tid + s * NT < LH * LW) (&lt[0][0])[tid + s * NT] = A.ls * lv[s];
                double vc[FS];
#pragma unroll
                for (int s = 0; s < FS; ++s) vc[s] = fv[s];
                __syncthreads();
                if (k < k1) fetch(k + 1);
                // the velocity adjoints of the faces this thread owns, requested
                // before the face arithmetic that updates them
                double vbo[FS];
                bool own[FS];
#pragma unroll
                for (int s = 0; s < FS; ++s) {
                    const int e = tid + s * NT;
                    own[s] = false;
                    vbo[s] = 0.0;
                    if (e >= NF) continue;
                    if (!THREE_D || e < NXF) {
                        const int r = e / NFX, f = i0 - FO + (e - r * NFX);
                        own[s] = f >= i0 && f < i0 + TX && f <= nx - 2 && (!THREE_D || j0 + r <= ny - 2);
                        if (own[s]) vbo[s] = A.ub[kb + f_off(s)];
                    } else {
                        const int e2 = e - NXF, fr = e2 / TX, c = e2 - fr * TX, g = j0 - FO + fr;
                        own[s] = g >= j0 && g < j0 + TY && g <= ny - 2 && i0 + c <= nx - 2;
                        if (own[s]) vbo[s] = A.vb[kb + f_off(s)];
                    }
                }
#pragma unroll
                for (int s = 0; s < FS; ++s) {
                    const int e = tid + s * NT;
                    if (e >= NF) continue;
                    double gs[NS];
#pragma unroll
                    for (int m = 0; m < NS; ++m) gs[m] = 0.0;
                    if (!THREE_D || e < NXF) {
                        // x face f + 1/2 (velocity u(f+1)), f = i0-FO+fl, of tile row r
                        const int r = e / NFX, fl = e - r * NFX;
                        const double uf = vc[s];
                        const int lr = THREE_D ? r + LHL : 0, qr = THREE_D ? r + QHL : 0;
                        const double Fb = (lt[lr][fl + 1] - lt[lr][fl]) * A.dx1;
                        double R;
                        if constexpr (S5) {
                            double sv[NS];
#pragma unroll
                            for (int m = 0; m < NS; ++m) sv[m] = qt[qr][fl + m];
                            R = face5_vjp(sv, !(uf < 0.0), Fb * uf, gs);
                        } else {
                            R = face_vjp(qt[qr][fl], qt[qr][fl + 1], qt[qr][fl + 2], qt[qr][fl + 3], !(uf < 0.0), 0,
                                         true, Fb * uf, gs[0], gs[1], gs[2], gs[3]);
                        }
                        if (own[s]) A.ub[kb + f_off(s)] = vbo[s] + Fb * R;
                    } else if constexpr (THREE_D) {
                        // y face g + 1/2 (velocity v(g+1)), g = j0-FO+fr, of tile column c
                        const int e2 = e - NXF, fr = e2 / TX, c = e2 - fr * TX;
                        const double vf = vc[s];
                        const double Fb = (lt[fr + 1][c + LHL] - lt[fr][c + LHL]) * A.dy1;
                        double R;
                        if constexpr (S5) {
                            double sv[NS];
#pragma unroll
                            for (int m = 0; m < NS; ++m) sv[m] = qt[fr + m][c + QHL];
                            R = face5_vjp(sv, !(vf < 0.0), Fb * vf, gs);
                        } else {
                            R = face_vjp(qt[fr][c + 3], qt[fr + 1][c + 3], qt[fr + 2][c + 3], qt[fr + 3][c + 3],
                                         !(vf < 0.0), 0, true, Fb * vf, gs[0], gs[1], gs[2], gs[3]);
                        }
                        if (own[s]) A.vb[kb + f_off(s)] = vbo[s] + Fb * R;
                    }
#pragma unroll
                    for (int m = 0; m < NS; ++m) cf[m][e] = gs[m];
                }
                __syncthreads();
                // cell (tx, ty) is stencil slot m of x face i+NS/2-1-m+1/2
                const int ex = ty * NFX + tx;
                double xy = 0.0;
#pragma unroll
                for (int m = 0; m < NS; ++m) xy += cf[m][ex + NS - 1 - m];
                if constexpr (THREE_D) {
                    const int ey = NXF + ty * TX + tx;
#pragma unroll
                    for (int m = 0; m < NS; ++m) xy += cf[m][ey + (NS - 1 - m) * TX];
                }
                acc[ZL] += xy;
                // no barrier before the next level's tile writes: qt, lt and
                // cf are read before the barrier above or by the gather,
                // and cf is rewritten only after the next level's first
                // barrier, which every thread reaches after its gather
            }
            // z face k+1/2 (velocity w(k+1)); no flux through floor and lid
            if (k <= nz - 1) {
                const double Fb = A.ls * (qbw[ZL + 1] * A.dz1w[k] - qbw[ZL] * A.dz1w[k - 1]);
                const bool upos = !(wz < 0.0);
                double R;
                if (S5 && wz > 0.0 && k >= 3 && k <= nz - 2) {
                    // WENO5 from q(k-2..k+2): levels k-2..k+2 are acc[0..4]
                    double h[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
                    R = weno5_vjp(zq[0], zq[1], zq[2], zq[3], zq[4], Fb * wz, h);
#pragma unroll
                    for (int m = 0; m < 5; ++m) acc[m] += h[m];
                } else if (S5 && wz < 0.0 && k >= 2 && k <= nz - 3) {
                    // WENO5 from q(k+3..k-1): levels k+3..k-1 are acc[5..1]
                    double h[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
                    R = weno5_vjp(zq[NS - 1], zq[4], zq[3], zq[2], zq[1], Fb * wz, h);
#pragma unroll
                    for (int m = 0; m < 5; ++m) acc[NS - 1 - m] += h[m];
                } else {
                    // the reference face of q(k-1..k+2)
                    double ga, gb, gc, gd;
                    if (k == nz - 1)
                        R = face_vjp(zq[ZL - 1], zq[ZL], zq[ZL + 1], 0.0, upos, 2, false, Fb * wz, ga, gb, gc, gd);
                    else
                        R = face_vjp(zq[ZL - 1], zq[ZL], zq[ZL + 1], zq[ZL + 2], upos, 0, k >= 3 && k <= nz - 3,
                                     Fb * wz, ga, gb, gc, gd);
                    // the first face reads x(0) as x(1)
                    if (k == 1) acc[ZL] += ga;
                    else acc[ZL - 1] += ga;
                    acc[ZL] += gb;
                    acc[ZL + 1] += gc;
                    acc[ZL + 2] += gd;
                }
                if (valid && k >= k0 && k <= k1) A.wb[col + (size_t)k * plane] += Fb * R;
            }
            finalize(k - ZL, acc[0], qbw[0]);
#pragma unroll
            for (int m = 0; m < NS - 1; ++m) acc[m] = acc[m + 1];
            acc[NS - 1] = 0.0;
#pragma unroll
            for (int m = 0; m < NS - 1; ++m) zq[m] = zq[m + 1];
            zq[NS - 1] = zn;
#pragma unroll
            for (int m = 0; m < ZL + 1; ++m) qbw[m] = qbw[m + 1];
            qbw[ZL + 1] = qbn1;
            wz = wn;
        }
        if (ke == nz) {
#pragma unroll
            for (int m = 0; m < ZL; ++m) finalize(nz - ZL + 1 + m, acc[m], qbw[m]);
        }
    }
}

template <int TX, int TY, bool S5>
hipError_t launch_tile(AdjArgs A, hipStream_t s)
{
    A.ntx = (A.nx - 5 + TX - 1) / TX;
    A.nty = TY > 1 ? (A.ny - 5 + TY - 1) / TY : 1;
    const int nchunk = (A.nz + A.kchunk - 1) / A.kchunk;
    A.nwork = A.ntx * A.nty * nchunk;
    hipLaunchKernelGGL((adjoint_kernel<TX, TY, S5>), dim3(A.nwork), dim3(TX * TY), 0, s, A);
    return hipGetLastError();
}

__global__ void interior_copy_kernel(int nx, int ny, size_t n, const double *src, double *dst)
{
    for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < n; id += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(id % nx) + 1, j = (int)((id / nx) % ny) + 1;
        const bool in = i >= 4 && i <= nx - 2 && (ny == 1 || (j >= 4 && j <= ny - 2));
        dst[id] = in ? src[id] : 0.0;
    }
}

}  // namespace

hipError_t launch_interior_copy(const GridDev &g, const double *src, double *dst, hipStream_t s)
{
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(interior_copy_kernel, dim3(blocks), dim3(256), 0, s, g.nx, g.ny, n, src, dst);
    return hipGetLastError();
}

hipError_t launch_adjoint_stage(const AdjArgs &a0, hipStream_t s)
{
    AdjArgs A = a0;
    if (A.kchunk <= 0) A.kchunk = 16;
    if (A.kchunk > A.nz) A.kchunk = A.nz;
    if (A.weno5) {
        // 6-cell stencils: a 16x16 tile has the fewest faces per cell
        // (672 for 256 cells), so its face contributions, tiles and slot
        // offsets take 49.7 KB of LDS and three groups share a CU
        if (A.ny > 1) return launch_tile<16, 16, true>(A, s);
        return launch_tile<64, 1, true>(A, s);
    }
    if (A.ny > 1) return launch_tile<64, 8, false>(A, s);
    return launch_tile<64, 1, false>(A, s);
}

}  // namespace mimadv
