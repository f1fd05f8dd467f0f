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

// periodic interior image of a 1-based index (period m, interior from 4)
__device__ __forceinline__ int perio(int i, int n, int m)
{
    if (i >= 4 && i <= n - 2) return i;
    int r = (i - 4) % m;
    return (r < 0 ? r + m : r) + 4;
}

template <int TX, int TY>
__global__ __launch_bounds__(TX *TY) __attribute__((amdgpu_waves_per_eu(4, 8))) void adjoint_kernel(AdjArgs A)
{
    constexpr bool THREE_D = TY > 1;
    constexpr int NT = TX * TY;
    constexpr int QW = TX + 6, QH = THREE_D ? TY + 6 : 1;   // q tile, halo 3
    constexpr int LW = TX + 4, LH = THREE_D ? TY + 4 : 1;   // lam tile, halo 2
    constexpr int NFX = TX + 3, NFY = THREE_D ? TY + 3 : 0; // faces per row / column
    constexpr int NXF = TY * NFX, NYF = NFY * TX, NF = NXF + NYF;
    constexpr int QS = (QH * QW + NT - 1) / NT, LS = (LH * LW + NT - 1) / NT, FS = (NF + NT - 1) / NT;
    __shared__ double qt[QH][QW];
    __shared__ double lt[LH][LW];
    __shared__ double cf[4][NF];  // stencil contributions per face, slot-major
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
                const int ii = perio(i0 - 3 + c, nx, mx), jj = THREE_D ? perio(j0 - 3 + r, ny, my) : 1;
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
                const int ii = perio(i0 - 2 + c, nx, mx), jj = THREE_D ? perio(j0 - 2 + r, ny, my) : 1;
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
                    // x face f + 1/2 (velocity u(f+1)), f = i0-2+fl, of tile row r
                    const int r = e / NFX, fl = e - r * NFX, f = i0 - 2 + fl;
                    const int jr = THREE_D ? perio(j0 + r, ny, my) : 1;
                    off = (unsigned)((perio(f + 1, nx, mx) - 1) + nx * (jr - 1));
                }
            } else if (e < NF) {
                // y face g + 1/2 (velocity v(g+1)), g = j0-2+fr, of tile column c
                const int e2 = e - NXF, fr = e2 / TX, c = e2 - fr * TX, g = j0 - 2 + fr;
                off = (unsigned)((perio(i0 + c, nx, mx) - 1) + nx * (perio(g + 1, ny, my) - 1));
            }
            slots[(QS + LS + s) * NT + tid] = off;
        }
        auto q_off = [&](int s) { return slots[s * NT + tid]; };
        auto l_off = [&](int s) { return slots[(QS + s) * NT + tid]; };
        auto f_off = [&](int s) { return slots[(QS + LS + s) * NT + tid]; };

        double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;  // levels k-1 .. k+2
        auto finalize = [&](int m, double val, double qb) {
            if (!valid || m < k0 || m > k1) return;
            const size_t id = col + (size_t)(m - 1) * plane;
            const double o = A.ic * qb + val;
            if (A.accumulate) A.out[id] += o;
            else A.out[id] = o;
            if (A.acc0_mode == 1) A.acc0[id] = A.c0 * qb;
            else if (A.acc0_mode == 2) A.acc0[id] += A.c0 * qb;
        };
        const int ks = max(1, k0 - 2), ke = min(nz, k1 + 1);
        // column values at level kk (clamped to 1..nz; q(0) := q(1) is the
        // floor convention of the first face)
        auto Qz = [&](int kk) { return A.q[col + (size_t)(min(max(kk, 1), nz) - 1) * plane]; };
        auto QBz = [&](int kk) { return A.qbn[col + (size_t)(min(max(kk, 1), nz) - 1) * plane]; };
        // register windows of iteration k: q(k-1..k+2), qbn(k-1..k+1), w(k+1)
        double z0 = Qz(ks - 1), z1 = Qz(ks), z2 = Qz(ks + 1), z3 = Qz(ks + 2);
        double qbm = QBz(ks - 1), qb0 = QBz(ks), qb1 = QBz(ks + 1);
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
                zn = Qz(k + 3);
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
                        const int r = e / NFX, f = i0 - 2 + (e - r * NFX);
                        own[s] = f >= i0 && f < i0 + TX && f <= nx - 2 && (!THREE_D || j0 + r <= ny - 2);
                        if (own[s]) vbo[s] = A.ub[kb + f_off(s)];
                    } else {
                        const int e2 = e - NXF, fr = e2 / TX, c = e2 - fr * TX, g = j0 - 2 + fr;
                        own[s] = g >= j0 && g < j0 + TY && g <= ny - 2 && i0 + c <= nx - 2;
                        if (own[s]) vbo[s] = A.vb[kb + f_off(s)];
                    }
                }
#pragma unroll
                for (int s = 0; s < FS; ++s) {
                    const int e = tid + s * NT;
                    if (e >= NF) continue;
                    double ga, gb, gc, gd;
                    if (!THREE_D || e < NXF) {
                        // x face f + 1/2 (velocity u(f+1)), f = i0-2+fl, of tile row r
                        const int r = e / NFX, fl = e - r * NFX, f = i0 - 2 + fl;
                        const double uf = vc[s];
                        const int lr = THREE_D ? r + 2 : 0, qr = THREE_D ? r + 3 : 0;
                        const double Fb = (lt[lr][fl + 1] - lt[lr][fl]) * A.dx1;
                        const double R = face_vjp(qt[qr][fl], qt[qr][fl + 1], qt[qr][fl + 2], qt[qr][fl + 3],
                                                  !(uf < 0.0), 0, true, Fb * uf, ga, gb, gc, gd);
                        (void)f;
                        if (own[s]) A.ub[kb + f_off(s)] = vbo[s] + Fb * R;
                    } else if constexpr (THREE_D) {
                        // y face g + 1/2 (velocity v(g+1)), g = j0-2+fr, of tile column c
                        const int e2 = e - NXF, fr = e2 / TX, c = e2 - fr * TX, g = j0 - 2 + fr;
                        const double vf = vc[s];
                        const double Fb = (lt[fr + 1][c + 2] - lt[fr][c + 2]) * A.dy1;
                        const double R = face_vjp(qt[fr][c + 3], qt[fr + 1][c + 3], qt[fr + 2][c + 3],
                                                  qt[fr + 3][c + 3], !(vf < 0.0), 0, true, Fb * vf, ga, gb, gc, gd);
                        (void)g;
                        if (own[s]) A.vb[kb + f_off(s)] = vbo[s] + Fb * R;
                    }
                    cf[0][e] = ga;
                    cf[1][e] = gb;
                    cf[2][e] = gc;
                    cf[3][e] = gd;
                }
                __syncthreads();
                // cell (tx, ty) is slot a of face i+1/2, b of i-1/2, c of i-3/2, d of i-5/2
                const int ex = ty * NFX + tx;
                double xy = cf[0][ex + 3] + cf[1][ex + 2] + cf[2][ex + 1] + cf[3][ex];
                if constexpr (THREE_D) {
                    const int ey = NXF + ty * TX + tx;
                    xy += cf[0][ey + 3 * TX] + cf[1][ey + 2 * TX] + cf[2][ey + TX] + cf[3][ey];
                }
                acc1 += xy;
                // no barrier before the next level's tile writes: qt, lt and
                // cf are read before the barrier above or by the gather,
                // and cf is rewritten only after the next level's first
                // barrier, which every thread reaches after its gather
            }
            // z face k+1/2 (velocity w(k+1)); no flux through floor and lid
            if (k <= nz - 1) {
                const double Fb = A.ls * (qb1 * A.dz1w[k] - qb0 * A.dz1w[k - 1]);
                const bool upos = !(wz < 0.0);
                double ga, gb, gc, gd, R;
                if (k == nz - 1) R = face_vjp(z0, z1, z2, 0.0, upos, 2, false, Fb * wz, ga, gb, gc, gd);
                else R = face_vjp(z0, z1, z2, z3, upos, 0, k >= 3 && k <= nz - 3, Fb * wz, ga, gb, gc, gd);
                // the first face reads x(0) as x(1)
                if (k == 1) acc1 += ga;
                else acc0 += ga;
                acc1 += gb;
                acc2 += gc;
                acc3 += gd;
                if (valid && k >= k0 && k <= k1) A.wb[col + (size_t)k * plane] += Fb * R;
            }
            finalize(k - 1, acc0, qbm);
            acc0 = acc1;
            acc1 = acc2;
            acc2 = acc3;
            acc3 = 0.0;
            z0 = z1;
            z1 = z2;
            z2 = z3;
            z3 = zn;
            qbm = qb0;
            qb0 = qb1;
            qb1 = qbn1;
            wz = wn;
        }
        if (ke == nz) finalize(nz, acc0, qbm);
    }
}

template <int TX, int TY>
hipError_t launch_tile(AdjArgs A, hipStream_t s)
{
    A.ntx = (A.nx - 5 + TX - 1) / TX;
    A.nty = TY > 1 ? (A.ny - 5 + TY - 1) / TY : 1;
    const int nchunk = (A.nz + A.kchunk - 1) / A.kchunk;
    A.nwork = A.ntx * A.nty * nchunk;
    hipLaunchKernelGGL((adjoint_kernel<TX, TY>), dim3(A.nwork), dim3(TX * TY), 0, s, A);
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
    if (A.ny > 1) return launch_tile<64, 8>(A, s);
    return launch_tile<64, 1>(A, s);
}

}  // namespace mimadv
