// momentum_fp64.hip -- momentum flux divergence (qnadv_uvw) on gfx950.
//
// Restates src/INTEG/windadv.f90: flux-form centred differences of the
// momenta (rho0 u, rho0 v, rho0 w) advected by wind_adv, 2nd or 4th order
// (mom_ord), with the 4th-difference artificial viscosity (lavisc, avisc =
// 1/16) and the nnf/ope vertical closures.  The reference sweeps lines per
// direction and accumulates into ua in the order x,y,z (even ntime) or
// z,y,x (odd ntime, windadv.f90:103,158); here one thread owns one cell and
// adds the same per-direction increments in the same order, recomputing the
// few face fluxes it needs (centred fluxes are cheap: the kernel is bound by
// the 9 input arrays it streams).  Output arrays are overwritten everywhere
// (ghosts and w(k=1) get 0, as `ua = 0.0` at windadv.f90:97).
//
// Compiled with -ffp-contract=off to keep the reference operation order.
#include "mimadv_internal.hpp"

namespace mimadv {

namespace {

constexpr double AVISC = 1.0 / 16.0;


struct MomCoefs {
    double c1s, c2s;  // self lines (4/3, 1/3)
    double c1c, c2c;  // cross lines (9/8, 1/8)
};

template <class F>
__device__ __forceinline__ double visc5(F f, int i)
{
    return -f(i + 2) + 4.0 * f(i + 1) - 6.0 * f(i) + 4.0 * f(i - 1) - f(i - 2);
}

// advux_fd / advvy_fd: self advection along a line (windadv.f90:526-548)
template <class MF, class TF>
__device__ __forceinline__ double self_h(double t, int i, double d, const MomCoefs &c, int lavisc, MF M, TF T)
{
    double d1 = 1.0 / d, d2 = 1.0 / (2.0 * d);
    auto flux = [&](int ii) { return ((T(ii + 1) + T(ii)) * 0.5) * ((M(ii + 1) + M(ii)) * 0.5); };
    auto advf = [&](int ii) { return ((T(ii + 2) + T(ii)) * 0.5) * ((M(ii + 2) + M(ii)) * 0.5); };
    t = t + c.c1s * (flux(i) - flux(i - 1)) * d1 - c.c2s * (advf(i) - advf(i - 2)) * d2;
    if (lavisc) {
        double mu = fabs(T(i)) * d1;
        t = t - AVISC * mu * visc5(M, i);
    }
    return t;
}

// advuy_fd / advvx_fd: cross advection along a line (windadv.f90:587-620);
// T1..T4 are the advecting component at transverse offsets -1, 0, -2, +1
template <class MF, class TF1, class TF2, class TF3, class TF4>
__device__ __forceinline__ double cross_h(double t, int i, double d, const MomCoefs &c, int lavisc, MF M, TF1 T1,
                                          TF2 T2, TF3 T3, TF4 T4)
{
    double d1 = 1.0 / d, d3 = 1.0 / (3.0 * d);
    auto flux = [&](int ii) {
        double vv = c.c1c * (T1(ii + 1) + T2(ii + 1)) * 0.5 - c.c2c * (T3(ii + 1) + T4(ii + 1)) * 0.5;
        double fr = (M(ii + 1) + M(ii)) * 0.5;
        return vv * fr;
    };
    t = t + c.c1c * (flux(i) - flux(i - 1)) * d1 - c.c2c * (flux(i + 1) - flux(i - 2)) * d3;
    if (lavisc) {
        double mu = 0.25 * fabs(T1(i) + T2(i) + T1(i + 1) + T2(i + 1)) * d1;
        t = t - AVISC * mu * visc5(M, i);
    }
    return t;
}

// advuz_fd / advvz_fd (windadv.f90:656-688); m2k = level of dz1w used on the
// k = 2 boundary line (1 for u, 2 for v)
template <class MF, class TF1, class TF2, class TF3, class TF4>
__device__ __forceinline__ double cross_z(double t, int k, int nz, const GridDev &g, const MomCoefs &c, int lavisc,
                                          int m2k, MF M, TF1 T1, TF2 T2, TF3 T3, TF4 T4)
{
    const double *dzw = g.zm.dzw, *dz1w = g.zm.dz1w;
    int bc5 = g.bcl[4], bc6 = g.bcl[5];
    auto flux = [&](int kk) -> double {
        if (kk == nz) {
            if (bc6 == 2) kk = nz - 1;  // 'ope': Flux(nz) = Flux(nz-1)
            else return 0.0;            // 'nnf': 0; otherwise never set
        }
        double ww = c.c1c * (T1(kk + 1) + T2(kk + 1)) * 0.5 - c.c2c * (T3(kk + 1) + T4(kk + 1)) * 0.5;
        double fr = (dzw[kk] * M(kk + 1) + dzw[kk - 1] * M(kk)) / (dzw[kk] + dzw[kk - 1]);
        return ww * fr;
    };
    if (k >= 3 && k <= nz - 1) {
        t = t + c.c1c * (flux(k) - flux(k - 1)) * dz1w[k - 1] -
            c.c2c * (flux(k + 1) - flux(k - 2)) / (dzw[k - 2] + dzw[k - 1] + dzw[k]);
    } else if (k == 1) {
        if (bc5 == 1) t = t + (flux(1) - 0.0) * dz1w[0];
        if (bc5 == 2) t = t + (flux(1) - flux(1)) * dz1w[0];
    } else if (k == 2) {
        double m2 = dz1w[m2k - 1];
        if (bc5 == 1) t = t + c.c1c * (flux(2) - flux(1)) * m2 - c.c2c * (flux(3) - 0.0) / (dzw[0] + dzw[1] + dzw[2]);
        if (bc5 == 2)
            t = t + c.c1c * (flux(2) - flux(1)) * m2 - c.c2c * (flux(3) - flux(1)) / (dzw[0] + dzw[1] + dzw[2]);
    } else {  // k == nz
        if (bc6 == 1) t = t + (0.0 - flux(nz - 1)) * dz1w[nz - 1];
        if (bc6 == 2) t = t + (flux(nz - 1) - flux(nz - 1)) * dz1w[nz - 1];
    }
    if (lavisc && k >= 3 && k <= nz - 2) {
        double mu = 0.25 * fabs(T1(k) + T2(k) + T1(k + 1) + T2(k + 1)) * dz1w[k - 1];
        t = t - AVISC * mu * visc5(M, k);
    }
    return t;
}

// advwx_fd / advwy_fd for level k in 2..nz (windadv.f90:927-962): M(ii, kk)
// and T(ii, kk) along the line
template <class MF, class TF>
__device__ __forceinline__ double w_h(double t, int i, int k, int nz, double d, const MomCoefs &c, int lavisc, MF M,
                                      TF T)
{
    double d1 = 1.0 / d, d3 = 1.0 / (3.0 * d);
    auto flux = [&](int ii) -> double {
        double uu;
        if (k == 2) uu = (T(ii + 1, 2) + T(ii + 1, 1)) * 0.5;
        else if (k == nz) uu = (T(ii + 1, nz) + T(ii + 1, nz - 1)) * 0.5;
        else uu = c.c1c * (T(ii + 1, k) + T(ii + 1, k - 1)) * 0.5 - c.c2c * (T(ii + 1, k + 1) + T(ii + 1, k - 2)) * 0.5;
        double fr = (M(ii + 1, k) + M(ii, k)) * 0.5;
        return uu * fr;
    };
    t = t + c.c1c * (flux(i) - flux(i - 1)) * d1 - c.c2c * (flux(i + 1) - flux(i - 2)) * d3;
    if (lavisc) {
        double mu = 0.25 * fabs(T(i, k) + T(i, k - 1) + T(i + 1, k) + T(i + 1, k - 1)) * d1;
        auto Mk = [&](int ii) { return M(ii, k); };
        t = t - AVISC * mu * visc5(Mk, i);
    }
    return t;
}

// advwz_fd (windadv.f90:1077-1121)
template <class MF, class TF>
__device__ __forceinline__ double w_z(double t, int k, int nz, const GridDev &g, const MomCoefs &c, int lavisc, MF M,
                                      TF T)
{
    const double *dzw = g.zm.dzw, *dz1 = g.zm.dz1;
    int bc5 = g.bcl[4], bc6 = g.bcl[5];
    auto flux = [&](int kk) -> double {
        if (kk == nz - 1) return 0.25 * (T(nz) + T(nz - 1)) * (M(nz) + M(nz - 1));
        if (kk >= nz || kk < 1) return 0.0;
        double ww = (T(kk + 1) + T(kk)) * 0.5;
        double fr = (M(kk + 1) + M(kk)) * 0.5;
        return ww * fr;
    };
    auto advf = [&](int kk) -> double {
        if (kk == nz - 1) {
            double s = dzw[nz - 1] + dzw[nz - 2];
            if (bc6 == 1) return (0.0 + dzw[nz - 2] * T(nz - 1)) * (0.0 + dzw[nz - 2] * M(nz - 1)) / (s * s);
            if (bc6 == 2)
                return (dzw[nz - 1] * T(nz) + dzw[nz - 2] * T(nz - 1)) * (dzw[nz - 1] * M(nz) + dzw[nz - 2] * M(nz - 1)) /
                       (s * s);
            return 0.0;
        }
        if (kk >= nz || kk < 1) return 0.0;
        double ww2 = (dzw[kk] * T(kk + 2) + dzw[kk - 1] * T(kk)) / (dzw[kk] + dzw[kk - 1]);
        double fr2 = (dzw[kk] * M(kk + 2) + dzw[kk - 1] * M(kk)) / (dzw[kk] + dzw[kk - 1]);
        return ww2 * fr2;
    };
    if (k >= 3 && k <= nz - 1) {
        t = t + c.c1s * (flux(k) - flux(k - 1)) * dz1[k - 2] - c.c2s * (advf(k) - advf(k - 2)) / (dzw[k - 1] + dzw[k - 2]);
    } else if (k == 2) {
        if (bc5 == 1) t = t + c.c1s * (flux(2) - flux(1)) * dz1[0] - c.c2s * (advf(2) - 0.0) / (dzw[1] + dzw[0]);
        if (bc5 == 2) t = t + c.c1s * (flux(2) - flux(1)) * dz1[0] - c.c2s * (advf(2) - flux(1)) / (dzw[1] + dzw[0]);
    } else if (k == nz) {
        if (bc6 == 1) t = t + (0.0 - flux(nz - 1)) * dz1[nz - 2];
        if (bc6 == 2) t = t + (flux(nz - 1) - flux(nz - 1)) * dz1[nz - 2];
    }
    if (lavisc && k >= 3 && k <= nz - 2) {
        double mu = fabs(T(k)) * dz1[k - 2];
        t = t - AVISC * mu * visc5(M, k);
    }
    return t;
}

// The three momentum increments of cell (i, j, k), as qnadv_uvw accumulates
// them (uadv_fd, vadv_fd, wadv_fd, windadv.f90:82-492): directions in the
// order x, y, z (even ntime) or z, y, x (odd).  M* and T* read the momenta
// and the advecting velocity at (ii, jj, kk) as doubles.
template <class MU, class MV, class MW, class UT, class VT, class WT>
__device__ __forceinline__ void mom_cell(const GridDev &g, const MomCoefs &c, int lavisc, bool even, int i, int j,
                                         int k, MU mu, MV mv, MW mw, UT ut, VT vt, WT wt, double &ua, double &va,
                                         double &wa)
{
    const int nz = g.nz;
    const bool three_d = g.ny > 1;
    int order[3];
    if (even) { order[0] = 0; order[1] = 1; order[2] = 2; }
    else { order[0] = 2; order[1] = 1; order[2] = 0; }

    // ---- u (uadv_fd, windadv.f90:82-220)
    {
        double t = 0.0;
        for (int s = 0; s < 3; ++s) {
            int d = order[s];
            if (d == 0) {
                t = self_h(t, i, g.dx, c, lavisc, [&](int ii) { return mu(ii, j, k); },
                           [&](int ii) { return ut(ii, j, k); });
            } else if (d == 1) {
                if (!three_d) continue;
                t = cross_h(t, j, g.dy, c, lavisc, [&](int jj) { return mu(i, jj, k); },
                            [&](int jj) { return vt(i - 1, jj, k); }, [&](int jj) { return vt(i, jj, k); },
                            [&](int jj) { return vt(i - 2, jj, k); }, [&](int jj) { return vt(i + 1, jj, k); });
            } else {
                t = cross_z(t, k, nz, g, c, lavisc, 1, [&](int kk) { return mu(i, j, kk); },
                            [&](int kk) { return wt(i - 1, j, kk); }, [&](int kk) { return wt(i, j, kk); },
                            [&](int kk) { return wt(i - 2, j, kk); }, [&](int kk) { return wt(i + 1, j, kk); });
            }
        }
        ua = t;
    }
    // ---- v (vadv_fd, windadv.f90:223-357), 3D only
    va = 0.0;
    if (three_d) {
        double t = 0.0;
        for (int s = 0; s < 3; ++s) {
            int d = order[s];
            if (d == 0) {
                t = cross_h(t, i, g.dx, c, lavisc, [&](int ii) { return mv(ii, j, k); },
                            [&](int ii) { return ut(ii, j - 1, k); }, [&](int ii) { return ut(ii, j, k); },
                            [&](int ii) { return ut(ii, j - 2, k); }, [&](int ii) { return ut(ii, j + 1, k); });
            } else if (d == 1) {
                t = self_h(t, j, g.dy, c, lavisc, [&](int jj) { return mv(i, jj, k); },
                           [&](int jj) { return vt(i, jj, k); });
            } else {
                t = cross_z(t, k, nz, g, c, lavisc, 2, [&](int kk) { return mv(i, j, kk); },
                            [&](int kk) { return wt(i, j - 1, kk); }, [&](int kk) { return wt(i, j, kk); },
                            [&](int kk) { return wt(i, j - 2, kk); }, [&](int kk) { return wt(i, j + 1, kk); });
            }
        }
        va = t;
    }
    // ---- w (wadv_fd, windadv.f90:360-492): levels 2..nz
    {
        double t = 0.0;
        if (k >= 2) {
            for (int s = 0; s < 3; ++s) {
                int d = order[s];
                if (d == 0) {
                    t = w_h(t, i, k, nz, g.dx, c, lavisc, [&](int ii, int kk) { return mw(ii, j, kk); },
                            [&](int ii, int kk) { return ut(ii, j, kk); });
                } else if (d == 1) {
                    if (!three_d) continue;
                    t = w_h(t, j, k, nz, g.dy, c, lavisc, [&](int jj, int kk) { return mw(i, jj, kk); },
                            [&](int jj, int kk) { return vt(i, jj, kk); });
                } else {
                    t = w_z(t, k, nz, g, c, lavisc, [&](int kk) { return mw(i, j, kk); },
                            [&](int kk) { return wt(i, j, kk); });
                }
            }
        }
        wa = t;
    }
}

__global__ __launch_bounds__(256) void momentum_fp64_kernel(GridDev g, MomCoefs c, int lavisc, int ntime,
                                                            const double *__restrict__ mu,
                                                            const double *__restrict__ mv,
                                                            const double *__restrict__ mw,
                                                            const double *__restrict__ ut,
                                                            const double *__restrict__ vt,
                                                            const double *__restrict__ wt, double *__restrict__ ua,
                                                            double *__restrict__ va, double *__restrict__ wa)
{
    const int nx = g.nx, ny = g.ny;
    const bool three_d = ny > 1;
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > nx) return;
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
    const size_t id = (size_t)(i - 1) + sy * (j - 1) + sz * (k - 1);
    const bool interior = i >= 4 && i <= nx - 2 && (!three_d || (j >= 4 && j <= ny - 2));
    if (!interior) {
        ua[id] = 0.0;
        if (three_d) va[id] = 0.0;
        wa[id] = 0.0;
        return;
    }
    auto acc = [&](const double *p) {
        return [=](int ii, int jj, int kk) { return p[(size_t)(ii - 1) + sy * (jj - 1) + sz * (kk - 1)]; };
    };
    double tu, tv, tw;
    mom_cell(g, c, lavisc, (ntime % 2) == 0, i, j, k, acc(mu), acc(mv), acc(mw), acc(ut), acc(vt), acc(wt), tu, tv,
             tw);
    ua[id] = tu;
    if (three_d) va[id] = tv;
    wa[id] = tw;
}

// Batched momentum stage: FP32 momenta advected by the FP32 velocities of
// the batch, FP64 arithmetic on the loaded values.  epi 0: out = -(div), the
// tendency the caller of qnadv_uvw forms (momentum.f90 subtracts winda);
// epi 1: out = c0*m0 + c1*(m + dt*(-div)) rounded once to FP32, with the
// periodic images of the cell.
//
// Column-marching form: a workgroup owns a TX x TY tile of columns and a
// chunk of levels.  Every operand the stencils of level k read (the three
// momenta and the three advecting velocities at levels k-2..k+2, columns
// within two cells of the tile) sits in a five-level LDS ring of FP32
// tiles with a 2-cell halo; each level is loaded once, coalesced,
// and serves five output levels.  The per-cell formulas are those of the
// FP64 kernel (mom_cell), so results do not depend on the mapping.  One
// work item per workgroup.  Tiles of 32x8 columns keep the ring at 51.8 KB,
// so three groups share a CU; the 168-VGPR cap that gives each SIMD three
// waves spills a few values and still wins (2.9 -> 2.5 ms per stage at
// 512x512x160 against 64x4 tiles at two waves).  The ring offsets are
// recomputed per load rather than held, which halves the spills.
template <int TX, int TY>
__global__ __launch_bounds__(TX *TY) __attribute__((amdgpu_waves_per_eu(3, 8))) void momentum_march_kernel(
    GridDev g, MomCoefs c, int lavisc, MomBatchArgs a, int kchunk, int ntx, int nwork)
{
    constexpr bool THREE_D = TY > 1;
    constexpr int NT = TX * TY;
    constexpr int H = 2, LW = TX + 2 * H, LH = THREE_D ? TY + 2 * H : 1, PL = LW * LH;
    constexpr int NR = 5, TS = (PL + NT - 1) / NT;
    __shared__ float ring[6][NR][PL];  // mu, mv, mw, ut, vt, wt

    const int nx = g.nx, ny = g.ny, nz = g.nz, mx = nx - 5, my = ny - 5;
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
    const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
    const int nchunk = (nz + kchunk - 1) / kchunk;
    const float *src[6] = {a.m[0], a.m[1], a.m[2], a.ut, a.vt, a.wt};

    if (const int wi = blockIdx.x; wi < nwork) {
        const int tile = wi / nchunk, kc = wi - tile * nchunk;
        const int i0 = 4 + (tile % ntx) * TX;
        const int j0 = THREE_D ? 4 + (tile / ntx) * TY : 1;
        const int k0 = 1 + kc * kchunk, k1 = min(nz, k0 + kchunk - 1);
        const int i = i0 + tx, j = j0 + ty;
        const bool valid = i <= nx - 2 && (!THREE_D || j <= ny - 2);
        // plane offset of ring entry s of this thread (periodic images for
        // positions outside the interior; -1: no entry), recomputed where it
        // is used rather than held in registers through the level loop
        auto t_off = [&](int s) -> int {
            const int e = tid + s * NT;
            if (e >= PL) return -1;
            const int r = e / LW, cc = e - r * LW;
            // ghost columns are read as the caller filled them, like the
            // reference; positions past the array (padding of a partial
            // tile) take a periodic image that feeds no interior cell
            int ii = i0 - H + cc;
            if (ii > nx) ii = (ii - 4) % mx + 4;
            int jj = 1;
            if (THREE_D) {
                jj = j0 - H + r;
                if (jj > ny) jj = (jj - 4) % my + 4;
            }
            return (ii - 1) + nx * (jj - 1);
        };
        // level kk of the six operands into registers (zeros outside 1..nz),
        // then into its ring slot
        float pre[6][TS];
        auto fetch = [&](int kk) {
            const bool in = kk >= 1 && kk <= nz;
            const size_t kb = in ? sz * (kk - 1) : 0;
#pragma unroll
            for (int s = 0; s < TS; ++s) {
                const int off = t_off(s);
#pragma unroll
                for (int n = 0; n < 6; ++n) pre[n][s] = (in && off >= 0) ? src[n][kb + off] : 0.0f;
            }
        };
        auto store = [&](int kk) {
            const int slot = (kk + 2 * NR) % NR;
#pragma unroll
            for (int s = 0; s < TS; ++s) {
                if (tid + s * NT >= PL) continue;
#pragma unroll
                for (int n = 0; n < 6; ++n) ring[n][slot][tid + s * NT] = pre[n][s];
            }
        };
#pragma unroll 1
        for (int kk = k0 - 2; kk <= k0 + 1; ++kk) {
            fetch(kk);
            store(kk);
        }
        fetch(k0 + 2);
#pragma unroll 1
        for (int k = k0; k <= k1; ++k) {
            // level k+2 arrived during the previous level's arithmetic; the
            // slot it takes (level k-3) was last read before the barrier
            // that ended that level
            store(k + 2);
            __syncthreads();
            if (k < k1) fetch(k + 3);
            if (valid) {
                const size_t kb = sz * (k - 1);
                const size_t id = (size_t)(i - 1) + sy * (j - 1) + kb;
                // the step-start momenta, requested before the cell arithmetic
                float m0v[3] = {0.0f, 0.0f, 0.0f};
                if (a.epi != 0 && a.c0 != 0.0) {
#pragma unroll
                    for (int n = 0; n < 3; ++n)
                        if (n != 1 || THREE_D) m0v[n] = a.m0[n][id];
                }
                // ring offset of level kk (k-2 <= kk <= k+2)
                auto lev = [&](int kk) { return ((kk + 2 * NR) % NR) * PL; };
                auto acc = [&](int n) {
                    return [&, n](int ii, int jj, int kk) {
                        const int e = lev(kk) + (THREE_D ? (jj - j0 + H) * LW : 0) + (ii - i0 + H);
                        return (double)ring[n][0][e];
                    };
                };
                double d[3];
                mom_cell(g, c, lavisc, true, i, j, k, acc(0), acc(1), acc(2), acc(3), acc(4), acc(5), d[0], d[1],
                         d[2]);
                const int ce = lev(k) + (THREE_D ? (ty + H) * LW : 0) + tx + H;
#pragma unroll
                for (int n = 0; n < 3; ++n) {
                    if (n == 1 && !THREE_D) continue;
                    const double L = -d[n];
                    float qv;
                    if (a.epi == 0) {
                        qv = (float)L;
                    } else {
                        const double m = (double)ring[n][0][ce];
                        const double qn = a.c0 != 0.0 ? a.c0 * (double)m0v[n] + a.c1 * (m + a.dt * L)
                                                      : a.c1 * (m + a.dt * L);
                        qv = (float)qn;
                    }
                    float *o = a.out[n];
                    o[id] = qv;
                    if (a.epi != 0 && a.fill) {
                        // periodic images (qnbc_p4 / lbcu_p4 for the periodic case)
                        int ix = 0, jy = 0;
                        if (i >= nx - 4) ix = i - mx;
                        else if (i <= 5) ix = i + mx;
                        if (THREE_D) {
                            if (j >= ny - 4) jy = j - my;
                            else if (j <= 5) jy = j + my;
                        }
                        if (ix) o[kb + (size_t)(ix - 1) + sy * (j - 1)] = qv;
                        if (jy) {
                            o[kb + (size_t)(i - 1) + sy * (jy - 1)] = qv;
                            if (ix) o[kb + (size_t)(ix - 1) + sy * (jy - 1)] = qv;
                        }
                    }
                }
            }
            // the next level's load overwrites the ring slot of level k-2
            __syncthreads();
        }
    }
}

MomCoefs mom_coefs(int mom_ord)
{
    MomCoefs c;
    if (mom_ord == 4) {
        c.c1s = 4.0 / 3.0; c.c2s = 1.0 / 3.0;
        c.c1c = 9.0 / 8.0; c.c2c = 1.0 / 8.0;
    } else {
        c.c1s = 1.0; c.c2s = 0.0;
        c.c1c = 1.0; c.c2c = 0.0;
    }
    return c;
}

}  // namespace

hipError_t launch_momentum_fp64(const GridDev &g, int mom_ord, int lavisc, int ntime, const double *mu,
                                const double *mv, const double *mw, const double *ut, const double *vt,
                                const double *wt, double *ua, double *va, double *wa, hipStream_t s)
{
    const MomCoefs c = mom_coefs(mom_ord);
    dim3 block(256);
    dim3 grid((g.nx + 255) / 256, g.ny, g.nz);
    hipLaunchKernelGGL(momentum_fp64_kernel, grid, block, 0, s, g, c, lavisc, ntime, mu, mv, mw, ut, vt, wt, ua, va,
                       wa);
    return hipGetLastError();
}

hipError_t launch_momentum_batch(const GridDev &g, int mom_ord, int lavisc, const MomBatchArgs &a, hipStream_t s)
{
    const MomCoefs c = mom_coefs(mom_ord);
    const int kchunk = g.nz < 16 ? g.nz : 16;
    const int nchunk = (g.nz + kchunk - 1) / kchunk;
    if (g.ny > 1) {
        constexpr int TX = 32, TY = 8;
        const int ntx = (g.nx - 5 + TX - 1) / TX, nty = (g.ny - 5 + TY - 1) / TY;
        const int nwork = ntx * nty * nchunk;
        hipLaunchKernelGGL((momentum_march_kernel<TX, TY>), dim3(nwork), dim3(TX * TY), 0, s, g, c, lavisc, a, kchunk,
                           ntx, nwork);
    } else {
        constexpr int TX = 64, TY = 1;
        const int ntx = (g.nx - 5 + TX - 1) / TX;
        const int nwork = ntx * nchunk;
        hipLaunchKernelGGL((momentum_march_kernel<TX, TY>), dim3(nwork), dim3(TX), 0, s, g, c, lavisc, a, kchunk, ntx,
                           nwork);
    }
    return hipGetLastError();
}

}  // namespace mimadv
