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

}  // namespace mimadv
