// momentum_batch.hip -- the momenta of the batched FP32 path on gfx950.
//
// The three momenta (rho0 u, rho0 v, rho0 w) in FP32, advected by the FP32
// velocities of the batch with the flux-form centred operator of
// qnadv_uvw (src/INTEG/windadv.f90:82-492, 494-1127): mom_ord 2 or 4,
// the 4th-difference artificial viscosity (lavisc, avisc = 1/16), 'nnf'
// floor and lid, directions in the even-ntime order x, y, z.  Arithmetic
// is FP32 with the level metrics the reference divides by taken as
// precomputed factors (momentum_z_table); the Shu-Osher stage update
// c0*m0 + c1*(m + dt*L) is evaluated in FP64 and rounded once.  The FP64
// parity operator (momentum_fp64.hip) keeps the reference operation order;
// this one is checked against it with tolerances.
//
// Column-marching form: a workgroup owns a TX x TY tile of columns and a
// chunk of levels.  The operands the stencils of level k read with
// horizontal offsets (the momenta at k..k+2, the horizontal velocities at
// k-2..k+1, w at k-1..k+2, columns within two cells of the tile) sit in
// LDS rings of FP32 tiles with a 2-cell halo, each level loaded once,
// coalesced; the values read only in the thread's own column (the momenta
// at k-2, k-1, w at k-2) are carried in registers.  36 KB of LDS per
// 256-thread group: four groups per CU.  One work item per workgroup.
#include "mimadv_internal.hpp"

namespace mimadv {

std::vector<float> momentum_z_table(int nz, const double *dzw, const double *dz1w, const double *dz1)
{
    std::vector<float> t((size_t)MOMZ_N * nz, 0.0f);
    auto Z = [&](int k) { return dzw[k - 1]; };  // 1-based
    for (int k = 1; k <= nz; ++k) {
        float *e = t.data() + (size_t)MOMZ_N * (k - 1);
        if (k < nz) {
            const double s = Z(k + 1) + Z(k);
            e[0] = (float)(Z(k + 1) / s);
            e[1] = (float)(Z(k) / s);
        }
        e[2] = (float)dz1w[k - 1];
        e[3] = (float)dz1[k - 1];
        if (k >= 2 && k < nz) e[4] = (float)(1.0 / (Z(k - 1) + Z(k) + Z(k + 1)));
        if (k >= 2) e[5] = (float)(1.0 / (Z(k) + Z(k - 1)));
        if (k == nz) {
            const double s = Z(nz) + Z(nz - 1);
            e[6] = (float)(Z(nz - 1) * Z(nz - 1) / (s * s));
        }
    }
    return t;
}

namespace {

constexpr float AVISC = 1.0f / 16.0f;

struct MomF {
    float c1s, c2s, c1c, c2c;      // self lines (4/3, 1/3), cross lines (9/8, 1/8)
    float dx1, dx2, dx3, dy1, dy2, dy3;  // 1/d, 1/(2d), 1/(3d)
    int lavisc;
};

template <class F>
__device__ __forceinline__ float visc5(F f, int i)
{
    return -f(i + 2) + 4.0f * f(i + 1) - 6.0f * f(i) + 4.0f * f(i - 1) - f(i - 2);
}

// advux_fd / advvy_fd (windadv.f90:494-554): self advection along a line
template <class MF, class TF>
__device__ __forceinline__ float self_line(int i, float d1, float d2, const MomF &c, MF M, TF T)
{
    auto flux = [&](int ii) { return (T(ii + 1) + T(ii)) * (M(ii + 1) + M(ii)) * 0.25f; };
    auto advf = [&](int ii) { return (T(ii + 2) + T(ii)) * (M(ii + 2) + M(ii)) * 0.25f; };
    float t = c.c1s * (flux(i) - flux(i - 1)) * d1 - c.c2s * (advf(i) - advf(i - 2)) * d2;
    if (c.lavisc) t -= AVISC * (fabsf(T(i)) * d1) * visc5(M, i);
    return t;
}

// advuy_fd / advvx_fd (windadv.f90:556-626): cross advection along a line;
// V(jj, o) is the advecting component at transverse offset o (-2..1)
template <class MF, class VF>
__device__ __forceinline__ float cross_line(int i, float d1, float d3, const MomF &c, MF M, VF V)
{
    auto flux = [&](int ii) {
        const float vv = (c.c1c * (V(ii + 1, -1) + V(ii + 1, 0)) - c.c2c * (V(ii + 1, -2) + V(ii + 1, 1))) * 0.5f;
        return vv * ((M(ii + 1) + M(ii)) * 0.5f);
    };
    float t = c.c1c * (flux(i) - flux(i - 1)) * d1 - c.c2c * (flux(i + 1) - flux(i - 2)) * d3;
    if (c.lavisc) {
        const float mu = 0.25f * fabsf(V(i, -1) + V(i, 0) + V(i + 1, -1) + V(i + 1, 0)) * d1;
        t -= AVISC * mu * visc5(M, i);
    }
    return t;
}

// advuz_fd / advvz_fd with 'nnf' floor and lid (windadv.f90:628-694):
// W(kk, o) the vertical velocity at level kk and transverse offset o;
// m2k the level of dz1w on the k = 2 line (1 for u, 2 for v)
template <class MF, class WF>
__device__ __forceinline__ float cross_z(int k, int nz, const float *mz, int m2k, const MomF &c, MF M, WF W)
{
    auto Zc = [&](int kk, int n) { return mz[MOMZ_N * (kk - 1) + n]; };
    auto flux = [&](int kk) -> float {
        if (kk >= nz || kk < 1) return 0.0f;  // 'nnf' lid: Flux(nz) = 0
        const float ww = (c.c1c * (W(kk + 1, -1) + W(kk + 1, 0)) - c.c2c * (W(kk + 1, -2) + W(kk + 1, 1))) * 0.5f;
        return ww * (Zc(kk, 0) * M(kk + 1) + Zc(kk, 1) * M(kk));
    };
    float t;
    if (k >= 3 && k <= nz - 1) {
        t = c.c1c * (flux(k) - flux(k - 1)) * Zc(k, 2) - c.c2c * (flux(k + 1) - flux(k - 2)) * Zc(k, 4);
    } else if (k == 1) {
        t = flux(1) * Zc(1, 2);
    } else if (k == 2) {
        t = c.c1c * (flux(2) - flux(1)) * Zc(m2k, 2) - c.c2c * flux(3) * Zc(2, 4);
    } else {
        t = -flux(nz - 1) * Zc(nz, 2);
    }
    if (c.lavisc && k >= 3 && k <= nz - 2) {
        const float mu = 0.25f * fabsf(W(k, -1) + W(k, 0) + W(k + 1, -1) + W(k + 1, 0)) * Zc(k, 2);
        t -= AVISC * mu * visc5(M, k);
    }
    return t;
}

// advwx_fd / advwy_fd at level k in 2..nz (windadv.f90:896-1056): M(ii)
// the vertical momentum at level k, T(ii, kk) the advecting component
template <class MF, class TF>
__device__ __forceinline__ float w_line(int i, int k, int nz, float d1, float d3, const MomF &c, MF M, TF T)
{
    auto uu = [&](int ii) -> float {
        if (k == 2) return (T(ii, 2) + T(ii, 1)) * 0.5f;
        if (k == nz) return (T(ii, nz) + T(ii, nz - 1)) * 0.5f;
        return (c.c1c * (T(ii, k) + T(ii, k - 1)) - c.c2c * (T(ii, k + 1) + T(ii, k - 2))) * 0.5f;
    };
    auto flux = [&](int ii) { return uu(ii + 1) * ((M(ii + 1) + M(ii)) * 0.5f); };
    float t = c.c1c * (flux(i) - flux(i - 1)) * d1 - c.c2c * (flux(i + 1) - flux(i - 2)) * d3;
    if (c.lavisc) {
        const float mu = 0.25f * fabsf(T(i, k) + T(i, k - 1) + T(i + 1, k) + T(i + 1, k - 1)) * d1;
        t -= AVISC * mu * visc5(M, i);
    }
    return t;
}

// advwz_fd with 'nnf' floor and lid (windadv.f90:1058-1127)
template <class MF, class TF>
__device__ __forceinline__ float w_z(int k, int nz, const float *mz, const MomF &c, MF M, TF T)
{
    auto Zc = [&](int kk, int n) { return mz[MOMZ_N * (kk - 1) + n]; };
    auto flux = [&](int kk) -> float {
        if (kk >= nz || kk < 1) return 0.0f;
        return (T(kk + 1) + T(kk)) * (M(kk + 1) + M(kk)) * 0.25f;
    };
    auto advf = [&](int kk) -> float {
        if (kk < 1 || kk >= nz) return 0.0f;
        if (kk == nz - 1) return Zc(nz, 6) * T(nz - 1) * M(nz - 1);
        return (Zc(kk, 0) * T(kk + 2) + Zc(kk, 1) * T(kk)) * (Zc(kk, 0) * M(kk + 2) + Zc(kk, 1) * M(kk));
    };
    float t = 0.0f;
    if (k >= 3 && k <= nz - 1) {
        t = c.c1s * (flux(k) - flux(k - 1)) * Zc(k - 1, 3) - c.c2s * (advf(k) - advf(k - 2)) * Zc(k, 5);
    } else if (k == 2) {
        t = c.c1s * (flux(2) - flux(1)) * Zc(1, 3) - c.c2s * advf(2) * Zc(2, 5);
    } else if (k == nz) {
        t = -flux(nz - 1) * Zc(nz - 1, 3);
    }
    if (c.lavisc && k >= 3 && k <= nz - 2) t -= AVISC * (fabsf(T(k)) * Zc(k - 1, 3)) * visc5(M, k);
    return t;
}

template <int TX, int TY>
__global__ __launch_bounds__(TX *TY) __attribute__((amdgpu_waves_per_eu(4, 8))) void momentum_batch_kernel(
    GridDev g, MomF c, MomBatchArgs a, int kchunk, int ntx, int nwork)
{
    constexpr bool THREE_D = TY > 1;
    constexpr int NT = TX * TY;
    constexpr int H = 2, LW = TX + 2 * H, LH = THREE_D ? TY + 2 * H : 1, PL = LW * LH;
    constexpr int TS = (PL + NT - 1) / NT;
    // LDS rings of FP32 tiles with a 2-cell halo, each holding the levels
    // its stencils read with horizontal offsets: the momenta at k..k+2, the
    // horizontal velocities at k-2..k+1, w at k-1..k+2.  The stencils read
    // the momenta at k-2, k-1 and w at k-2 only in the thread's own column:
    // those stay in registers.
    __shared__ float ringm[3][3][PL];  // mu, mv, mw
    __shared__ float ringh[2][4][PL];  // ut, vt
    __shared__ float ringw[4][PL];     // wt

    const int nx = g.nx, ny = g.ny, nz = g.nz, mx = nx - 5, my = ny - 5;
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
    const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
    const int nchunk = (nz + kchunk - 1) / kchunk;
    const float *src[6] = {a.m[0], a.m[1], a.m[2], a.ut, a.vt, a.wt};
    const float *mz = g.zm.momz;

    if (const int wi = blockIdx.x; wi < nwork) {
        // Work order: level chunk outermost, tiles x-fastest within it.  The
        // hardware deals workgroups round-robin over the 8 XCDs; with
        // nwork % 8 == 0 every XCD takes a contiguous range of that list,
        // so a tile's x and y neighbours of the same chunk run on the same
        // XCD at about the same time and its halo rows come from that L2.
        const int ntile = nwork / nchunk;
        const int w = (nwork % 8 == 0) ? (wi % 8) * (nwork / 8) + wi / 8 : wi;
        const int kc = w / ntile, tile = w - kc * ntile;
        const int i0 = 4 + (tile % ntx) * TX;
        const int j0 = THREE_D ? 4 + (tile / ntx) * TY : 1;
        const int k0 = 1 + kc * kchunk, k1 = min(nz, k0 + kchunk - 1);
        const int i = i0 + tx, j = j0 + ty;
        const bool valid = i <= nx - 2 && (!THREE_D || j <= ny - 2);
        // ring position of this thread's own column
        const int own = (THREE_D ? (ty + H) * LW : 0) + tx + H;
        // plane offset of ring entry s of this thread (ghost columns read as
        // the caller filled them; positions past the array, the padding of a
        // partial tile, take a periodic image that feeds no interior cell)
        auto t_off = [&](int s) -> int {
            const int e = tid + s * NT;
            if (e >= PL) return -1;
            const int r = e / LW, cc = e - r * LW;
            int ii = i0 - H + cc;
            if (ii > nx) ii = (ii - 4) % mx + 4;
            int jj = 1;
            if (THREE_D) {
                jj = j0 - H + r;
                if (jj > ny) jj = (jj - 4) % my + 4;
            }
            return (ii - 1) + nx * (jj - 1);
        };
        // level km of the momenta and w, level kh of the horizontal
        // velocities (levels outside 1..nz read as zero)
        float pre[6][TS];
        auto fetch = [&](int km, int kh) {
            const bool inm = km >= 1 && km <= nz, inh = kh >= 1 && kh <= nz;
            const size_t kbm = inm ? sz * (km - 1) : 0, kbh = inh ? sz * (kh - 1) : 0;
#pragma unroll
            for (int s = 0; s < TS; ++s) {
                const int off = t_off(s);
#pragma unroll
                for (int n = 0; n < 6; ++n) {
                    const bool h = n == 3 || n == 4;
                    pre[n][s] = (off >= 0 && (h ? inh : inm)) ? src[n][(h ? kbh : kbm) + off] : 0.0f;
                }
            }
        };
        auto store = [&](int km, int kh, bool with_h) {
            const int sm = (km + 6) % 3;
#pragma unroll
            for (int s = 0; s < TS; ++s) {
                const int e = tid + s * NT;
                if (e >= PL) continue;
#pragma unroll
                for (int n = 0; n < 3; ++n) ringm[n][sm][e] = pre[n][s];
                ringw[km & 3][e] = pre[5][s];
                if (with_h) {
                    ringh[0][kh & 3][e] = pre[3][s];
                    ringh[1][kh & 3][e] = pre[4][s];
                }
            }
        };
        // fill: levels k0-2..k0 of every ring, then the own-column registers
        // of k0-2, k0-1 (their ring slots are reused), then the momenta and w
        // at k0+1; the momenta and w at k0+2 and the horizontal velocities at
        // k0+1 are requested for the first level
        float mk2_0, mk2_1, mk2_2, mk1_0, mk1_1, mk1_2, wkm2;
#pragma unroll 1
        for (int kk = k0 - 2; kk <= k0; ++kk) {
            fetch(kk, kk);
            store(kk, kk, true);
        }
        __syncthreads();
        {
            const int a2 = (k0 - 2 + 6) % 3, a1 = (k0 - 1 + 6) % 3;
            mk2_0 = ringm[0][a2][own];
            mk2_1 = ringm[1][a2][own];
            mk2_2 = ringm[2][a2][own];
            mk1_0 = ringm[0][a1][own];
            mk1_1 = ringm[1][a1][own];
            mk1_2 = ringm[2][a1][own];
        }
        wkm2 = ringw[(k0 - 2) & 3][own];
        __syncthreads();
        fetch(k0 + 1, k0 + 1);
        store(k0 + 1, k0 + 1, false);
        fetch(k0 + 2, k0 + 1);
        // step-start momenta of the thread's cell, requested one level ahead
        const bool need_m0 = valid && a.epi != 0 && a.c0 != 0.0;
        float m0n[3] = {0.0f, 0.0f, 0.0f};
        auto fetch_m0 = [&](int kk) {
            const size_t id = (size_t)(i - 1) + sy * (j - 1) + sz * (kk - 1);
#pragma unroll
            for (int n = 0; n < 3; ++n)
                if (n != 1 || THREE_D) m0n[n] = a.m0[n][id];
        };
        if (need_m0) fetch_m0(k0);
#pragma unroll 1
        for (int k = k0; k <= k1; ++k) {
            store(k + 2, k + 1, true);
            const float m0v[3] = {m0n[0], m0n[1], m0n[2]};
            __syncthreads();
            if (k < k1) {
                fetch(k + 3, k + 2);
                if (need_m0) fetch_m0(k + 1);
            }
            // ring slot of the momenta at level k
            const int s0 = k % 3;
            if (valid) {
                const size_t kb = sz * (k - 1);
                const size_t id = (size_t)(i - 1) + sy * (j - 1) + kb;
                // operand n (a literal) at column offsets (di, dj) from this
                // thread's column and level kk (k-2 <= kk <= k+2); the levels
                // kept in registers are read only in the own column
                auto R = [&](int n, int di, int dj, int kk) __attribute__((always_inline)) -> float {
                    const int e = own + (THREE_D ? dj * LW : 0) + di;
                    const int dk = kk - k;
                    if (n < 3) {
                        const float r2 = n == 0 ? mk2_0 : (n == 1 ? mk2_1 : mk2_2);
                        const float r1 = n == 0 ? mk1_0 : (n == 1 ? mk1_1 : mk1_2);
                        if (dk < 0) return dk == -2 ? r2 : r1;
                        const int sl = s0 + dk;  // dk in 0..2
                        return ringm[n][sl >= 3 ? sl - 3 : sl][e];
                    }
                    if (n < 5) return ringh[n - 3][kk & 3][e];
                    if (dk == -2) return wkm2;
                    return ringw[kk & 3][e];
                };
                float L[3];
                // ---- u (uadv_fd): x self, y cross, z cross
                {
                    float t = self_line(0, c.dx1, c.dx2, c, [&](int di) { return R(0, di, 0, k); },
                                        [&](int di) { return R(3, di, 0, k); });
                    if (THREE_D)
                        t += cross_line(0, c.dy1, c.dy3, c, [&](int dj) { return R(0, 0, dj, k); },
                                        [&](int dj, int o) { return R(4, o, dj, k); });
                    t += cross_z(k, nz, mz, 1, c, [&](int kk) { return R(0, 0, 0, kk); },
                                 [&](int kk, int o) { return R(5, o, 0, kk); });
                    L[0] = t;
                }
                // ---- v (vadv_fd), 3D only: x cross, y self, z cross
                L[1] = 0.0f;
                if (THREE_D) {
                    float t = cross_line(0, c.dx1, c.dx3, c, [&](int di) { return R(1, di, 0, k); },
                                         [&](int di, int o) { return R(3, di, o, k); });
                    t += self_line(0, c.dy1, c.dy2, c, [&](int dj) { return R(1, 0, dj, k); },
                                   [&](int dj) { return R(4, 0, dj, k); });
                    t += cross_z(k, nz, mz, 2, c, [&](int kk) { return R(1, 0, 0, kk); },
                                 [&](int kk, int o) { return R(5, 0, o, kk); });
                    L[1] = t;
                }
                // ---- w (wadv_fd), levels 2..nz: x, y, z
                L[2] = 0.0f;
                if (k >= 2) {
                    float t = w_line(0, k, nz, c.dx1, c.dx3, c, [&](int di) { return R(2, di, 0, k); },
                                     [&](int di, int kk) { return R(3, di, 0, kk); });
                    if (THREE_D)
                        t += w_line(0, k, nz, c.dy1, c.dy3, c, [&](int dj) { return R(2, 0, dj, k); },
                                    [&](int dj, int kk) { return R(4, 0, dj, kk); });
                    t += w_z(k, nz, mz, c, [&](int kk) { return R(2, 0, 0, kk); }, [&](int kk) { return R(5, 0, 0, kk); });
                    L[2] = t;
                }
#pragma unroll
                for (int n = 0; n < 3; ++n) {
                    if (n == 1 && !THREE_D) continue;
                    const double Ln = -(double)L[n];
                    float qv;
                    if (a.epi == 0) {
                        qv = (float)Ln;
                    } else {
                        const double m = (double)ringm[n][s0][own];
                        const double qn = a.c0 != 0.0 ? a.c0 * (double)m0v[n] + a.c1 * (m + a.dt * Ln)
                                                      : a.c1 * (m + a.dt * Ln);
                        qv = (float)qn;
                    }
                    float *o = a.out[n];
                    o[id] = qv;
                    if (a.epi != 0 && a.fill) {
                        // periodic images (lbcu_p4 for the periodic case)
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
            // own-column registers for level k+1; the next level's stores
            // overwrite the slots of the momenta at k and of w at k-1
            mk2_0 = mk1_0;
            mk2_1 = mk1_1;
            mk2_2 = mk1_2;
            mk1_0 = ringm[0][s0][own];
            mk1_1 = ringm[1][s0][own];
            mk1_2 = ringm[2][s0][own];
            wkm2 = ringw[(k - 1) & 3][own];
            __syncthreads();
        }
    }
}

}  // namespace

hipError_t launch_momentum_batch(const GridDev &g, int mom_ord, int lavisc, const MomBatchArgs &a, hipStream_t s)
{
    MomF c;
    if (mom_ord == 4) {
        c.c1s = 4.0f / 3.0f;
        c.c2s = 1.0f / 3.0f;
        c.c1c = 9.0f / 8.0f;
        c.c2c = 1.0f / 8.0f;
    } else {
        c.c1s = 1.0f;
        c.c2s = 0.0f;
        c.c1c = 1.0f;
        c.c2c = 0.0f;
    }
    c.dx1 = (float)(1.0 / g.dx);
    c.dx2 = (float)(1.0 / (2.0 * g.dx));
    c.dx3 = (float)(1.0 / (3.0 * g.dx));
    c.dy1 = (float)(1.0 / g.dy);
    c.dy2 = (float)(1.0 / (2.0 * g.dy));
    c.dy3 = (float)(1.0 / (3.0 * g.dy));
    c.lavisc = lavisc;
    // Levels per work item: the ring fill of an item (its first four levels)
    // is paid once per chunk, so an item marches its whole column when the
    // tiles alone fill the resident groups (four per CU); smaller grids split
    // the columns into chunks of at least 32 levels.  At 512x512x160 whole
    // columns are 4 % faster than 32-level chunks, and those 4 % faster than
    // 16-level ones.
    const bool three_d = g.ny > 1;
    const int ntx3 = (g.nx - 5 + 31) / 32, nty3 = (g.ny - 5 + 7) / 8;
    const int tiles = three_d ? ntx3 * nty3 : (g.nx - 5 + 63) / 64;
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
        ncu = 256;
    const int groups = 4 * ncu;
    int kchunk = g.nz;
    if (tiles < groups) {
        const int want = (groups + tiles - 1) / tiles;  // chunks per column
        kchunk = std::min(g.nz, std::max(32, (g.nz + want - 1) / want));
    }
    const int nchunk = (g.nz + kchunk - 1) / kchunk;
    if (three_d) {
        constexpr int TX = 32, TY = 8;
        const int ntx = ntx3;
        const int nwork = ntx * nty3 * nchunk;
        hipLaunchKernelGGL((momentum_batch_kernel<TX, TY>), dim3(nwork), dim3(TX * TY), 0, s, g, c, a, kchunk, ntx,
                           nwork);
    } else {
        constexpr int TX = 64, TY = 1;
        const int ntx = (g.nx - 5 + TX - 1) / TX;
        const int nwork = ntx * nchunk;
        hipLaunchKernelGGL((momentum_batch_kernel<TX, TY>), dim3(nwork), dim3(TX), 0, s, g, c, a, kchunk, ntx,
                           nwork);
    }
    return hipGetLastError();
}

}  // namespace mimadv
