// scalar_fp64.hip -- plain per-field FP64 scalar advection on gfx950: the
// per-plane PPM kernel and the scheme dispatch (the other schemes run on the
// column-marching kernels of weno_march.hip).
//
// One launch computes the fused x+y+z flux divergence of one field
// (advection_x/y/z of src/TOOLS/advection.f90:556-798 for all three
// directions) and applies one of the epilogues of ScalarArgs: accumulate
// into a tendency, overwrite it, or a full Heun / Shu-Osher stage update
// with the periodic ghost images written in the same pass (so no separate
// statebc/qnbc_p4 sweep is needed between stages).
//
// Work mapping: a workgroup of TX x TY threads owns a TX x TY tile of one
// k-plane.  The plane tile plus a 3-cell halo in x and y is staged in LDS
// with coalesced x-fastest loads.  The x and y cell records (Flux_r,
// Flux_l of limiter_ppm and the CW84 averages) of the tile's cells and one
// neighbour on each side are computed once each into LDS, and every cell
// takes its x and y fluxes from them; the z records of levels k-1..k+1 come
// from the thread's own column (coalesced across the wavefront).  Tiles for neighbouring k of the same
// (x,y) tile have linear block ids 256-apart-multiples, so they land on the
// same XCD and share its L2 for the column reads.
//
// This file is compiled with -ffp-contract=off: the arithmetic follows the
// reference operation order and matches the CPU restatement bit for bit.
#include "mimadv_internal.hpp"

namespace mimadv {

// z contribution of cell k (PPM)
template <int SCHEME, class QF, class WF>
__device__ __forceinline__ double z_contrib(int k, int nz, bool adv, const GridDev &g, const SchemeParams &sp,
                                            QF Q, WF W)
{
    static_assert(SCHEME == 1, "per-plane kernel: PPM only");
    const ZMetrics &zm = g.zm;
    int bc5 = g.bcl[4], bc6 = g.bcl[5];
    double bc51 = (bc5 == 0) ? 1.0 : 0.0, bc61 = (bc6 == 0) ? 1.0 : 0.0;
    // cells k-1, k, k+1 (level nz carries only its Flux_l)
    PPMRec cm{0.0, 0.0}, c0{0.0, 0.0}, cp{0.0, 0.0};
    auto cell = [&](int c) -> PPMRec {
        if (c == nz) {
            PPMRec r;
            r.fr = 0.0;
            r.fl = ppm_ztop_fl(nz, zm, sp, Q, W);
            return r;
        }
        return ppm_zcell(c, nz, zm, sp, bc51, bc61, Q, W);
    };
    bool need = (k >= 2 && k <= nz - 1) || (k == 1 && bc5 == 1) || (k == nz && bc6 == 1);
    if (!need) return 0.0;
    if (k >= 2) cm = cell(k - 1);
    c0 = cell(k);
    if (k <= nz - 1) cp = cell(k + 1);
    auto face = [&](int f) -> double {
        if (f == k - 1) return ppm_faceflux(cm.fr, c0.fl, W(k));
        return ppm_faceflux(c0.fr, cp.fl, W(k + 1));
    };
    return z_assemble(k, nz, adv, bc5, bc6, zm, face, Q, W);
}

template <int SCHEME, int TX, int TY>
__global__ __launch_bounds__(TX *TY) void scalar_fp64_kernel(GridDev g, SchemeParams sp, ScalarArgs a)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    constexpr bool THREE_D = TY > 1;
    constexpr int NT = TX * TY;
    constexpr int LW = TX + 6;
    constexpr int LH = THREE_D ? TY + 6 : 1;
    // cell records (Flux_r, Flux_l) of the x cells i0-1 .. i0+TX of each row
    // and the y cells j0-1 .. j0+TY of each column, each computed once
    constexpr int NXR = TY * (TX + 2), NYR = THREE_D ? (TY + 2) * TX : 0, NR = NXR + NYR;
    __shared__ double tile[LH][LW];
    __shared__ double rfr[NR], rfl[NR];

    const int k = blockIdx.z + 1;
    const int i0 = 4 + blockIdx.x * TX;
    const int j0 = THREE_D ? 4 + blockIdx.y * TY : 1;
    const size_t plane = (size_t)nx * ny;
    const size_t kbase = (size_t)(k - 1) * plane;
    const double *__restrict__ qk = a.q + kbase;
    const int mx = nx - 5, my = ny - 5;  // interior extents (periodic images)

    for (int t = threadIdx.y * TX + threadIdx.x; t < LH * LW; t += TX * TY) {
        int r = t / LW, c = t - r * LW;
        // positions outside the array are replaced by their periodic
        // interior image; the PPM closures never read them
        int ii = i0 - 3 + c;
        if (ii < 1 || ii > nx) {
            int rr = (ii - 4) % mx;
            ii = (rr < 0 ? rr + mx : rr) + 4;
        }
        int jj = 1;
        if (THREE_D) {
            jj = j0 - 3 + r;
            if (jj < 1 || jj > ny) {
                int rr = (jj - 4) % my;
                jj = (rr < 0 ? rr + my : rr) + 4;
            }
        }
        tile[r][c] = qk[(size_t)(ii - 1) + (size_t)nx * (jj - 1)];
    }
    __syncthreads();

    // positions past the array (padding of a partial tile) read a periodic
    // image; their records feed no interior cell
    auto wrap_hi = [](int x, int n, int m) { return x > n ? (x - 4) % m + 4 : x; };
    for (int e = threadIdx.y * TX + threadIdx.x; e < NR; e += NT) {
        PPMRec rec;
        if (e < NXR) {
            if (!(a.dirmask & 1)) continue;
            const int rr = e / (TX + 2), c = i0 - 1 + (e - rr * (TX + 2));
            const int jr = THREE_D ? wrap_hi(j0 + rr, ny, my) : 1;
            const double *__restrict__ urow = a.u + kbase + (size_t)nx * (jr - 1);
            const int lr = THREE_D ? rr + 3 : 0;
            auto Q = [&](int ii) { return tile[lr][ii - (i0 - 3)]; };
            auto U = [&](int ii) { return urow[wrap_hi(ii, nx, mx) - 1]; };
            rec = ppm_hcell(c, nx, 1.0 / g.dx, sp, Q, U);
        } else {
            if (!(THREE_D && (a.dirmask & 2))) continue;
            const int e2 = e - NXR, rr = e2 / TX, cc = e2 - rr * TX, c = j0 - 1 + rr;
            const double *__restrict__ vcol = a.v + kbase + (size_t)(wrap_hi(i0 + cc, nx, mx) - 1);
            auto Q = [&](int jj) { return tile[jj - (j0 - 3)][cc + 3]; };
            auto V = [&](int jj) { return vcol[(size_t)nx * (wrap_hi(jj, ny, my) - 1)]; };
            rec = ppm_hcell(c, ny, 1.0 / g.dy, sp, Q, V);
        }
        rfr[e] = rec.fr;
        rfl[e] = rec.fl;
    }
    __syncthreads();

    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i = i0 + tx, j = j0 + ty;
    if (i > nx - 2 || (THREE_D && j > ny - 2)) return;
    const size_t col = (size_t)(i - 1) + (size_t)nx * (j - 1);
    const size_t id = col + kbase;
    const bool adv = a.adv != 0;
    const int lrow = THREE_D ? ty + 3 : 0;

    double X = 0.0, Y = 0.0, Z = 0.0;
    const double qc = tile[lrow][tx + 3];
    if (a.dirmask & 1) {
        // ppm_h (advq_ppm.f90:136-163) on the records of cells i-1, i, i+1
        const double *__restrict__ urow = a.u + kbase + (size_t)nx * (j - 1);
        const double d1 = 1.0 / g.dx, ui = urow[i - 1], uip = urow[i];
        const int e = ty * (TX + 2) + tx;
        double fr_i = rfr[e + 1], fl_ip = rfl[e + 2];
        if (g.bcl[1] != 0 && i == nx - 2) {
            fr_i = rfl[e + 1];  // Flux_r(ii2) = Flux_l(ii2)
            fl_ip = rfr[e];     // Flux_l(ii2+1) = Flux_r(ii2-1)
        }
        const double flux_old = ppm_faceflux(rfr[e], rfl[e + 1], ui);
        const double flux = ppm_faceflux(fr_i, fl_ip, uip);
        X = (flux - flux_old) * d1;
        if (adv) X = X - qc * (uip - ui) * d1;
    }
    if (THREE_D && (a.dirmask & 2)) {
        const double *__restrict__ vcol = a.v + kbase + (size_t)(i - 1);
        const double d1 = 1.0 / g.dy, vj = vcol[(size_t)nx * (j - 1)], vjp = vcol[(size_t)nx * j];
        const int e = NXR + ty * TX + tx;
        double fr_j = rfr[e + TX], fl_jp = rfl[e + 2 * TX];
        if (g.bcl[3] != 0 && j == ny - 2) {
            fr_j = rfl[e + TX];
            fl_jp = rfr[e];
        }
        const double flux_old = ppm_faceflux(rfr[e], rfl[e + TX], vj);
        const double flux = ppm_faceflux(fr_j, fl_jp, vjp);
        Y = (flux - flux_old) * d1;
        if (adv) Y = Y - qc * (vjp - vj) * d1;
    }
    if (a.dirmask & 4) {
        const double *__restrict__ qc = a.q + col;
        const double *__restrict__ wc = a.w + col;
        auto Q = [&](int kk) { return qc[plane * (size_t)(kk - 1)]; };
        auto W = [&](int kk) { return wc[plane * (size_t)(kk - 1)]; };
        Z = z_contrib<SCHEME>(k, nz, adv, g, sp, Q, W);
    }

    if (a.epi == EPI_ACCUM || a.epi == EPI_OVERWRITE) {
        double r = (a.epi == EPI_ACCUM) ? a.xs[id] : 0.0;
        if (a.dirmask & 1) r = r - X;
        if (THREE_D && (a.dirmask & 2)) r = r - Y;
        if (a.dirmask & 4) r = r - Z;
        a.xs[id] = r;
        return;
    }
    double A = 0.0;
    if (a.dirmask & 1) A = A - X;
    if (THREE_D && (a.dirmask & 2)) A = A - Y;
    if (a.dirmask & 4) A = A - Z;
    double qn;
    if (a.epi == EPI_HEUN) {
        double T = (a.b_lm != 0.0) ? A + a.b_lm * a.tend[id] : A + 0.0;
        a.tend[id] = T;
        qn = a.q0[id] + a.adt * T;
    } else {
        qn = a.c0 * a.q0[id] + a.c1 * (a.q[id] + a.dt * A);
    }
    a.qn[id] = qn;
    if (a.fill) {
        // periodic images (qnbc_p4, bc.f90:409-414 and :431-436):
        // x(1:3) = x(nx-4:nx-2), x(nx-1:nx) = x(4:5), same in y, corners
        // from both
        int ix = 0, jy = 0;
        if (i >= nx - 4) ix = i - mx;
        else if (i <= 5) ix = i + mx;
        if (THREE_D) {
            if (j >= ny - 4) jy = j - my;
            else if (j <= 5) jy = j + my;
        }
        double *qk_out = a.qn + kbase;
        if (ix) qk_out[(size_t)(ix - 1) + (size_t)nx * (j - 1)] = qn;
        if (jy) {
            qk_out[(size_t)(i - 1) + (size_t)nx * (jy - 1)] = qn;
            if (ix) qk_out[(size_t)(ix - 1) + (size_t)nx * (jy - 1)] = qn;
        }
    }
}

template <int SCHEME>
static hipError_t launch_scheme(const GridDev &g, const SchemeParams &sp, const ScalarArgs &a, hipStream_t s)
{
    int mxi = g.nx - 5;
    if (g.ny > 1) {
        constexpr int TX = 64, TY = 4;
        dim3 grid((mxi + TX - 1) / TX, (g.ny - 5 + TY - 1) / TY, g.nz);
        hipLaunchKernelGGL((scalar_fp64_kernel<SCHEME, TX, TY>), grid, dim3(TX, TY), 0, s, g, sp, a);
    } else {
        constexpr int TX = 64, TY = 1;
        dim3 grid((mxi + TX - 1) / TX, 1, g.nz);
        hipLaunchKernelGGL((scalar_fp64_kernel<SCHEME, TX, TY>), grid, dim3(TX, TY), 0, s, g, sp, a);
    }
    return hipGetLastError();
}

hipError_t launch_scalar_fp64(const GridDev &g, const SchemeParams &sp, const ScalarArgs &a, hipStream_t s)
{
    switch (sp.scheme) {
    case 1: return launch_scheme<1>(g, sp, a, s);
    case 0: case 2: case 3: case 4: case 5: case 6: case 7: return launch_march(g, sp, a, 0, s);
    default: return hipErrorInvalidValue;
    }
}

std::vector<double> ppm_coefficients(int nz, const double *dz1w)
{
    std::vector<double> c((size_t)nz * 12, 0.0);
    for (int k = 3; k <= nz - 2; ++k) {
        double delta = dz1w[k - 1], deltam1 = dz1w[k - 2], deltap1 = dz1w[k], deltap2 = dz1w[k + 1];
        double *p = c.data() + 12 * (k - 1);
        p[0] = delta / (deltam1 + delta + deltap1);
        p[1] = (2.0 * deltam1 + delta) / (deltap1 + delta);
        p[2] = (delta + 2.0 * deltap1) / (deltam1 + delta);
        p[3] = deltap1 / (delta + deltap1 + deltap2);
        p[4] = (2.0 * delta + deltap1) / (deltap2 + deltap1);
        p[5] = (deltap1 + 2.0 * deltap2) / (delta + deltap1);
        p[6] = delta / (delta + deltap1);
        p[7] = 1.0 / (deltam1 + delta + deltap1 + deltap2);
        p[8] = 2.0 * deltap1 * delta / (delta + deltap1) *
               ((deltam1 + delta) / (2.0 * delta + deltap1) - (deltap2 + deltap1) / (2.0 * deltap1 + delta));
        p[9] = delta * (deltam1 + delta) / (2.0 * delta + deltap1);
        p[10] = deltap1 * (deltap1 + deltap2) / (2.0 * deltap1 + delta);
    }
    return c;
}

}  // namespace mimadv
