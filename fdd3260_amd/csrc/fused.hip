// fused.hip -- batched, mixed-precision scalar advection on gfx950.
//
// One persistent launch advects a batch of fields that share one velocity
// (u, v, w in FP32): up to MAX64 fields in FP64 (theta, water vapour: small
// increments on large values) and the rest in FP32.  Each workgroup owns a
// TX x TY tile of columns and a range of levels and marches up the column:
//
//   * per level the three velocity faces of every column are read once and
//     serve every field;
//   * per field the level's tile plus a 2-cell halo goes through LDS; every
//     thread evaluates the east x-face and north y-face of its cell, the
//     first TY + TX threads the west/south edge faces of the tile, and the
//     shared faces are exchanged through LDS (each face evaluated once);
//   * the z stencil is a register window q(k-1..k+2) per field, refilled
//     with one coalesced load per level; the top-face flux of level k is
//     carried to level k+1 as its bottom face;
//   * face fluxes of FP32 fields are FP32; the flux differences and the
//     divergence are accumulated in FP64 (exact differences of FP32 values,
//     a compensated sum) and the Shu-Osher stage update
//     q_new = c0*q0 + c1*(q + dt*div) is evaluated in FP64 and rounded once.
//
// Schemes: the reference WENO ('weno' of advq_weno.f90, including its
// limiter, the k = 3..nz-3 range of nonlinear vertical weights and the
// bottom/top closures) and MUSCL (advq_muscl.f90) in an operation-reduced
// form: the positivity rescale is applied to the weighted face value
// (b + phi*(F - b), equal to rescaling each candidate since the weights
// sum to one), only the upwind side's weights are evaluated, and divisions
// are reciprocals.  It is checked against the FP64 oracle with tolerances,
// not bit for bit.  Vertical boundaries must be no-flux ('nnf'), the
// lateral ones periodic.
#include "fused.hpp"

namespace mimadv {

namespace {

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
    return 1.0 / x;
}

template <class T>
__device__ __forceinline__ T tmin(T a, T b) { return a < b ? a : b; }
template <class T>
__device__ __forceinline__ T tmax(T a, T b) { return a > b ? a : b; }

// Reference WENO face flux, operation-reduced.  a..d = q(f-1..f+2), uu the
// face velocity, nonlin: whether the smoothness switch is active at this face
// (always in x/y, k = 3..nz-3 in z).
template <class T>
__device__ __forceinline__ T weno_face(T a, T b, T c, T d, T uu, bool nonlin)
{
    const T eps = T(1e-15), thr = T(20);
    T ffim1 = T(1.5) * b - T(0.5) * a;
    T ffip1 = T(0.5) * (b + c);
    T ffip2 = T(1.5) * c - T(0.5) * d;
    // limiter_weno: rescale about b (the left/lower cell)
    T minp = tmin(tmin(ffim1, ffip1), ffip2);
    T phi = tmin(fabs(b * rcp((minp - b) + eps)), T(1));
    bool pos = uu > T(0);
    T e2 = (c - b) * (c - b);
    T oth = pos ? (b - a) : (d - c);
    T eo = oth * oth;
    T fo = pos ? ffim1 : ffip2;
    T wo = T(1.0 / 3.0);
    if (nonlin) {
        T mx = tmax(eo, e2), mn = tmin(eo, e2);
        if (mx > thr * (mn + eps)) {
            T so = eps + eo;
            so = so * so;
            T si = eps + e2;
            si = si * si;
            // (g1/so) / (g1/so + g2/si) with g2 = 2 g1
            wo = si * rcp(si + T(2) * so);
        }
    }
    T F = ffip1 + wo * (fo - ffip1);
    F = b + phi * (F - b);
    return uu * F;
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

template <class T>
__device__ __forceinline__ T muscl_face(T a, T b, T c, T d, T uu, int lim, T lt, T d1)
{
    const T eps = T(1e-18);
    T sp = c - b, sp2 = d - c, sm = b - a;  // slopes times dx
    T fl, fr;
    T tol = lt * fabs(b) + eps;
    if (fabs(sm) > tol || fabs(sp2) > tol) {
        // limiter_muscl compares per-metre slopes with lim_tol*|x|
        // (advq_muscl.f90:361-362)
        T rp1 = T(1), rp2 = T(1);
        if (fabs(sm) * d1 > lt * fabs(b) && fabs(sp) < fabs(sm) * rcp(lt)) rp1 = sp * rcp(sm);
        if (fabs(sp2) * d1 > lt * fabs(b) && fabs(sp) < fabs(sp2) * rcp(lt)) rp2 = sp * rcp(sp2);
        fr = c - T(0.5) * muscl_phi(lim, rp2) * sp2;
        fl = b + T(0.5) * muscl_phi(lim, rp1) * sm;
    } else {
        fr = c - T(0.5) * sp;
        fl = b + T(0.5) * sp;
    }
    return uu * (uu > T(0) ? fl : fr);
}

// MUSCL z face f (1..nz-1) with the reference's stretched metrics
template <class T>
__device__ __forceinline__ T muscl_zface(int f, int nz, T a, T b, T c, T d, T ww, const FusedArgs &A)
{
    const T eps = T(1e-18);
    int lim = A.muscl_limiter;
    T lt = T(A.lim_tol), dz = T(A.dz);
    const double *dz1 = A.dz1, *fdz0 = A.fdz0;
    T m = T(A.dzw[f - 1]);  // 1/dz1w(f)
    T sp = (c - b) * T(dz1[f - 1]);
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
    T fl, fr;
    if (limited) {
        T rp1 = T(1), rp2 = T(1);
        if (fabs(sm) > lt * fabs(b) && fabs(sp) < fabs(sm) * rcp(lt)) rp1 = sp * rcp(sm);
        if (fabs(sp2) > lt * fabs(b) && fabs(sp) < fabs(sp2) * rcp(lt)) rp2 = sp * rcp(sp2);
        fr = c - T(0.5) * muscl_phi(lim, rp2) * sp2 * m;
        fl = b + T(0.5) * muscl_phi(lim, rp1) * sm * m;
    } else {
        fr = c - T(0.5) * sp * m;
        fl = b + T(0.5) * sp * m;
    }
    return ww * (ww > T(0) ? fl : fr);
}

template <int SCHEME, class T>
__device__ __forceinline__ T hface(T a, T b, T c, T d, T uu, const FusedArgs &A, double d1)
{
    if constexpr (SCHEME == 0) return weno_face(a, b, c, d, uu, true);
    else return muscl_face(a, b, c, d, uu, A.muscl_limiter, T(A.lim_tol), T(d1));
}

template <int SCHEME, class T>
__device__ __forceinline__ T zface(int f, int nz, const T *zw, T ww, const FusedArgs &A)
{
    // zw = q(f-1), q(f), q(f+1), q(f+2) (q(0) := q(1))
    if (f >= nz) return T(0);
    if constexpr (SCHEME == 0) {
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

__device__ __forceinline__ int wrap(int i, int n, int m)
{
    // 1-based index i mapped into [1, n]; outside positions take their
    // periodic interior image (period m, interior starts at 4)
    if (i >= 1 && i <= n) return i;
    int r = (i - 4) % m;
    return (r < 0 ? r + m : r) + 4;
}

template <class T>
struct FieldPtrs;
template <>
struct FieldPtrs<double> {
    static __device__ __forceinline__ const double *q(const FusedArgs &A, int f) { return A.q64[f]; }
    static __device__ __forceinline__ const double *q0(const FusedArgs &A, int f) { return A.q0_64[f]; }
    static __device__ __forceinline__ double *out(const FusedArgs &A, int f) { return A.out64[f]; }
};
template <>
struct FieldPtrs<float> {
    static __device__ __forceinline__ const float *q(const FusedArgs &A, int f) { return A.q32[f]; }
    static __device__ __forceinline__ const float *q0(const FusedArgs &A, int f) { return A.q0_32[f]; }
    static __device__ __forceinline__ float *out(const FusedArgs &A, int f) { return A.out32[f]; }
};

template <int SCHEME, int TX, int TY, class T>
__device__ __forceinline__ void do_field(const FusedArgs &A, int f, int k, int tid, int tx, int ty, int i0, int j0,
                                         bool valid, size_t id, T (&zw)[4], T &fzb, T ue, T vn, T uw_e, T vs_e,
                                         T wt, double dz1w_k, T (*tile)[TX + 4], T (*fx)[TX + 1], T (*fy)[TX],
                                         const int *hal_r, const int *hal_c, const T *hal_v, const bool *hal_ok)
{
    constexpr int NT = TX * TY;
    constexpr int HPT = (4 * TX + 4 * TY + 16 + NT - 1) / NT;
    using P = FieldPtrs<T>;
    const int nz = A.nz;
    // own value and halo into LDS (the previous field's readers finished at
    // its second barrier)
    tile[ty + 2][tx + 2] = zw[1];
#pragma unroll
    for (int h = 0; h < HPT; ++h)
        if (hal_ok[h]) tile[hal_r[h]][hal_c[h]] = hal_v[h];
    __syncthreads();
    T Fx = hface<SCHEME, T>(tile[ty + 2][tx + 1], tile[ty + 2][tx + 2], tile[ty + 2][tx + 3], tile[ty + 2][tx + 4], ue,
                            A, A.dx1);
    fx[ty][tx + 1] = Fx;
    T Fy = T(0);
    if (A.three_d) {
        Fy = hface<SCHEME, T>(tile[ty + 1][tx + 2], tile[ty + 2][tx + 2], tile[ty + 3][tx + 2], tile[ty + 4][tx + 2], vn,
                              A, A.dy1);
        fy[ty + 1][tx] = Fy;
    }
    if (tid < TY) {  // west edge face of row tid
        int r = tid;
        fx[r][0] = hface<SCHEME, T>(tile[r + 2][0], tile[r + 2][1], tile[r + 2][2], tile[r + 2][3], uw_e, A, A.dx1);
    } else if (A.three_d && tid < TY + TX) {  // south edge face of column tid-TY
        int c = tid - TY;
        fy[0][c] = hface<SCHEME, T>(tile[0][c + 2], tile[1][c + 2], tile[2][c + 2], tile[3][c + 2], vs_e, A, A.dy1);
    }
    T Fz = zface<SCHEME, T>(k, nz, zw, wt, A);
    __syncthreads();
    if (!valid) {
        fzb = Fz;
        return;
    }
    double div = 0.0;
    if (A.dirmask & 1) div += (double(Fx) - double(fx[ty][tx])) * A.dx1;
    if (A.three_d && (A.dirmask & 2)) div += (double(Fy) - double(fy[ty][tx])) * A.dy1;
    if (A.dirmask & 4) div += (double(Fz) - double(fzb)) * dz1w_k;
    fzb = Fz;
    double a = -div;
    T *out = P::out(A, f);
    if (A.epi == 0) {
        out[id] = T(a);
        return;
    }
    double qn;
    if (A.c0 != 0.0) qn = A.c0 * double(P::q0(A, f)[id]) + A.c1 * (double(zw[1]) + A.dt * a);
    else qn = A.c1 * (double(zw[1]) + A.dt * a);
    T qv = T(qn);
    out[id] = qv;
    if (A.fill) {
        const int nx = A.nx, ny = A.ny, mx = A.mx, my = A.my;
        int i = i0 + tx, j = j0 + ty;
        int ix = 0, jy = 0;
        if (i >= nx - 4) ix = i - mx;
        else if (i <= 5) ix = i + mx;
        if (A.three_d) {
            if (j >= ny - 4) jy = j - my;
            else if (j <= 5) jy = j + my;
        }
        size_t kb = id - ((size_t)(i - 1) + (size_t)nx * (j - 1));
        if (ix) out[kb + (size_t)(ix - 1) + (size_t)nx * (j - 1)] = qv;
        if (jy) {
            out[kb + (size_t)(i - 1) + (size_t)nx * (jy - 1)] = qv;
            if (ix) out[kb + (size_t)(ix - 1) + (size_t)nx * (jy - 1)] = qv;
        }
    }
}

template <int SCHEME, int N64, int N32, int TX, int TY>
__global__ __launch_bounds__(TX *TY) void fused_kernel(FusedArgs A)
{
    constexpr int NT = TX * TY;
    constexpr int LW = TX + 4, LH = TY + 4;
    constexpr int NHALO = LW * LH - TX * TY;
    constexpr int HPT = (NHALO + NT - 1) / NT;  // halo cells per thread
    __shared__ double tile64[N64 > 0 ? LH : 1][LW];
    __shared__ double fx64[N64 > 0 ? TY : 1][TX + 1];
    __shared__ double fy64[N64 > 0 ? TY + 1 : 1][TX];
    __shared__ float tile32[N32 > 0 ? LH : 1][LW];
    __shared__ float fx32[N32 > 0 ? TY : 1][TX + 1];
    __shared__ float fy32[N32 > 0 ? TY + 1 : 1][TX];

    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const size_t plane = (size_t)nx * ny;
    const int tid = threadIdx.x;
    const int tx = tid % TX, ty = tid / TX;

    for (int wi = blockIdx.x; wi < A.nwork; wi += gridDim.x) {
        // XCD-aware order: consecutive work ids go round-robin over the 8
        // XCDs; give every XCD a contiguous range of tiles so that tile halos
        // and the k-chunks of one column stay in one L2.
        int tiles = A.ntx * A.nty;
        int chunk = wi / tiles;
        int t = wi - chunk * tiles;
        if (A.xcd_remap) {
            int per = tiles / 8;
            t = (t % 8) * per + t / 8;
        }
        int bx = t % A.ntx, by = t / A.ntx;
        int i0 = 4 + bx * TX;
        int j0 = A.three_d ? 4 + by * TY : 1;
        int ka = 1 + chunk * A.kchunk;
        int kb = min(nz, ka + A.kchunk - 1);
        int i = i0 + tx, j = j0 + ty;
        bool valid = i <= nx - 2 && (!A.three_d || j <= ny - 2) && (A.three_d || ty == 0);
        int il = wrap(i, nx, A.mx);
        int jl = A.three_d ? wrap(j, ny, A.my) : 1;
        size_t col = (size_t)(il - 1) + (size_t)nx * (jl - 1);

        // halo cells owned by this thread (fixed for the work item); slot h
        // is static so the bookkeeping stays in registers
        int hal_r[HPT], hal_c[HPT];
        size_t hal_col[HPT];
        bool hal_ok[HPT];
#pragma unroll
        for (int h = 0; h < HPT; ++h) {
            int e = tid + h * NT;
            // ring enumeration: rows 0,1 and LH-2,LH-1 full; middle rows
            // columns 0,1,LW-2,LW-1
            int r = 0, c = 0;
            const int top = 2 * LW;
            if (e < top) { r = e / LW; c = e % LW; }
            else if (e < 2 * top) { int e2 = e - top; r = LH - 2 + e2 / LW; c = e2 % LW; }
            else {
                int e3 = e - 2 * top;
                r = 2 + e3 / 4;
                int cc = e3 % 4;
                c = cc < 2 ? cc : LW - 4 + cc;
            }
            hal_ok[h] = e < NHALO && (A.three_d || r == 2);
            int ii = wrap(i0 - 2 + c, nx, A.mx);
            int jj = A.three_d ? wrap(j0 - 2 + r, ny, A.my) : 1;
            hal_r[h] = r;
            hal_c[h] = c;
            hal_col[h] = hal_ok[h] ? (size_t)(ii - 1) + (size_t)nx * (jj - 1) : 0;
        }

        // register windows q(k-1..k+2) with q(0) := q(1); carried bottom face
        double z64[N64 > 0 ? N64 : 1][4];
        double fzb64[N64 > 0 ? N64 : 1];
        float z32[N32 > 0 ? N32 : 1][4];
        float fzb32[N32 > 0 ? N32 : 1];
        auto lev = [&](int kk) { return (size_t)(min(max(kk, 1), nz) - 1) * plane; };
#pragma unroll
        for (int f = 0; f < N64; ++f) {
            fzb64[f] = 0.0;
            if (f >= A.n64_active) continue;
            const double *q = A.q64[f];
#pragma unroll
            for (int s = 0; s < 3; ++s) z64[f][s + 1] = q[col + lev(ka - 1 + s)];
            z64[f][0] = z64[f][1];
        }
#pragma unroll
        for (int f = 0; f < N32; ++f) {
            fzb32[f] = 0.0f;
            if (f >= A.n32_active) continue;
            const float *q = A.q32[f];
#pragma unroll
            for (int s = 0; s < 3; ++s) z32[f][s + 1] = q[col + lev(ka - 1 + s)];
            z32[f][0] = z32[f][1];
        }
        // carried bottom face of the first level (zero at the floor)
        if (ka > 1) {
            float wb = A.w[col + lev(ka)];
#pragma unroll
            for (int f = 0; f < N64; ++f) {
                if (f >= A.n64_active) continue;
                const double *q = A.q64[f];
                double zz[4] = {q[col + lev(ka - 2)], z64[f][1], z64[f][2], z64[f][3]};
                fzb64[f] = zface<SCHEME, double>(ka - 1, nz, zz, double(wb), A);
            }
#pragma unroll
            for (int f = 0; f < N32; ++f) {
                if (f >= A.n32_active) continue;
                const float *q = A.q32[f];
                float zz[4] = {q[col + lev(ka - 2)], z32[f][1], z32[f][2], z32[f][3]};
                fzb32[f] = zface<SCHEME, float>(ka - 1, nz, zz, wb, A);
            }
        }

        for (int k = ka; k <= kb; ++k) {
            size_t kbase = (size_t)(k - 1) * plane;
            size_t id = col + kbase;
            // velocities of the level, shared by every field
            int ie = wrap(il + 1, nx, A.mx);
            float ue = A.u[kbase + (size_t)(ie - 1) + (size_t)nx * (jl - 1)];
            float vn = 0.0f;
            if (A.three_d) {
                int je = wrap(jl + 1, ny, A.my);
                vn = A.v[kbase + (size_t)(il - 1) + (size_t)nx * (je - 1)];
            }
            float wt = (k < nz) ? A.w[col + kbase + plane] : 0.0f;
            float uw_e = 0.0f, vs_e = 0.0f;
            if (tid < TY) {
                int jr = A.three_d ? wrap(j0 + tid, ny, A.my) : 1;
                uw_e = A.u[kbase + (size_t)(i0 - 1) + (size_t)nx * (jr - 1)];
            } else if (A.three_d && tid < TY + TX) {
                int ic = wrap(i0 + tid - TY, nx, A.mx);
                vs_e = A.v[kbase + (size_t)(ic - 1) + (size_t)nx * (j0 - 1)];
            }
            double dz1w_k = A.dz1w[k - 1];
            // issue every field's window refill and halo loads first
            size_t kn = lev(k + 2);
            double h64[N64 > 0 ? N64 : 1][HPT];
            float h32[N32 > 0 ? N32 : 1][HPT];
#pragma unroll
            for (int f = 0; f < N64; ++f) {
                if (f >= A.n64_active) continue;
                z64[f][0] = z64[f][1]; z64[f][1] = z64[f][2]; z64[f][2] = z64[f][3];
                z64[f][3] = A.q64[f][col + kn];
#pragma unroll
                for (int h = 0; h < HPT; ++h)
                    if (hal_ok[h]) h64[f][h] = A.q64[f][hal_col[h] + kbase];
            }
#pragma unroll
            for (int f = 0; f < N32; ++f) {
                if (f >= A.n32_active) continue;
                z32[f][0] = z32[f][1]; z32[f][1] = z32[f][2]; z32[f][2] = z32[f][3];
                z32[f][3] = A.q32[f][col + kn];
#pragma unroll
                for (int h = 0; h < HPT; ++h)
                    if (hal_ok[h]) h32[f][h] = A.q32[f][hal_col[h] + kbase];
            }
            if (k == 1) {
                // q(0) := q(1) for the first window
#pragma unroll
                for (int f = 0; f < N64; ++f) z64[f][0] = z64[f][1];
#pragma unroll
                for (int f = 0; f < N32; ++f) z32[f][0] = z32[f][1];
            }
#pragma unroll
            for (int f = 0; f < N64; ++f)
                if (f < A.n64_active)
                do_field<SCHEME, TX, TY, double>(A, f, k, tid, tx, ty, i0, j0, valid, id, z64[f], fzb64[f], double(ue),
                                                 double(vn), double(uw_e), double(vs_e), double(wt), dz1w_k, tile64,
                                                 fx64, fy64, hal_r, hal_c, h64[f], hal_ok);
#pragma unroll
            for (int f = 0; f < N32; ++f)
                if (f < A.n32_active)
                do_field<SCHEME, TX, TY, float>(A, f, k, tid, tx, ty, i0, j0, valid, id, z32[f], fzb32[f], ue, vn,
                                                uw_e, vs_e, wt, dz1w_k, tile32, fx32, fy32, hal_r, hal_c, h32[f], hal_ok);
        }
        __syncthreads();
    }
}

template <int SCHEME, int N64, int N32>
hipError_t launch_combo(const FusedArgs &a0, hipStream_t s, int blocks_per_cu)
{
    FusedArgs A = a0;
    constexpr int TX = 64, TY = 4;
    A.ntx = (A.nx - 5 + TX - 1) / TX;
    A.nty = A.three_d ? (A.ny - 5 + TY - 1) / TY : 1;
    int tiles = A.ntx * A.nty;
    A.xcd_remap = (tiles % 8 == 0) ? 1 : 0;
    if (A.kchunk <= 0) A.kchunk = A.nz;
    int nch = (A.nz + A.kchunk - 1) / A.kchunk;
    A.nwork = tiles * nch;
    int dev = 0, ncu = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    int occ = 0;
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fused_kernel<SCHEME, N64, N32, TX, TY>, TX * TY, 0);
    if (occ < 1) occ = 1;
    if (blocks_per_cu > 0 && blocks_per_cu < occ) occ = blocks_per_cu;
    int grid = min(A.nwork, ncu * occ);
    hipLaunchKernelGGL((fused_kernel<SCHEME, N64, N32, TX, TY>), dim3(grid), dim3(TX * TY), 0, s, A);
    return hipGetLastError();
}

template <int SCHEME, int N64>
hipError_t launch_n32(int n32, const FusedArgs &A, hipStream_t s, int bpc)
{
    switch (n32) {
    case 0: return launch_combo<SCHEME, N64, 0>(A, s, bpc);
    case 1: return launch_combo<SCHEME, N64, 1>(A, s, bpc);
    case 2: return launch_combo<SCHEME, N64, 2>(A, s, bpc);
    case 4: return launch_combo<SCHEME, N64, 4>(A, s, bpc);
    case 8: return launch_combo<SCHEME, N64, 8>(A, s, bpc);
    case 16: return launch_combo<SCHEME, N64, 16>(A, s, bpc);
    case 18: return launch_combo<SCHEME, N64, 18>(A, s, bpc);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

int fused_padded_count(int n32)
{
    static const int sizes[] = {0, 1, 2, 4, 8, 16, 18};
    for (int s : sizes)
        if (s >= n32) return s;
    return -1;
}

hipError_t launch_fused(int scheme, int n64, int n32, const FusedArgs &A, hipStream_t s, int blocks_per_cu)
{
    if (scheme == 0) {
        switch (n64) {
        case 0: return launch_n32<0, 0>(n32, A, s, blocks_per_cu);
        case 1: return launch_n32<0, 1>(n32, A, s, blocks_per_cu);
        case 2: return launch_n32<0, 2>(n32, A, s, blocks_per_cu);
        default: return hipErrorInvalidValue;
        }
    }
    if (scheme == 2) {
        switch (n64) {
        case 0: return launch_n32<2, 0>(n32, A, s, blocks_per_cu);
        case 1: return launch_n32<2, 1>(n32, A, s, blocks_per_cu);
        case 2: return launch_n32<2, 2>(n32, A, s, blocks_per_cu);
        default: return hipErrorInvalidValue;
        }
    }
    return hipErrorInvalidValue;
}

}  // namespace mimadv
