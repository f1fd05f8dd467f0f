// fused.hip -- batched, mixed-precision scalar advection on gfx950.
//
// One persistent launch advects a batch of fields that share one velocity
// (u, v, w in FP32): up to MAX64 fields in FP64 (theta, water vapour: small
// increments on large values) and the rest in FP32.  A work item is a
// TX x TY tile of columns and a chunk of KC levels:
//
//   * the chunk's face velocities are read once into registers and serve
//     every field of the batch (fields are the outer loop, levels the
//     inner one, so register use does not grow with the field count);
//   * per field and level the tile plus a 2-cell halo goes through LDS
//     (the next level's tile is already in flight), the TY*(TX+1) x faces
//     and (TY+1)*TX y faces of the tile are evaluated once each by a
//     flattened loop over the workgroup, and the z face k+1/2 of every
//     column comes from a register window q(k-1..k+2) and is carried to
//     level k+1 as its bottom face;
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
    // hardware estimate refined by two Newton steps (error ~1 ulp)
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
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

// Geometry of one work item (a TX x TY column tile and a chunk of at most
// KC levels) shared by every field it advects.
template <int TX, int TY>
struct Item {
    static constexpr bool THREE_D = TY > 1;
    static constexpr int NT = TX * TY;
    static constexpr int LW = TX + 4, LH = THREE_D ? TY + 4 : 1;  // tile + 2-cell halo
    static constexpr int NXF = TY * (TX + 1), NYF = THREE_D ? (TY + 1) * TX : 0, NF = NXF + NYF;
    static constexpr int FS = (NF + NT - 1) / NT;        // face slots per thread
    static constexpr int TS = (LH * LW + NT - 1) / NT;   // tile slots per thread
};

// Per-thread slot geometry of a work item, computed once and shared by
// every field and level: the thread's tile slots (LDS index, column offset
// in a level plane) and face slots of the flattened list
// [x faces row-major | y faces] (LDS offset of the stencil q(f-1..f+2),
// step 1 for x and LW for y, 0 for an empty slot; plane offset of the face
// velocity u(f+1) or v(g+1)).  Plane offsets fit in 32 bits.
template <int TX, int TY>
struct Slots {
    using I = Item<TX, TY>;
    int t_off[I::TS];
    int f_base[I::FS], f_step[I::FS], f_off[I::FS];
    __device__ __forceinline__ void init(int tid, int i0, int j0, const FusedArgs &A)
    {
        const int nx = A.nx, ny = A.ny;
#pragma unroll
        for (int s = 0; s < I::TS; ++s) {
            const int e = tid + s * I::NT;
            t_off[s] = -1;
            if (e < I::LH * I::LW) {
                const int r = e / I::LW, c = e - r * I::LW;
                const int ii = wrap(i0 - 2 + c, nx, A.mx);
                const int jj = I::THREE_D ? wrap(j0 - 2 + r, ny, A.my) : 1;
                t_off[s] = (ii - 1) + nx * (jj - 1);
            }
        }
#pragma unroll
        for (int s = 0; s < I::FS; ++s) {
            const int e = tid + s * I::NT;
            f_base[s] = 0;
            f_step[s] = 0;
            f_off[s] = 0;
            if (e < I::NXF) {
                const int r = e / (TX + 1), fl = e - r * (TX + 1);  // x face f = i0-1+fl, row j0+r
                f_base[s] = (I::THREE_D ? r + 2 : 0) * I::LW + fl;
                f_step[s] = 1;
                const int jr = I::THREE_D ? wrap(j0 + r, ny, A.my) : 1;
                f_off[s] = (wrap(i0 + fl, nx, A.mx) - 1) + nx * (jr - 1);
            } else if (e < I::NF) {
                const int e2 = e - I::NXF, fr = e2 / TX, c = e2 - fr * TX;  // y face g = j0-1+fr, column i0+c
                f_base[s] = fr * I::LW + c + 2;
                f_step[s] = I::LW;
                f_off[s] = (wrap(i0 + c, nx, A.mx) - 1) + nx * (wrap(j0 + fr, ny, A.my) - 1);
            }
        }
    }
};

// One field marched up the item's levels.  Per level: the prefetched tile
// goes to LDS, the next level's tile, face velocities, window value and
// step-start value are requested, every thread evaluates its face slots
// and the top z face of its column (register window q(k-1..k+2)), then
// finishes its cell: FP64 flux divergence of the FP32/FP64 face values and
// the Shu-Osher stage update, rounded once to the field's precision, with
// the periodic ghost images.
template <int SCHEME, int TX, int TY, class T>
__device__ __forceinline__ void march_field(const FusedArgs &A, const T *__restrict__ q, const T *__restrict__ q0,
                                            T *__restrict__ out, T *lt, T *lf, int k0, int nk, int i0, int j0,
                                            int tid, int tx, int ty, bool valid, int col,
                                            const Slots<TX, TY> &sl)
{
    using I = Item<TX, TY>;
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const size_t plane = (size_t)nx * ny;
    auto lev = [&](int kk) { return (size_t)(min(max(kk, 1), nz) - 1) * plane; };
    // window q(k0-1..k0+2) with q(0) := q(1), carried bottom face of k0
    T zw[4] = {q[col + lev(k0 - 1)], q[col + lev(k0)], q[col + lev(k0 + 1)], q[col + lev(k0 + 2)]};
    if (k0 == 1) zw[0] = zw[1];
    T fzb = T(0);
    if (k0 > 1 && (A.dirmask & 4)) {
        T zz[4] = {q[col + lev(k0 - 2)], zw[0], zw[1], zw[2]};
        if (k0 == 2) zz[0] = zz[1];
        fzb = zface<SCHEME, T>(k0 - 1, nz, zz, T(A.w[col + lev(k0)]), A);
    }
    T tv[I::TS];
    float vv[I::FS], wv;
    T q0v = T(0), zn = T(0);
    double d1v = 0.0;
    const bool need_q0 = A.epi != 0 && A.c0 != 0.0;
    // everything level k needs from memory, requested one level ahead
    auto fetch = [&](int k) {
        const size_t kb = lev(k);
        const T *qk = q + kb;
#pragma unroll
        for (int s = 0; s < I::TS; ++s)
            if (sl.t_off[s] >= 0) tv[s] = qk[sl.t_off[s]];
#pragma unroll
        for (int s = 0; s < I::FS; ++s)
            vv[s] = sl.f_step[s] == 1 ? A.u[kb + sl.f_off[s]] : (sl.f_step[s] != 0 ? A.v[kb + sl.f_off[s]] : 0.0f);
        wv = k < nz ? A.w[col + kb + plane] : 0.0f;
        if (need_q0 && valid) q0v = q0[col + kb];
        zn = q[col + lev(k + 2)];
        d1v = A.dz1w[k - 1];
    };
    fetch(k0);
#pragma unroll 1
    for (int kk = 0; kk < nk; ++kk) {
        const int k = k0 + kk;
#pragma unroll
        for (int s = 0; s < I::TS; ++s)
            if (sl.t_off[s] >= 0) lt[tid + s * I::NT] = tv[s];
        float vc[I::FS];
#pragma unroll
        for (int s = 0; s < I::FS; ++s) vc[s] = vv[s];
        const float wc = wv;
        const T q0c = q0v;
        const double dz1w_k = d1v;
        // window q(k-1..k+2): the level's top value arrived one level ahead
        if (kk > 0) {
            zw[0] = zw[1];
            zw[1] = zw[2];
            zw[2] = zw[3];
            zw[3] = zn;
        }
        __syncthreads();
        if (kk + 1 < nk) fetch(k + 1);
#pragma unroll
        for (int s = 0; s < I::FS; ++s) {
            const int st = sl.f_step[s];
            if (st == 0) continue;
            const int b = sl.f_base[s];
            lf[tid + s * I::NT] = hface<SCHEME, T>(lt[b], lt[b + st], lt[b + 2 * st], lt[b + 3 * st], T(vc[s]), A,
                                                   st == 1 ? A.dx1 : A.dy1);
        }
        const T fzt = (A.dirmask & 4) ? zface<SCHEME, T>(k, nz, zw, T(wc), A) : T(0);
        __syncthreads();
        if (valid) {
            const size_t id = col + (size_t)(k - 1) * plane;
            double div = 0.0;
            if (A.dirmask & 1) {
                const int e = ty * (TX + 1) + tx;
                div += (double(lf[e + 1]) - double(lf[e])) * A.dx1;
            }
            if (I::THREE_D && (A.dirmask & 2)) {
                const int e = I::NXF + ty * TX + tx;
                div += (double(lf[e + TX]) - double(lf[e])) * A.dy1;
            }
            if (A.dirmask & 4) div += (double(fzt) - double(fzb)) * dz1w_k;
            const double a = -div;
            if (A.epi == 0) {
                out[id] = T(a);
            } else {
                double qn;
                if (A.c0 != 0.0) qn = A.c0 * double(q0c) + A.c1 * (double(zw[1]) + A.dt * a);
                else qn = A.c1 * (double(zw[1]) + A.dt * a);
                const T qv = T(qn);
                out[id] = qv;
                if (A.fill) {
                    const int i = i0 + tx, j = j0 + ty;
                    int ix = 0, jy = 0;
                    if (i >= nx - 4) ix = i - A.mx;
                    else if (i <= 5) ix = i + A.mx;
                    if (I::THREE_D) {
                        if (j >= ny - 4) jy = j - A.my;
                        else if (j <= 5) jy = j + A.my;
                    }
                    const size_t kb = (size_t)(k - 1) * plane;
                    if (ix) out[kb + (size_t)(ix - 1) + (size_t)nx * (j - 1)] = qv;
                    if (jy) {
                        out[kb + (size_t)(i - 1) + (size_t)nx * (jy - 1)] = qv;
                        if (ix) out[kb + (size_t)(ix - 1) + (size_t)nx * (jy - 1)] = qv;
                    }
                }
            }
        }
        fzb = fzt;
    }
}

// T: the precision of every field of the launch (FP64 and FP32 members of
// a batch go to two launches, so that the FP32 kernel keeps the register
// budget of FP32 arithmetic).
template <int SCHEME, int TX, int TY, class T>
__global__ __launch_bounds__(TX *TY) void fused_kernel(FusedArgs A)
{
    using I = Item<TX, TY>;
    __shared__ T tile_lds[I::LH * I::LW];
    __shared__ T face_lds[I::FS * I::NT];

    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
    const int tiles = A.ntx * A.nty;
    const int kc = A.kchunk;

    for (int wi = blockIdx.x; wi < A.nwork; wi += gridDim.x) {
        // XCD-aware order: consecutive work ids go round-robin over the 8
        // XCDs; every XCD gets a contiguous range of tiles so that tile
        // halos and velocity rows stay in one L2.
        const int chunk = wi / tiles;
        int t = wi - chunk * tiles;
        if (A.xcd_remap) {
            const int per = tiles / 8;
            t = (t % 8) * per + t / 8;
        }
        const int i0 = 4 + (t % A.ntx) * TX;
        const int j0 = I::THREE_D ? 4 + (t / A.ntx) * TY : 1;
        const int k0 = 1 + chunk * kc;
        const int nk = min(kc, nz - k0 + 1);
        const int i = i0 + tx, j = j0 + ty;
        const bool valid = i <= nx - 2 && (!I::THREE_D || j <= ny - 2);
        const int il = wrap(i, nx, A.mx), jl = I::THREE_D ? wrap(j, ny, A.my) : 1;
        const int col = (il - 1) + nx * (jl - 1);
        Slots<TX, TY> sl;
        sl.init(tid, i0, j0, A);
        if constexpr (sizeof(T) == 8) {
#pragma unroll 1
            for (int f = 0; f < A.n64_active; ++f)
                march_field<SCHEME, TX, TY, T>(A, A.q64[f], A.q0_64[f], A.out64[f], tile_lds, face_lds, k0, nk, i0, j0,
                                               tid, tx, ty, valid, col, sl);
        } else {
#pragma unroll 1
            for (int f = 0; f < A.n32_active; ++f)
                march_field<SCHEME, TX, TY, T>(A, A.q32[f], A.q0_32[f], A.out32[f], tile_lds, face_lds, k0, nk, i0, j0,
                                               tid, tx, ty, valid, col, sl);
        }
    }
}

template <int SCHEME, int TX, int TY, class T>
hipError_t launch_tiles(const FusedArgs &a0, hipStream_t s, int blocks_per_cu)
{
    FusedArgs A = a0;
    A.ntx = (A.nx - 5 + TX - 1) / TX;
    A.nty = TY > 1 ? (A.ny - 5 + TY - 1) / TY : 1;
    const int tiles = A.ntx * A.nty;
    A.xcd_remap = (tiles % 8 == 0) ? 1 : 0;
    if (A.kchunk <= 0 || A.kchunk > A.nz) A.kchunk = A.kchunk <= 0 ? min(32, A.nz) : A.nz;
    A.nwork = tiles * ((A.nz + A.kchunk - 1) / A.kchunk);
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
        ncu = 256;
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fused_kernel<SCHEME, TX, TY, T>, TX * TY, 0) !=
            hipSuccess ||
        occ < 1)
        occ = 1;
    if (blocks_per_cu > 0 && blocks_per_cu < occ) occ = blocks_per_cu;
    const int grid = min(A.nwork, ncu * occ);
    hipLaunchKernelGGL((fused_kernel<SCHEME, TX, TY, T>), dim3(grid), dim3(TX * TY), 0, s, A);
    return hipGetLastError();
}

template <int SCHEME, class T>
hipError_t launch_prec(const FusedArgs &A, hipStream_t s, int bpc)
{
    if (!A.three_d) return launch_tiles<SCHEME, 64, 1, T>(A, s, bpc);
    switch (A.tile_y) {
    case 2: return launch_tiles<SCHEME, 64, 2, T>(A, s, bpc);
    case 4: return launch_tiles<SCHEME, 64, 4, T>(A, s, bpc);
    default: return launch_tiles<SCHEME, 64, 8, T>(A, s, bpc);
    }
}

template <int SCHEME>
hipError_t launch_scheme(const FusedArgs &A, hipStream_t s, int bpc)
{
    hipError_t e = hipSuccess;
    if (A.n64_active > 0) e = launch_prec<SCHEME, double>(A, s, bpc);
    if (e == hipSuccess && A.n32_active > 0) e = launch_prec<SCHEME, float>(A, s, bpc);
    return e;
}

}  // namespace

int fused_padded_count(int n32)
{
    return (n32 >= 0 && n32 <= FUSED_MAX32) ? n32 : -1;
}

hipError_t launch_fused(int scheme, int n64, int n32, const FusedArgs &A, hipStream_t s, int blocks_per_cu)
{
    if (n64 < 0 || n64 > FUSED_MAX64 || n32 < 0 || n32 > FUSED_MAX32) return hipErrorInvalidValue;
    if (scheme == 0) return launch_scheme<0>(A, s, blocks_per_cu);
    if (scheme == 2) return launch_scheme<2>(A, s, blocks_per_cu);
    return hipErrorInvalidValue;
}

}  // namespace mimadv
