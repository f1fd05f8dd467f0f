// fused.hip -- batched, mixed-precision scalar advection on gfx950.
//
// One persistent launch advects a batch of fields that share one velocity
// (u, v, w in FP32): up to MAX64 fields in FP64 (theta, water vapour: small
// increments on large values) and the rest in FP32.  A work item is a
// TX x TY tile of columns and a chunk of KC levels:
//
//   * the chunk's face velocities are staged in LDS once and serve every
//     field of the batch (fields are the outer loop, levels the inner one,
//     so register use does not grow with the field count);
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
// bottom/top closures), MUSCL (advq_muscl.f90) and the WENO5-JS extension
// (3-cell halo, in z where the upwind stencil fits the column) in an
// operation-reduced form: limiter_weno's rescale is applied to the
// weighted face value in a form that resolves its eps scale in FP32 (see
// weno_face), only the upwind side's weights are evaluated, and divisions
// are reciprocals.  It is checked against the FP64 oracle with tolerances,
// not bit for bit.  Vertical boundaries must be no-flux ('nnf'), the
// lateral ones periodic.
#include "fused.hpp"
#include "fused_faces.hpp"

namespace mimadv {

namespace {

using namespace faces;

// Geometry of one work item (a TX x TY column tile and a chunk of at most
// KC levels) shared by every field it advects.
template <int TX, int TY, int H>
struct Item {
    static constexpr bool THREE_D = TY > 1;
    static constexpr int NT = TX * TY;
    static constexpr int LW = TX + 2 * H, LH = THREE_D ? TY + 2 * H : 1;  // tile + H-cell halo
    static constexpr int NXF = TY * (TX + 1), NYF = THREE_D ? (TY + 1) * TX : 0, NF = NXF + NYF;
    static constexpr int FS = (NF + NT - 1) / NT;        // face slots per thread
    static constexpr int TS = (LH * LW + NT - 1) / NT;   // tile slots per thread
    static constexpr int NV = NF + NT;                   // velocities per level: faces, then w of every column
};

// Levels of a work item whose velocities are staged in LDS once and read by
// every field (the chunk length is capped at this).
constexpr int FUSED_KC = 8;

// The FP32 256-thread tiles of the WENO and MUSCL schemes: six-level chunks
// take 25 KB of LDS, so six groups of four waves fit a CU, and the kernel
// is held to 80 VGPRs (a few values of the per-item velocity staging spill;
// the level loop does not).  The FP64 and WENO5 instances are bound by
// their registers and keep eight-level chunks.
template <int SCHEME, int TX, int TY, class T>
constexpr bool fused_six_waves()
{
    return SCHEME != 3 && TX * TY == 256 && TY > 1 && sizeof(T) == 4;
}
template <int SCHEME, int TX, int TY, class T>
constexpr int fused_kc()
{
    return fused_six_waves<SCHEME, TX, TY, T>() ? 6 : FUSED_KC;
}

// Face e of the flattened list [x faces row-major | y faces] of a tile:
// kind (1 x, 2 y), LDS offset/step of its stencil and the column of its
// velocity (u(f+1) of x face f + 1/2, v(g+1) of y face g + 1/2).
template <int TX, int TY, int H>
__device__ __forceinline__ int face_slot_geom(int e, int i0, int j0, const FusedArgs &A, int &base, int &step,
                                              size_t &vcol)
{
    using I = Item<TX, TY, H>;
    const int nx = A.nx, ny = A.ny;
    if (e < I::NXF) {
        const int r = e / (TX + 1), fl = e - r * (TX + 1);
        base = (I::THREE_D ? r + H : 0) * I::LW + fl;
        step = 1;
        const int jr = I::THREE_D ? wrap(j0 + r, ny, A.my) : 1;
        vcol = (size_t)(wrap(i0 + fl, nx, A.mx) - 1) + (size_t)nx * (jr - 1);
        return 1;
    }
    const int e2 = e - I::NXF, fr = e2 / TX, c = e2 - fr * TX;
    base = fr * I::LW + c + H;
    step = I::LW;
    vcol = (size_t)(wrap(i0 + c, nx, A.mx) - 1) + (size_t)nx * (wrap(j0 + fr, ny, A.my) - 1);
    return 2;
}

// Per-thread slot geometry of a work item, computed once and shared by
// every field and level: the thread's tile slots (column offset in a level
// plane, -1 if none) and face slots of the flattened list
// [x faces row-major | y faces] (LDS offset of the stencil q(f-1..f+2),
// step 1 for x and LW for y, 0 for an empty slot).  Plane offsets fit in
// 32 bits.
template <int TX, int TY, int H>
struct Slots {
    using I = Item<TX, TY, H>;
    int t_off[I::TS];
    int f_base[I::FS], f_step[I::FS];
    __device__ __forceinline__ void init(int tid, int i0, int j0, const FusedArgs &A)
    {
        const int nx = A.nx, ny = A.ny;
#pragma unroll
        for (int s = 0; s < I::TS; ++s) {
            const int e = tid + s * I::NT;
            t_off[s] = -1;
            if (e < I::LH * I::LW) {
                const int r = e / I::LW, c = e - r * I::LW;
                const int ii = wrap(i0 - H + c, nx, A.mx);
                const int jj = I::THREE_D ? wrap(j0 - H + r, ny, A.my) : 1;
                t_off[s] = (ii - 1) + nx * (jj - 1);
            }
        }
#pragma unroll
        for (int s = 0; s < I::FS; ++s) {
            const int e = tid + s * I::NT;
            f_base[s] = 0;
            f_step[s] = 0;
            if (e < I::NXF) {
                const int r = e / (TX + 1), fl = e - r * (TX + 1);  // x face f = i0-1+fl, row j0+r
                f_base[s] = (I::THREE_D ? r + H : 0) * I::LW + fl;
                f_step[s] = 1;
            } else if (e < I::NF) {
                const int e2 = e - I::NXF, fr = e2 / TX, c = e2 - fr * TX;  // y face g = j0-1+fr, column i0+c
                f_base[s] = fr * I::LW + c + H;
                f_step[s] = I::LW;
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
                                            T *__restrict__ out, T *lt, T *lf, const float *lv, int k0, int nk,
                                            int i0, int j0, int tid, int tx, int ty, bool valid, int col,
                                            const Slots<TX, TY, stencil_h<SCHEME>()> &sl)
{
    constexpr int H = stencil_h<SCHEME>();
    using I = Item<TX, TY, H>;
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const size_t plane = (size_t)nx * ny;
    auto lev = [&](int kk) { return (size_t)(min(max(kk, 1), nz) - 1) * plane; };
    // A chunk above the floor first evaluates its bottom face k0-1/2 in a
    // z-only pass of the level loop (kk = -1), so that every z face of the
    // column is evaluated by the same code whatever the chunking.
    const bool zdir = (A.dirmask & 4) != 0;
    const int kstart = (k0 > 1 && zdir) ? -1 : 0;
    const int ks = k0 + kstart;
    // byte offsets in a level plane: this thread's column, its tile slots
    const uint32_t colb = (uint32_t)col * (uint32_t)sizeof(T);
    uint32_t tb[I::TS];
#pragma unroll
    for (int s = 0; s < I::TS; ++s) tb[s] = (uint32_t)sl.t_off[s] * (uint32_t)sizeof(T);
    // window q(ks-H+1..ks+H), levels outside 1..nz clamped
    T zw[2 * H];
#pragma unroll
    for (int m = 0; m < 2 * H; ++m) zw[m] = load_at(q + lev(ks - H + 1 + m), colb);
    const float wz0 = load_at(A.w + lev(ks + 1), (uint32_t)col * 4u);
    T fzb = T(0);
    T tv[I::TS];
    T q0v = T(0), zn = T(0);
    double d1v = 0.0;
    const bool need_q0 = A.epi != 0 && A.c0 != 0.0;
    // everything level k needs from memory, requested one level ahead; zn
    // = q(k+H) enters the window after level k-1
    auto fetch = [&](int k) {
        const size_t kb = lev(k);
        const T *qk = q + kb;
#pragma unroll
        for (int s = 0; s < I::TS; ++s)
            if (sl.t_off[s] >= 0) tv[s] = load_at(qk, tb[s]);
        if (need_q0 && valid) q0v = load_at(q0 + kb, colb);
        zn = load_at(q + lev(k + H), colb);
        d1v = A.dz1w[k - 1];
    };
    fetch(k0);
#pragma unroll 1
    for (int kk = kstart; kk < nk; ++kk) {
        const int k = k0 + kk;
        // window q(k-H+1..k+H); this level's staged velocities
        const float *lvk = lv + (kk < 0 ? 0 : kk) * I::NV;
        const T fzt = zdir ? zface<SCHEME, T>(k, nz, zw, T(kk < 0 ? wz0 : lvk[I::NF + tid]), A) : T(0);
        if (kk >= 0) {
#pragma unroll
            for (int s = 0; s < I::TS; ++s)
                if (sl.t_off[s] >= 0) lt[tid + s * I::NT] = tv[s];
            const T q0c = q0v;
            const double dz1w_k = d1v;
            __syncthreads();
            if (kk + 1 < nk) fetch(k + 1);
#pragma unroll
            for (int s = 0; s < I::FS; ++s) {
                const int st = sl.f_step[s];
                if (st == 0) continue;
                const int b = sl.f_base[s];
                T sv[2 * H];
#pragma unroll
                for (int m = 0; m < 2 * H; ++m) sv[m] = lt[b + m * st];
                lf[tid + s * I::NT] = hface<SCHEME, T>(sv, T(lvk[tid + s * I::NT]), A, st == 1 ? A.dx1 : A.dy1);
            }
            __syncthreads();
            if (valid) {
                // flux divergence in FP64: the differences of FP32 face
                // values are exact there, and the three terms are summed
                // with one rounding per term (a compensated FP32 sum)
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
                T *outk = out + (size_t)(k - 1) * plane;
                if (A.epi == 0) {
                    store_at(outk, colb, T(a));
                } else {
                    // Shu-Osher stage in FP64, rounded once
                    double qn;
                    if (A.c0 != 0.0) qn = A.c0 * double(q0c) + A.c1 * (double(zw[H - 1]) + A.dt * a);
                    else qn = A.c1 * (double(zw[H - 1]) + A.dt * a);
                    const T qv = T(qn);
                    store_at(outk, colb, qv);
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
        }
        fzb = fzt;
#pragma unroll
        for (int m = 0; m < 2 * H - 1; ++m) zw[m] = zw[m + 1];
        zw[2 * H - 1] = zn;
    }
}

// T: the precision of every field of the launch (FP64 and FP32 members of
// a batch go to two launches, so that the FP32 kernel keeps the register
// budget of FP32 arithmetic).
// Waves per SIMD the register allocation must allow: six for the FP32
// 256-thread tiles (see fused_kc).  The other instances are bound by LDS or
// need their registers (WENO5's six-value stencils among them).
template <int SCHEME, int TX, int TY, class T>
constexpr int fused_min_waves()
{
    return fused_six_waves<SCHEME, TX, TY, T>() ? 6 : 1;
}

template <int SCHEME, int TX, int TY, class T>
__global__ __launch_bounds__(TX *TY) __attribute__((amdgpu_waves_per_eu(fused_min_waves<SCHEME, TX, TY, T>(), 8))) void
fused_kernel(FusedArgs A)
{
    constexpr int H = stencil_h<SCHEME>();
    using I = Item<TX, TY, H>;
    __shared__ T tile_lds[I::LH * I::LW];
    __shared__ T face_lds[I::FS * I::NT];
    constexpr int KC = fused_kc<SCHEME, TX, TY, T>();
    __shared__ float vel_lds[KC * I::NV];

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
        Slots<TX, TY, H> sl;
        sl.init(tid, i0, j0, A);
        // the chunk's face velocities and column w(k+1), read once for every
        // field: the source array and plane offset of each of this thread's
        // staging slots are fixed for the item, and all the chunk's loads are
        // in flight together (issued before the barrier that waits for the
        // previous item's fields) before they go to LDS
        {
            constexpr int VS = (I::NV + I::NT - 1) / I::NT;
            const size_t plane = (size_t)nx * ny;
            const float *vp[VS];
            int voff[VS];
            bool vw[VS];
#pragma unroll
            for (int s = 0; s < VS; ++s) {
                const int e = tid + s * I::NT;
                vp[s] = A.u;
                voff[s] = -1;
                vw[s] = false;
                if (e < I::NF) {
                    int base, step;
                    size_t vcol;
                    const int kind = face_slot_geom<TX, TY, H>(e, i0, j0, A, base, step, vcol);
                    vp[s] = kind == 1 ? A.u : A.v;
                    voff[s] = (int)vcol;
                } else if (e < I::NV) {
                    const int c = e - I::NF, cx = c % TX, cy = c / TX;
                    const int ci = wrap(i0 + cx, nx, A.mx), cj = I::THREE_D ? wrap(j0 + cy, ny, A.my) : 1;
                    vp[s] = A.w;  // w(k+1) at level k
                    voff[s] = (ci - 1) + nx * (cj - 1);
                    vw[s] = true;
                }
            }
            float vb[KC][VS];
#pragma unroll
            for (int kk = 0; kk < KC; ++kk) {
                const int k = k0 + kk;
#pragma unroll
                for (int s = 0; s < VS; ++s) {
                    vb[kk][s] = 0.0f;
                    if (kk < nk && voff[s] >= 0 && (!vw[s] || k < nz))
                        vb[kk][s] = vp[s][(size_t)(vw[s] ? k : k - 1) * plane + (size_t)voff[s]];
                }
            }
            __syncthreads();  // the previous item's fields are done with vel_lds
#pragma unroll
            for (int kk = 0; kk < KC; ++kk)
#pragma unroll
                for (int s = 0; s < VS; ++s)
                    if (kk < nk && tid + s * I::NT < I::NV) vel_lds[kk * I::NV + tid + s * I::NT] = vb[kk][s];
            __syncthreads();
        }
        if constexpr (sizeof(T) == 8) {
#pragma unroll 1
            for (int f = 0; f < A.n64_active; ++f)
                march_field<SCHEME, TX, TY, T>(A, A.q64[f], A.q0_64[f], A.out64[f], tile_lds, face_lds, vel_lds, k0, nk,
                                               i0, j0, tid, tx, ty, valid, col, sl);
        } else {
#pragma unroll 1
            for (int f = 0; f < A.n32_active; ++f)
                march_field<SCHEME, TX, TY, T>(A, A.q32[f], A.q0_32[f], A.out32[f], tile_lds, face_lds, vel_lds, k0, nk,
                                               i0, j0, tid, tx, ty, valid, col, sl);
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
    constexpr int KC = fused_kc<SCHEME, TX, TY, T>();
    A.kchunk = A.kchunk <= 0 ? KC : min(A.kchunk, KC);
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
    // default 32x8: the fewest x/y faces per cell of the 256-thread tiles
    // (552 for 256 cells against 580 for 64x4) and the smallest halo tile
    if (A.tile_x == 32 || (A.tile_x == 0 && A.tile_y == 0)) return launch_tiles<SCHEME, 32, 8, T>(A, s, bpc);
    if (A.tile_x == 16) return launch_tiles<SCHEME, 16, 16, T>(A, s, bpc);
    const int ty = A.tile_y > 0 ? A.tile_y : 4;
    switch (ty) {
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
    if (scheme == 3) return launch_scheme<3>(A, s, blocks_per_cu);
    return hipErrorInvalidValue;
}

}  // namespace mimadv
