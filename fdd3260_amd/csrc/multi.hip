// multi.hip -- one launch advecting many independent periodic subdomains.
//
// A batch of unequal subdomains (each with its own grid, metrics, FP32
// velocity, dt, scheme and set of FP32 fields) is advanced by one
// persistent launch per Shu-Osher stage.  The subdomains' work items (a
// tile of columns and a chunk of levels) are numbered consecutively; a
// workgroup finds its subdomain by a binary search of the first work id of
// every subdomain, reads that subdomain's argument block from a device
// table, and runs the same per-field column march as the single-grid
// batched kernel (fused.hip): tile plus a 2-cell (WENO5: 3-cell) halo in
// LDS, flattened x/y face list, z face carried up the column, FP64
// divergence and stage update rounded once to FP32.
//
// The tile shape is chosen per subdomain (TX x TY with TX*TY = 256:
// 16x16, 32x8, 64x4) so that small or narrow subdomains waste fewer lanes.
// 256-thread groups keep the LDS of a work item at 32 KB, so five groups
// share a CU (five waves per SIMD at 94 VGPRs); faces are evaluated with the same formulas as the
// single-grid kernel, so the results do not depend on the shape.
#include "fused.hpp"
#include "fused_faces.hpp"

namespace mimadv {

namespace {

using namespace faces;

constexpr int NT = MULTI_NT;
constexpr int FS = 3;                  // 2*NT + TX + TY faces per tile
constexpr int TS = 3;                  // (TX+4)*(TY+4) tile values
constexpr int TILE_MAX = 70 * 10;      // largest (TX+2H)*(TY+2H) of the shapes (64x4, H = 3)
constexpr int KC_MAX = MULTI_KC;       // levels of a work item (velocities staged in LDS)
constexpr int NV_MAX = 2 * NT + 68 + NT;  // face velocities and column w of one level

struct Geom {
    int TX, TY, H, LW, LH, NXF, NF, NV;
    __device__ __forceinline__ void set(int tx_, int ty_, int h_)
    {
        TX = tx_;
        TY = ty_;
        H = h_;
        LW = TX + 2 * H;
        LH = TY + 2 * H;
        NXF = TY * (TX + 1);
        NF = NXF + (TY + 1) * TX;
        NV = NF + NT;
    }
};

struct SlotsR {
    int t_off[TS];
    int f_base[FS], f_step[FS];
    __device__ __forceinline__ void init(const Geom &G, int tid, int i0, int j0, const FusedArgs &A)
    {
        const int nx = A.nx, ny = A.ny;
#pragma unroll
        for (int s = 0; s < TS; ++s) {
            const int e = tid + s * NT;
            t_off[s] = -1;
            if (e < G.LH * G.LW) {
                const int r = e / G.LW, c = e - r * G.LW;
                t_off[s] = (wrap(i0 - G.H + c, nx, A.mx) - 1) + nx * (wrap(j0 - G.H + r, ny, A.my) - 1);
            }
        }
#pragma unroll
        for (int s = 0; s < FS; ++s) {
            const int e = tid + s * NT;
            f_base[s] = 0;
            f_step[s] = 0;
            if (e < G.NXF) {
                const int r = e / (G.TX + 1), fl = e - r * (G.TX + 1);  // x face f = i0-1+fl, row j0+r
                f_base[s] = (r + G.H) * G.LW + fl;
                f_step[s] = 1;
            } else if (e < G.NF) {
                const int e2 = e - G.NXF, fr = e2 / G.TX, c = e2 - fr * G.TX;  // y face g = j0-1+fr
                f_base[s] = fr * G.LW + c + G.H;
                f_step[s] = G.LW;
            }
        }
    }
};

// One field of one work item; SCHEME is the subdomain's scheme (uniform
// over the work item).
template <int SCHEME>
__device__ __forceinline__ void march_multi(const FusedArgs &A, const Geom &G, const float *__restrict__ q,
                                            const float *__restrict__ q0, float *__restrict__ out, float *lt,
                                            float *lf, const float *lv, int k0, int nk, int i0, int j0, int tid,
                                            int tx, int ty, bool valid, int col, const SlotsR &sl)
{
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const size_t plane = (size_t)nx * ny;
    auto lev = [&](int kk) { return (size_t)(min(max(kk, 1), nz) - 1) * plane; };
    // a chunk above the floor first evaluates its bottom face in a z-only
    // pass of the level loop: every z face comes from the same code
    constexpr int H = stencil_h<SCHEME>();
    const int kstart = k0 > 1 ? -1 : 0;
    const int ks = k0 + kstart;
    // byte offsets in a level plane: this thread's column, its tile slots
    const uint32_t colb = (uint32_t)col * 4u;
    uint32_t tb[TS];
#pragma unroll
    for (int s = 0; s < TS; ++s) tb[s] = (uint32_t)sl.t_off[s] * 4u;
    // window q(ks-H+1..ks+H), levels below 1 read as q(1)
    float zw[2 * H];
#pragma unroll
    for (int m = 0; m < 2 * H; ++m) zw[m] = load_at(q + lev(ks - H + 1 + m), colb);
    const float wz0 = ks < nz ? load_at(A.w + lev(ks + 1), colb) : 0.0f;
    float fzb = 0.0f;
    float tv[TS], q0v = 0.0f, zn = 0.0f;
    double d1v = 0.0;
    const bool need_q0 = A.epi != 0 && A.c0 != 0.0;
    auto fetch = [&](int k) {
        const size_t kb = lev(k);
        const float *qk = q + kb;
#pragma unroll
        for (int s = 0; s < TS; ++s)
            if (sl.t_off[s] >= 0) tv[s] = load_at(qk, tb[s]);
        if (need_q0 && valid) q0v = load_at(q0 + kb, colb);
        zn = load_at(q + lev(k + H), colb);
        d1v = A.dz1w[k - 1];
    };
    const int ex = ty * (G.TX + 1) + tx, ey = G.NXF + ty * G.TX + tx;
    fetch(k0);
#pragma unroll 1
    for (int kk = kstart; kk < nk; ++kk) {
        const int k = k0 + kk;
        if (kk > kstart) {
#pragma unroll
            for (int m = 0; m < 2 * H - 1; ++m) zw[m] = zw[m + 1];
            zw[2 * H - 1] = zn;
        }
        const float *lvk = lv + (kk < 0 ? 0 : kk) * G.NV;  // this level's staged velocities
        const float fzt = zface<SCHEME, float>(k, nz, zw, kk < 0 ? wz0 : lvk[G.NF + tid], A);
        if (kk < 0) {
            fzb = fzt;
            continue;
        }
#pragma unroll
        for (int s = 0; s < TS; ++s)
            if (sl.t_off[s] >= 0) lt[tid + s * NT] = tv[s];
        const float q0c = q0v;
        const double dz1w_k = d1v;
        __syncthreads();
        if (kk + 1 < nk) fetch(k + 1);
#pragma unroll
        for (int s = 0; s < FS; ++s) {
            const int st = sl.f_step[s];
            if (st == 0) continue;
            const int b = sl.f_base[s];
            float sv[2 * H];
#pragma unroll
            for (int m = 0; m < 2 * H; ++m) sv[m] = lt[b + m * st];
            lf[tid + s * NT] = hface<SCHEME, float>(sv, lvk[tid + s * NT], A, st == 1 ? A.dx1 : A.dy1);
        }
        __syncthreads();
        if (valid) {
            float *outk = out + (size_t)(k - 1) * plane;
            double div = (double(lf[ex + 1]) - double(lf[ex])) * A.dx1;
            div += (double(lf[ey + G.TX]) - double(lf[ey])) * A.dy1;
            div += (double(fzt) - double(fzb)) * dz1w_k;
            const double a = -div;
            if (A.epi == 0) {
                store_at(outk, colb, float(a));
            } else {
                double qn;
                if (A.c0 != 0.0) qn = A.c0 * double(q0c) + A.c1 * (double(zw[H - 1]) + A.dt * a);
                else qn = A.c1 * (double(zw[H - 1]) + A.dt * a);
                const float qv = float(qn);
                store_at(outk, colb, qv);
                if (A.fill) {
                    const int i = i0 + tx, j = j0 + ty;
                    int ix = 0, jy = 0;
                    if (i >= nx - 4) ix = i - A.mx;
                    else if (i <= 5) ix = i + A.mx;
                    if (j >= ny - 4) jy = j - A.my;
                    else if (j <= 5) jy = j + A.my;
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

// Six-level work items take 25 KB of LDS; the register budget is held to
// five waves per SIMD (96 VGPRs; three values are reloaded per level).
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(5, 8))) void multi_kernel(
    const FusedArgs *__restrict__ table, int ndom, int nwork)
{
    __shared__ float tile_lds[TILE_MAX];
    __shared__ float face_lds[FS * NT];
    __shared__ float vel_lds[KC_MAX * NV_MAX];
    const int tid = threadIdx.x;
    for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
        // subdomain of this work id: last d with work_begin <= wi
        int lo = 0, hi = ndom - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (table[mid].work_begin <= wi) lo = mid;
            else hi = mid - 1;
        }
        // a private copy of the argument block: its scalars stay in registers
        // instead of being reloaded after every store (the table is global
        // memory); the field pointers are read from the table once per field
        const FusedArgs A = table[lo];
        Geom G;
        G.set(A.tile_x, A.tile_y, A.scheme == 3 ? 3 : 2);
        const int local = wi - A.work_begin;
        const int tiles = A.ntx * A.nty;
        const int chunk = local / tiles, t = local - chunk * tiles;
        const int i0 = 4 + (t % A.ntx) * G.TX, j0 = 4 + (t / A.ntx) * G.TY;
        const int k0 = 1 + chunk * A.kchunk, nk = min(A.kchunk, A.nz - k0 + 1);
        const int tx = tid % G.TX, ty = tid / G.TX;
        const int i = i0 + tx, j = j0 + ty;
        const bool valid = i <= A.nx - 2 && j <= A.ny - 2;
        const int col = (wrap(i, A.nx, A.mx) - 1) + A.nx * (wrap(j, A.ny, A.my) - 1);
        SlotsR sl;
        sl.init(G, tid, i0, j0, A);
        // the chunk's face velocities and column w(k+1), read once for every
        // field: each staging slot's source array and plane offset are fixed
        // for the item, and all the chunk's loads are in flight together
        // (issued before the barrier that waits for the previous item's
        // fields) before they go to LDS
        {
            constexpr int VS = (NV_MAX + NT - 1) / NT;
            const int nx = A.nx, ny = A.ny;
            const size_t plane = (size_t)nx * ny;
            const float *vp[VS];
            int voff[VS];
            bool vw[VS];
#pragma unroll
            for (int s = 0; s < VS; ++s) {
                const int e = tid + s * NT;
                vp[s] = A.u;
                voff[s] = -1;
                vw[s] = false;
                if (e < G.NXF) {
                    const int r = e / (G.TX + 1), fl = e - r * (G.TX + 1);
                    voff[s] = (wrap(i0 + fl, nx, A.mx) - 1) + nx * (wrap(j0 + r, ny, A.my) - 1);
                } else if (e < G.NF) {
                    const int e2 = e - G.NXF, fr = e2 / G.TX, c = e2 - fr * G.TX;
                    vp[s] = A.v;
                    voff[s] = (wrap(i0 + c, nx, A.mx) - 1) + nx * (wrap(j0 + fr, ny, A.my) - 1);
                } else if (e < G.NV) {
                    const int c = e - G.NF, cx = c % G.TX, cy = c / G.TX;
                    vp[s] = A.w;  // w(k+1) at level k
                    voff[s] = (wrap(i0 + cx, nx, A.mx) - 1) + nx * (wrap(j0 + cy, ny, A.my) - 1);
                    vw[s] = true;
                }
            }
            float vb[KC_MAX][VS];
#pragma unroll
            for (int kk = 0; kk < KC_MAX; ++kk) {
                const int k = k0 + kk;
#pragma unroll
                for (int s = 0; s < VS; ++s) {
                    vb[kk][s] = 0.0f;
                    if (kk < nk && voff[s] >= 0 && (!vw[s] || k < A.nz))
                        vb[kk][s] = vp[s][(size_t)(vw[s] ? k : k - 1) * plane + (size_t)voff[s]];
                }
            }
            __syncthreads();  // the previous item's fields are done with vel_lds
#pragma unroll
            for (int kk = 0; kk < KC_MAX; ++kk)
#pragma unroll
                for (int s = 0; s < VS; ++s)
                    if (kk < nk && tid + s * NT < G.NV) vel_lds[kk * G.NV + tid + s * NT] = vb[kk][s];
            __syncthreads();
        }
#pragma unroll 1
        for (int f = 0; f < A.n32_active; ++f) {
            const float *qf = table[lo].q32[f], *q0f = table[lo].q0_32[f];
            float *of = table[lo].out32[f];
            if (A.scheme == 0)
                march_multi<0>(A, G, qf, q0f, of, tile_lds, face_lds, vel_lds, k0, nk, i0, j0, tid, tx, ty, valid, col, sl);
            else if (A.scheme == 3)
                march_multi<3>(A, G, qf, q0f, of, tile_lds, face_lds, vel_lds, k0, nk, i0, j0, tid, tx, ty, valid, col, sl);
            else
                march_multi<2>(A, G, qf, q0f, of, tile_lds, face_lds, vel_lds, k0, nk, i0, j0, tid, tx, ty, valid, col, sl);
        }
    }
}

}  // namespace

bool multi_tile_ok(int tile_x, int tile_y)
{
    return tile_x * tile_y == MULTI_NT && (tile_x == 16 || tile_x == 32 || tile_x == 64);
}

void multi_prepare(FusedArgs &A, int work_begin)
{
    A.ntx = (A.nx - 5 + A.tile_x - 1) / A.tile_x;
    A.nty = (A.ny - 5 + A.tile_y - 1) / A.tile_y;
    A.kchunk = A.kchunk <= 0 ? KC_MAX : min(A.kchunk, KC_MAX);
    A.nwork = A.ntx * A.nty * ((A.nz + A.kchunk - 1) / A.kchunk);
    A.work_begin = work_begin;
}

hipError_t launch_multi(const FusedArgs *d_table, int ndom, int nwork, hipStream_t s)
{
    if (ndom <= 0 || nwork <= 0) return hipSuccess;
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
        ncu = 256;
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, multi_kernel, NT, 0) != hipSuccess || occ < 1) occ = 1;
    const int grid = min(nwork, ncu * occ);
    hipLaunchKernelGGL(multi_kernel, dim3(grid), dim3(NT), 0, s, d_table, ndom, nwork);
    return hipGetLastError();
}

}  // namespace mimadv
