// util.hip -- ghost fill, stage combine and interior reductions.
#include <algorithm>
#include <cmath>
#include <mutex>

#include "mimadv_internal.hpp"

namespace mimadv {

namespace {

// qnbc_p4 (bc.f90:378-503), serial branch.  The reference runs the x copy,
// then the y copy over full rows, then the z copy, then the mirror fills;
// each pass is one launch here so later passes see earlier results.
__global__ void fill_x_kernel(int nx, int ny, int nz, int bc_w, int bc_e, double *x)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int k = blockIdx.y + 1;
    if (j > ny) return;
    double *r = x + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (k - 1));
    auto X = [&](int i) -> double & { return r[i - 1]; };
    if (bc_w == 0 || bc_e == 0) {
        X(1) = X(nx - 4);
        X(2) = X(nx - 3);
        X(3) = X(nx - 2);
        X(nx) = X(5);
        X(nx - 1) = X(4);
    }
}

__global__ void fill_y_kernel(int nx, int ny, int nz, double *x)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int k = blockIdx.y + 1;
    if (i > nx) return;
    size_t sy = (size_t)nx;
    double *c = x + (size_t)(i - 1) + (size_t)nx * ny * (k - 1);
    auto X = [&](int j) -> double & { return c[sy * (j - 1)]; };
    X(1) = X(ny - 4);
    X(2) = X(ny - 3);
    X(3) = X(ny - 2);
    X(ny) = X(5);
    X(ny - 1) = X(4);
}

__global__ void fill_z_kernel(int nx, int ny, int nz, double *x)
{
    size_t n = (size_t)nx * ny;
    size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    auto X = [&](int k) -> double & { return x[p + n * (k - 1)]; };
    X(1) = X(nz - 4);
    X(2) = X(nz - 3);
    X(3) = X(nz - 2);
    X(nz) = X(5);
    X(nz - 1) = X(4);
}

// no-flux / open mirror fills (bc.f90:467-497)
__global__ void mirror_x_kernel(int nx, int ny, int nz, int bc_w, int bc_e, double *x)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int k = blockIdx.y + 1;
    if (j > ny) return;
    double *r = x + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (k - 1));
    if (bc_w == 1 || bc_w == 2) {
        r[1] = r[3];
        r[0] = r[4];
    }
    if (bc_e == 1 || bc_e == 2) {
        r[nx - 1] = r[nx - 5];
        r[nx - 2] = r[nx - 4];
    }
}

__global__ void mirror_y_kernel(int nx, int ny, int nz, int bc_s, int bc_n, double *x)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int k = blockIdx.y + 1;
    if (i > nx) return;
    size_t sy = (size_t)nx;
    double *c = x + (size_t)(i - 1) + (size_t)nx * ny * (k - 1);
    if (bc_s == 1 || bc_s == 2) {
        c[sy * 1] = c[sy * 3];
        c[0] = c[sy * 4];
    }
    if (bc_n == 1 || bc_n == 2) {
        c[sy * (ny - 1)] = c[sy * (ny - 5)];
        c[sy * (ny - 2)] = c[sy * (ny - 4)];
    }
}

// lbcu_p4 side fills of the normal wind component (bc.f90:255-281 for u in
// x, :307-333 for v in y): 'nnf' mirrors with a sign change, 'ope' copies.
// stride 1 for u along x, nx for v along y; one thread per line.
__global__ void normal_fill_kernel(int n, size_t stride, size_t nlines, size_t line_step_a, int line_count_a,
                                   size_t line_step_b, int bc_lo, int bc_hi, double *x)
{
    size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlines) return;
    const size_t la = l % (size_t)line_count_a, lb = l / (size_t)line_count_a;
    double *r = x + la * line_step_a + lb * line_step_b;
    auto X = [&](int i) -> double & { return r[stride * (size_t)(i - 1)]; };
    if (bc_lo == 1 || bc_lo == 2) {
        const double sg = bc_lo == 1 ? -1.0 : 1.0;
        X(3) = sg * X(4);
        X(2) = sg * X(5);
        X(1) = sg * X(6);
    }
    if (bc_hi == 1 || bc_hi == 2) {
        const double sg = bc_hi == 1 ? -1.0 : 1.0;
        X(n) = sg * X(n - 3);
        X(n - 1) = sg * X(n - 2);
    }
}

// integ_mom (integpack.f90:54-135) for one component: the integ_scal
// combine, with the step scaled by the sponge factor tw for w (tw = 1 for
// u and v)
__global__ void integ_mom_kernel(long n, int sorder, double a_lm, double b_lm, double dt, double tw,
                                 const double *xold, const double *__restrict__ a, double *__restrict__ tend,
                                 double *x)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double upd;
    if (sorder > 1) {
        double t = a[i] + b_lm * tend[i];
        tend[i] = t;
        upd = xold[i] + a_lm * dt * t;
    } else {
        upd = xold[i] + dt * a[i];
    }
    x[i] = tw == 1.0 ? upd : tw * upd;
}

__global__ void integ_kernel(long n, int sorder, double a_lm, double b_lm, double dt, const double *__restrict__ xold,
                             const double *__restrict__ a, double *__restrict__ tend, double *__restrict__ x)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (sorder > 1) {
        double t = a[i] + b_lm * tend[i];
        tend[i] = t;
        x[i] = xold[i] + a_lm * dt * t;
    } else {
        x[i] = xold[i] + dt * a[i];
    }
}

// Interior sums per level: plain and dzw-weighted, tree reduction within a
// block (deterministic), one block per level.
__global__ __launch_bounds__(256) void level_sums_kernel(int nx, int ny, int nz, const double *__restrict__ x,
                                                         double *__restrict__ part)
{
    const bool three_d = ny > 1;
    int k = blockIdx.x + 1;
    int j0 = three_d ? 4 : 1, j1 = three_d ? ny - 2 : 1;
    int mx = nx - 5, my = j1 - j0 + 1;
    long cnt = (long)mx * my;
    const double *p = x + (size_t)nx * ny * (k - 1);
    double s = 0.0;
    for (long t = threadIdx.x; t < cnt; t += blockDim.x) {
        int jj = (int)(t / mx), ii = (int)(t - (long)jj * mx);
        s += p[(size_t)(ii + 3) + (size_t)nx * (jj + j0 - 1)];
    }
    __shared__ double sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[k - 1] = sh[0];
}

__global__ void finish_sums_kernel(int nz, long ncells, const double *__restrict__ part, const double *__restrict__ dzw,
                                   double *__restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0, sw = 0.0;
    for (int k = 0; k < nz; ++k) {
        s += part[k];
        sw += dzw[k] * part[k];
    }
    out[0] = s / (double)ncells;
    out[1] = sw;
}

// dzw-weighted interior sum per level of an FP32 or FP64 field (FP64
// accumulation, deterministic tree)
template <class T>
__global__ __launch_bounds__(256) void level_wsums_kernel(int nx, int ny, int nz, const T *__restrict__ x,
                                                          const double *__restrict__ dzw, double *__restrict__ part)
{
    const bool three_d = ny > 1;
    int k = blockIdx.x + 1;
    int j0 = three_d ? 4 : 1, j1 = three_d ? ny - 2 : 1;
    int mx = nx - 5, my = j1 - j0 + 1;
    long cnt = (long)mx * my;
    const T *p = x + (size_t)nx * ny * (k - 1);
    double s = 0.0;
    for (long t = threadIdx.x; t < cnt; t += blockDim.x) {
        int jj = (int)(t / mx), ii = (int)(t - (long)jj * mx);
        s += double(p[(size_t)(ii + 3) + (size_t)nx * (jj + j0 - 1)]);
    }
    __shared__ double sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[k - 1] = dzw[k - 1] * sh[0];
}

__global__ void sum_levels_kernel(int nz, const double *__restrict__ part, double *__restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int k = 0; k < nz; ++k) s += part[k];
    out[0] = s;
}

}  // namespace

namespace {

// Sum over the 64 lanes of a wavefront with cross-lane shuffles (no LDS,
// no barrier); the order of the butterfly is fixed, so the result is
// deterministic.
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Domain budgets of nf fields, first pass: block (b, f) sums dzw(k) x over
// the interior rows r = b, b + nb, ... of field f (a row is the mx interior
// values of one (j, k), read coalesced), reduces its four waves with
// shuffles and one LDS exchange, and writes part[f*nb + b].
__global__ __launch_bounds__(256) void budget_part_kernel(BudgetArgs a)
{
    const int f = blockIdx.y, b = blockIdx.x;
    const bool three_d = a.ny > 1;
    const int j0 = three_d ? 4 : 1, my = three_d ? a.ny - 5 : 1, mx = a.nx - 5;
    const long rows = (long)my * a.nz;
    const bool f32 = a.fp32[f] != 0;
    const float *x32 = (const float *)a.q[f];
    const double *x64 = (const double *)a.q[f];
    double s = 0.0;
    for (long r = b; r < rows; r += a.nb) {
        const int k = (int)(r / my), jj = (int)(r - (long)k * my);
        const size_t base = (size_t)a.nx * ((size_t)(jj + j0 - 1) + (size_t)a.ny * k) + 3;
        double rs = 0.0;
        for (int i = threadIdx.x; i < mx; i += blockDim.x) rs += f32 ? (double)x32[base + i] : x64[base + i];
        s += a.dzw[k] * rs;
    }
    s = wave_sum(s);
    __shared__ double w[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) w[wv] = s;
    __syncthreads();
    if (threadIdx.x == 0) a.part[(size_t)f * a.nb + b] = ((w[0] + w[1]) + w[2]) + w[3];
}

// Second pass: one wave per field sums its nb partials (strided lanes,
// then the shuffle tree).
__global__ __launch_bounds__(64) void budget_final_kernel(int nb, const double *__restrict__ part,
                                                          double *__restrict__ out)
{
    const int f = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += 64) s += part[(size_t)f * nb + b];
    s = wave_sum(s);
    if (threadIdx.x == 0) out[f] = s;
}

}  // namespace

int budget_blocks(const GridDev &g)
{
    const long rows = (long)(g.ny > 1 ? g.ny - 5 : 1) * g.nz;
    return (int)std::min<long>(rows, BUDGET_MAXNB);
}

// Scratch of the allocation-free budget entry point (mimadv_batch_budget_
// device): the per-block partials of up to BUDGET_MAXF fields,
// BUDGET_SCRATCH doubles (262 KB).  Each GPU holds its own copy until the
// process ends; host code uses only the address hipGetSymbolAddress gives
// for the current device, never this name.
static __device__ __align__(256) double scratch[BUDGET_SCRATCH];

namespace {

std::mutex scratch_mutex;
double *scratch_ptr[BUDGET_MAX_DEVICES] = {};

}  // namespace

hipError_t budget_scratch_setup(int device, double **out)
{
    *out = nullptr;
    if (device < 0 || device >= BUDGET_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[device]) {
        // the caller made `device` current
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        scratch_ptr[device] = (double *)p;
    }
    *out = scratch_ptr[device];
    return hipSuccess;
}

hipError_t launch_budgets(const BudgetArgs &a, int nf, double *d_out, hipStream_t s)
{
    hipLaunchKernelGGL(budget_part_kernel, dim3(a.nb, nf), dim3(256), 0, s, a);
    hipLaunchKernelGGL(budget_final_kernel, dim3(nf), dim3(64), 0, s, a.nb, a.part, d_out);
    return hipGetLastError();
}

hipError_t launch_wsum(const GridDev &g, const void *x, bool fp32, double *d_part, double *d_out, hipStream_t s)
{
    if (fp32)
        hipLaunchKernelGGL(level_wsums_kernel<float>, dim3(g.nz), dim3(256), 0, s, g.nx, g.ny, g.nz, (const float *)x,
                           g.zm.dzw, d_part);
    else
        hipLaunchKernelGGL(level_wsums_kernel<double>, dim3(g.nz), dim3(256), 0, s, g.nx, g.ny, g.nz,
                           (const double *)x, g.zm.dzw, d_part);
    hipLaunchKernelGGL(sum_levels_kernel, dim3(1), dim3(64), 0, s, g.nz, d_part, d_out);
    return hipGetLastError();
}

hipError_t launch_fill_scalar(const GridDev &g, double *x, hipStream_t s)
{
    int nx = g.nx, ny = g.ny, nz = g.nz;
    bool three_d = ny > 1;
    if (g.bcl[0] == 0 || g.bcl[1] == 0)
        hipLaunchKernelGGL(fill_x_kernel, dim3((ny + 63) / 64, nz), dim3(64), 0, s, nx, ny, nz, g.bcl[0], g.bcl[1], x);
    if (three_d && (g.bcl[2] == 0 || g.bcl[3] == 0))
        hipLaunchKernelGGL(fill_y_kernel, dim3((nx + 63) / 64, nz), dim3(64), 0, s, nx, ny, nz, x);
    if (g.bcl[4] == 0 || g.bcl[5] == 0) {
        size_t n = (size_t)nx * ny;
        hipLaunchKernelGGL(fill_z_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nx, ny, nz, x);
    }
    if (g.bcl[0] != 0 || g.bcl[1] != 0)
        hipLaunchKernelGGL(mirror_x_kernel, dim3((ny + 63) / 64, nz), dim3(64), 0, s, nx, ny, nz, g.bcl[0], g.bcl[1], x);
    if (three_d && (g.bcl[2] != 0 || g.bcl[3] != 0))
        hipLaunchKernelGGL(mirror_y_kernel, dim3((nx + 63) / 64, nz), dim3(64), 0, s, nx, ny, nz, g.bcl[2], g.bcl[3], x);
    return hipGetLastError();
}

hipError_t launch_fill_winds(const GridDev &g, double *u, double *v, double *w, hipStream_t s)
{
    int nx = g.nx, ny = g.ny, nz = g.nz;
    bool three_d = ny > 1;
    double *f[3] = {u, three_d ? v : nullptr, w};
    // periodic copies in x, y, z (each pass sees the previous ones), then
    // the normal-component fills of the non-periodic sides
    for (double *x : f)
        if (x && (g.bcl[0] == 0 || g.bcl[1] == 0))
            hipLaunchKernelGGL(fill_x_kernel, dim3((ny + 63) / 64, nz), dim3(64), 0, s, nx, ny, nz, 0, 0, x);
    for (double *x : f)
        if (x && three_d && (g.bcl[2] == 0 || g.bcl[3] == 0))
            hipLaunchKernelGGL(fill_y_kernel, dim3((nx + 63) / 64, nz), dim3(64), 0, s, nx, ny, nz, x);
    for (double *x : f)
        if (x && (g.bcl[4] == 0 || g.bcl[5] == 0)) {
            size_t n = (size_t)nx * ny;
            hipLaunchKernelGGL(fill_z_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nx, ny, nz, x);
        }
    if (g.bcl[0] != 0 || g.bcl[1] != 0) {
        // u along x: one line per (j, k)
        size_t nl = (size_t)ny * nz;
        hipLaunchKernelGGL(normal_fill_kernel, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, s, nx, (size_t)1,
                           nl, (size_t)nx, ny, (size_t)nx * ny, g.bcl[0], g.bcl[1], u);
    }
    if (three_d && (g.bcl[2] != 0 || g.bcl[3] != 0)) {
        // v along y: one line per (i, k)
        size_t nl = (size_t)nx * nz;
        hipLaunchKernelGGL(normal_fill_kernel, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, s, ny, (size_t)nx,
                           nl, (size_t)1, nx, (size_t)nx * ny, g.bcl[2], g.bcl[3], v);
    }
    return hipGetLastError();
}

hipError_t launch_integ_mom_one(const GridDev &g, int comp, int sorder, int im, double dt, double tw,
                                const double *old, const double *a, double *tend, double *x, hipStream_t s)
{
    (void)comp;
    double a_lm = 1.0, b_lm = 0.0;
    if (sorder > 1 && im != 1) {
        a_lm = 0.5;
        b_lm = 1.0;
    }
    long n = (long)g.nx * g.ny * g.nz;
    unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(integ_mom_kernel, dim3(nb), dim3(256), 0, s, n, sorder, a_lm, b_lm, dt, tw, old, a, tend, x);
    return hipGetLastError();
}

hipError_t launch_integ_mom(const GridDev &g, int sorder, int im, double dt, double tw, const double *const *old,
                            const double *const *a, double *const *tend, double *const *x, hipStream_t s)
{
    for (int c = 0; c < 3; ++c) {
        if (c == 1 && g.ny == 1) continue;
        hipError_t e = launch_integ_mom_one(g, c, sorder, im, dt, c == 2 ? tw : 1.0, old[c], a[c],
                                            tend ? tend[c] : nullptr, x[c], s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

double sponge_factor(int nz, int sorder, int im, double dt, const double *z0, double zmax, double zdamp, int sponge)
{
    // integpack.f90:83-96 as the reference runs it: the level loop assigns
    // the whole tauw array, so the last level at or above zdamp sets one
    // factor for every level (pi and taud are the reference's constants)
    const double pi = 3.14159265, taud = 0.25;
    const double a_lm = (sorder > 1 && im != 1) ? 0.5 : 1.0;
    double tauw = 1.0;
    if (sponge == 2) {
        const double H = zmax - zdamp;
        for (int k = 0; k < nz; ++k)
            if (z0[k] >= zdamp) {
                const double c = std::cos(0.5 * pi * (zmax - z0[k]) / H);
                tauw = taud * std::max(std::min(std::pow(c, 2.0), 1.0), 0.0);
            }
        tauw = 1.0 / (1.0 + a_lm * dt * tauw);
    }
    return tauw;
}

hipError_t launch_integ_scal(long n, int sorder, int im, double dt, const double *xold, const double *a,
                             double *tend, double *x, hipStream_t s)
{
    double a_lm = 1.0, b_lm = 0.0;
    if (sorder > 1 && im != 1) {
        a_lm = 0.5;
        b_lm = 1.0;
    }
    unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(integ_kernel, dim3(nb), dim3(256), 0, s, n, sorder, a_lm, b_lm, dt, xold, a, tend, x);
    return hipGetLastError();
}

hipError_t launch_totalav(const GridDev &g, const double *x, double *d_out, hipStream_t s)
{
    // d_out: [0] mean, [1] dzw-weighted sum, [2..] per-level partials (nz)
    double *part = d_out + 2;
    hipLaunchKernelGGL(level_sums_kernel, dim3(g.nz), dim3(256), 0, s, g.nx, g.ny, g.nz, x, part);
    long my = g.ny > 1 ? g.ny - 5 : 1;
    long ncells = (long)(g.nx - 5) * my * g.nz;
    hipLaunchKernelGGL(finish_sums_kernel, dim3(1), dim3(64), 0, s, g.nz, ncells, part, g.zm.dzw, d_out);
    return hipGetLastError();
}

namespace {

// out = in * p(k) (or in / p(k)) - shift, level profile p (funcpack.f90
// anelastic branches); in and out may alias
__global__ void level_scale_kernel(size_t plane, size_t n, const double *in, double *out, const double *prof,
                                   int divide, double shift)
{
    for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < n; id += (size_t)gridDim.x * blockDim.x) {
        const double p = prof[id / plane];
        out[id] = (divide ? in[id] / p : in[id] * p) - shift;
    }
}

// q = q * d (or q / d), d a 3D array (get_conserved / get_scalars)
__global__ void field_scale_kernel(size_t n, double *q, const double *d, int divide)
{
    for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < n; id += (size_t)gridDim.x * blockDim.x)
        q[id] = divide ? q[id] / d[id] : q[id] * d[id];
}

unsigned elementwise_blocks(size_t n)
{
    size_t b = (n + 255) / 256;
    return (unsigned)(b < 65536 ? b : 65536);
}

}  // namespace

hipError_t launch_level_scale(const GridDev &g, const double *prof, const double *in, double *out, int divide,
                              double shift, hipStream_t s)
{
    size_t plane = (size_t)g.nx * g.ny, n = plane * g.nz;
    hipLaunchKernelGGL(level_scale_kernel, dim3(elementwise_blocks(n)), dim3(256), 0, s, plane, n, in, out, prof,
                       divide, shift);
    return hipGetLastError();
}

hipError_t launch_field_scale(const GridDev &g, double *q, const double *d, int divide, hipStream_t s)
{
    size_t n = (size_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(field_scale_kernel, dim3(elementwise_blocks(n)), dim3(256), 0, s, n, q, d, divide);
    return hipGetLastError();
}

namespace {

// advection_ac_1d (advection.f90:993-1042) with grad_1d (gradients.f90:
// 400-473): xs(i,j,k) -= dref(k)*w0(k)*gz(k), gz(k) the derivative at level
// k of the quadratic through the values x/dref of levels k-1, k, k+1 (the
// floor reads x(1) at distance dz*fdz0(1); gz(nz) = 0).  One thread per
// interior column, marching up with a three-level register window.
__global__ __launch_bounds__(64) void adv1d_kernel(GridDev g, const double *__restrict__ w0,
                                                   const double *__restrict__ dprof, const double *__restrict__ d3,
                                                   const double *__restrict__ q, double *__restrict__ xs)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const bool three_d = ny > 1;
    const int i = 4 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = three_d ? 4 + (int)blockIdx.y : 1;
    if (i > nx - 2) return;
    const size_t plane = (size_t)nx * ny, col = (size_t)(i - 1) + (size_t)nx * (j - 1);
    const double *fdz0 = g.zm.fdz0;
    const double dz = g.zm.dz;
    auto dref = [&](int k) { return d3 ? d3[col + (size_t)(k - 1) * plane] : (dprof ? dprof[k - 1] : 1.0); };
    double dk = dref(1);
    double x0 = q[col] / dk, xm = 0.0;
    for (int k = 1; k <= nz; ++k) {
        double gz = 0.0, dn = 0.0, xp = 0.0;
        if (k <= nz - 1) {
            dn = dref(k + 1);
            xp = q[col + (size_t)k * plane] / dn;
            const double deltad = 0.5 * dz * (fdz0[k] + fdz0[k - 1]);
            const double deltau = k == 1 ? dz * fdz0[0] : 0.5 * dz * (fdz0[k - 1] + fdz0[k - 2]);
            const double deltac = (deltad + deltau) * (deltad * deltad + deltau * deltau) -
                                  (deltad - deltau) * (deltad * deltad - deltau * deltau);
            const double ffip1 = xp - x0;
            const double ffim1 = k == 1 ? 0.0 : x0 - xm;
            gz = ((ffip1 + ffim1) * (deltad * deltad + deltau * deltau) -
                  (ffip1 - ffim1) * (deltad * deltad - deltau * deltau)) / deltac;
        }
        const size_t id = col + (size_t)(k - 1) * plane;
        xs[id] = xs[id] - dk * w0[k - 1] * gz;
        xm = x0;
        x0 = xp;
        dk = dn;
    }
}

}  // namespace

hipError_t launch_advection_1d(const GridDev &g, const double *w0, const double *dprof, const double *d3,
                               const double *q, double *xs, hipStream_t s)
{
    const int mx = g.nx - 5, my = g.ny > 1 ? g.ny - 5 : 1;
    hipLaunchKernelGGL(adv1d_kernel, dim3((mx + 63) / 64, my, 1), dim3(64), 0, s, g, w0, dprof, d3, q, xs);
    return hipGetLastError();
}

}  // namespace mimadv
