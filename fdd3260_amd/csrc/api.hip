// api.hip -- the C ABI of libmimadv (declared in include/mimadv.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/mimadv.h"
#include "mimadv_internal.hpp"
#include "fused.hpp"

using namespace mimadv;

struct mimadv_ctx {
    GridDev g{};
    SchemeParams sp{};
    mimadv_opts opts{};
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    double *d_metrics = nullptr;  // fdz0, dzw, dz1w, dz1, ppmc
    double *d_red = nullptr;      // reduction scratch [2 + nz]
    double *d_budget = nullptr;   // budget results and per-block partials
    size_t budget_elems = 0;
    double *budget_scratch = nullptr;  // this device's static budget scratch
    // staging buffers for MIMADV_HOST_PTRS calls
    double *stage = nullptr;
    size_t stage_elems = 0;
    // advance scratch: two arrays per field
    double *work = nullptr;
    size_t work_elems = 0;
    // batched path: scratch bytes, tuning, captured step graph
    char *bwork = nullptr;
    size_t bwork_bytes = 0;
    mimadv_tuning tuning{0, 0, 0, 0};
    hipGraphExec_t graph_exec = nullptr;
    std::vector<const void *> graph_key;
    // adjoint: adjoint arrays, checkpoints and stage states
    double *adj = nullptr;
    size_t adj_elems = 0;
    // density weighting: den0, avden0 on the device (2*nz), or null
    double *d_den0 = nullptr;
    double *d_w0 = nullptr;  // subsidence profile of mimadv_advection_1d (nz)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    bool timing_pending = false;
    // MIMADV_PROFILE: an event pair around every batched launch of the last
    // profiled call, with its kind (PROF_MOMENTUM, PROF_F64, PROF_F32)
    std::vector<hipEvent_t> prof_ev;
    std::vector<int> prof_kind;
    size_t prof_used = 0;
    bool profiling = false;
    size_t n3() const { return (size_t)g.nx * g.ny * g.nz; }
};

namespace {

int fail(mimadv_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

int hipfail(mimadv_ctx *c, hipError_t e, const char *what)
{
    return fail(c, MIMADV_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(ctx, expr)                                   \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return hipfail(ctx, e_, #expr); \
    } while (0)

int check_opts(mimadv_ctx *c, const mimadv_opts *o)
{
    if (!o) return fail(c, MIMADV_ERR_ARG, "opts is NULL");
    if (o->scheme < MIMADV_WENO || o->scheme > MIMADV_LW) return fail(c, MIMADV_ERR_ARG, "unknown scheme");
    if (o->muscl_limiter < 0 || o->muscl_limiter > 3) return fail(c, MIMADV_ERR_ARG, "unknown MUSCL limiter");
    if (o->weno_bottom != 0 && o->weno_bottom != 1) return fail(c, MIMADV_ERR_ARG, "unknown weno_bottom convention");
    if (o->mom_ord != 2 && o->mom_ord != 4) return fail(c, MIMADV_ERR_ARG, "mom_ord must be 2 or 4");
    if (!(o->lim_tol > 0.0)) return fail(c, MIMADV_ERR_ARG, "lim_tol must be positive");
    return MIMADV_OK;
}

// The line kernel advection_{x,y,z} runs (advection.f90:580-620): limit =
// .false. gives unlimited QUICK for 'quick' and unlimited Lax-Wendroff for
// every other scheme; sorder = 1 gives limited Lax-Wendroff.  Kernel codes
// 6 and 7 are the unlimited QUICK and Lax-Wendroff.
int line_scheme(int scheme, unsigned flags)
{
    if (flags & MIMADV_UNLIMITED) return scheme == MIMADV_QUICK ? 6 : 7;
    if (flags & MIMADV_SORDER1) return MIMADV_LW;
    return scheme;
}

bool uses_dt0(int line) { return line == MIMADV_PPM || line == MIMADV_LW || line == 7; }

// The captured batch step bakes in the scheme, its options and the tile
// shapes: any change of those drops it, and the next MIMADV_GRAPH call
// captures again.
void drop_graph(mimadv_ctx *c)
{
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    c->graph_exec = nullptr;
    c->graph_key.clear();
}

void apply_opts(mimadv_ctx *c, const mimadv_opts *o)
{
    drop_graph(c);
    c->opts = *o;
    c->sp.scheme = o->scheme;
    c->sp.muscl_limiter = o->muscl_limiter;
    c->sp.ppm_limiter = o->ppm_limiter ? 1 : 0;
    c->sp.weno_bottom = o->weno_bottom;
    c->sp.lim_tol = o->lim_tol;
    c->sp.dt0 = 0.0;
}

// Device staging for host-pointer calls: returns a device buffer of n3()*count.
int ensure_stage(mimadv_ctx *c, size_t count)
{
    size_t need = c->n3() * count;
    if (need <= c->stage_elems) return MIMADV_OK;
    if (c->stage) (void)hipFree(c->stage);
    c->stage = nullptr;
    c->stage_elems = 0;
    hipError_t e = hipMalloc(&c->stage, need * sizeof(double));
    if (e != hipSuccess) return fail(c, MIMADV_ERR_NOMEM, "staging allocation failed");
    c->stage_elems = need;
    return MIMADV_OK;
}

int ensure_work(mimadv_ctx *c, size_t count)
{
    size_t need = c->n3() * count;
    if (need <= c->work_elems) return MIMADV_OK;
    if (c->work) (void)hipFree(c->work);
    c->work = nullptr;
    c->work_elems = 0;
    hipError_t e = hipMalloc(&c->work, need * sizeof(double));
    if (e != hipSuccess) return fail(c, MIMADV_ERR_NOMEM, "work allocation failed");
    c->work_elems = need;
    return MIMADV_OK;
}

// MIMADV_POSITIVE steps: Shu-Osher stages whose outgoing fluxes are scaled
// per donor cell (launch_positive_stage), two stage arrays per field plus
// the flux and theta scratch in the work buffer.
int advance_positive(mimadv_ctx *c, int integrator, int nsteps, double dt, const double *u, const double *v,
                     const double *w, int nf, double *const *q, const int *adv_form, unsigned flags)
{
    if (integrator != MIMADV_INT_SSPRK3 && integrator != MIMADV_INT_HEUN)
        return fail(c, MIMADV_ERR_UNSUPPORTED, "MIMADV_POSITIVE needs MIMADV_INT_SSPRK3 or MIMADV_INT_HEUN");
    if (c->sp.scheme != MIMADV_WENO || (flags & MIMADV_UNLIMITED))
        return fail(c, MIMADV_ERR_UNSUPPORTED, "MIMADV_POSITIVE implements the limited 'weno' scheme");
    if (c->g.bcl[4] != MIMADV_BC_NNF || c->g.bcl[5] != MIMADV_BC_NNF)
        return fail(c, MIMADV_ERR_UNSUPPORTED, "MIMADV_POSITIVE needs 'nnf' floor and lid");
    for (int f = 0; f < nf; ++f)
        if (adv_form && adv_form[f])
            return fail(c, MIMADV_ERR_UNSUPPORTED, "MIMADV_POSITIVE works on flux-form fields");
    int rc = ensure_work(c, (size_t)2 * nf + 4);
    if (rc) return rc;
    const size_t n3 = c->n3();
    for (int f = 0; f < nf; ++f) {
        HIPCHK(c, hipMemcpyAsync(c->work + (2 * f) * n3, q[f], n3 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->work + (2 * f + 1) * n3, q[f], n3 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    }
    double *scr = c->work + (size_t)2 * nf * n3;
    struct Stage {
        double c0, c1;
        int in, out;  // 0: q, 1: w1, 2: w2
    };
    std::vector<Stage> st;
    if (integrator == MIMADV_INT_HEUN) st = {{0.0, 1.0, 0, 1}, {0.5, 0.5, 1, 0}};
    else st = {{0.0, 1.0, 0, 1}, {0.75, 0.25, 1, 2}, {1.0 / 3.0, 2.0 / 3.0, 2, 0}};
    for (int s = 0; s < nsteps; ++s)
        for (const Stage &g : st)
            for (int f = 0; f < nf; ++f) {
                double *w1 = c->work + (2 * f) * n3, *w2 = c->work + (2 * f + 1) * n3;
                auto pick = [&](int which) { return which == 0 ? q[f] : (which == 1 ? w1 : w2); };
                PosArgs a{};
                a.q = pick(g.in);
                a.u = u;
                a.v = v ? v : u;
                a.w = w;
                a.q0 = q[f];
                a.qn = pick(g.out);
                a.c0 = g.c0;
                a.c1 = g.c1;
                a.dt = dt;
                a.dx1 = 1.0 / c->g.dx;
                a.dy1 = 1.0 / c->g.dy;
                a.fx = scr;
                a.fy = scr + n3;
                a.fz = scr + 2 * n3;
                a.th = scr + 3 * n3;
                HIPCHK(c, launch_positive_stage(c->g, c->sp, a, c->stream));
            }
    return MIMADV_OK;
}

// nsteps steps of the advection-only loop for nf fields (mimadv_advance
// without its checks and timing): per stage, tendency of every field,
// stage combine and periodic ghost refresh in one launch per field.
int advance_steps(mimadv_ctx *c, int integrator, int nsteps, double dt, const double *u, const double *v,
                  const double *w, int nf, double *const *q, const int *adv_form, unsigned flags = 0)
{
    if (flags & MIMADV_POSITIVE) return advance_positive(c, integrator, nsteps, dt, u, v, w, nf, q, adv_form, flags);
    int rc = ensure_work(c, (size_t)2 * nf);
    if (rc) return rc;
    size_t n3 = c->n3();
    SchemeParams sp = c->sp;
    sp.scheme = line_scheme(c->sp.scheme, flags | (integrator == MIMADV_INT_EULER ? MIMADV_SORDER1 : 0u));
    sp.dt0 = dt;
    // the scratch arrays start from the state so their ghosts are defined
    for (int f = 0; f < nf; ++f) {
        HIPCHK(c, hipMemcpyAsync(c->work + (2 * f) * n3, q[f], n3 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->work + (2 * f + 1) * n3, q[f], n3 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    }
    for (int s = 0; s < nsteps; ++s) {
        if (integrator == MIMADV_INT_EULER) {
            // sorder = 1 (integ.h): x = xold + dt*a, into w1, then back
            for (int f = 0; f < nf; ++f) {
                double *w1 = c->work + (2 * f) * n3, *w2 = c->work + (2 * f + 1) * n3;
                ScalarArgs a{};
                a.u = u;
                a.v = v;
                a.w = w;
                a.adv = adv_form ? adv_form[f] : 0;
                a.dirmask = 7;
                a.fill = 1;
                a.epi = EPI_HEUN;
                a.tend = w2;
                a.q0 = q[f];
                a.q = q[f];
                a.qn = w1;
                a.b_lm = 0.0;
                a.adt = dt;
                HIPCHK(c, launch_scalar_fp64(c->g, sp, a, c->stream));
                HIPCHK(c, hipMemcpyAsync(q[f], w1, n3 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            }
            continue;
        }
        int nst = integrator == MIMADV_INT_HEUN ? 2 : 3;
        for (int st = 1; st <= nst; ++st) {
            for (int f = 0; f < nf; ++f) {
                double *w1 = c->work + (2 * f) * n3, *w2 = c->work + (2 * f + 1) * n3;
                ScalarArgs a{};
                a.u = u;
                a.v = v;
                a.w = w;
                a.adv = adv_form ? adv_form[f] : 0;
                a.dirmask = 7;
                a.fill = 1;
                if (integrator == MIMADV_INT_HEUN) {
                    // w1: stage state x*, w2: accumulated tendency (integ.h)
                    a.epi = EPI_HEUN;
                    a.tend = w2;
                    a.q0 = q[f];
                    if (st == 1) {
                        a.q = q[f];
                        a.qn = w1;
                        a.b_lm = 0.0;
                        a.adt = 1.0 * dt;
                    } else {
                        a.q = w1;
                        a.qn = q[f];
                        a.b_lm = 1.0;
                        a.adt = 0.5 * dt;
                    }
                } else {
                    a.epi = EPI_SO;
                    a.dt = dt;
                    a.q0 = q[f];
                    if (st == 1) {
                        a.q = q[f]; a.qn = w1; a.c0 = 0.0; a.c1 = 1.0;
                    } else if (st == 2) {
                        a.q = w1; a.qn = w2; a.c0 = 0.75; a.c1 = 0.25;
                    } else {
                        a.q = w2; a.qn = q[f]; a.c0 = 1.0 / 3.0; a.c1 = 2.0 / 3.0;
                    }
                }
                HIPCHK(c, launch_scalar_fp64(c->g, sp, a, c->stream));
            }
        }
    }
    return MIMADV_OK;
}

}  // namespace

extern "C" {

int mimadv_create(const mimadv_grid *grid, const mimadv_opts *opts, mimadv_ctx **out)
{
    if (!out) return MIMADV_ERR_ARG;
    *out = nullptr;
    if (!grid) return MIMADV_ERR_ARG;
    mimadv_ctx *c = new (std::nothrow) mimadv_ctx();
    if (!c) return MIMADV_ERR_NOMEM;
    auto bad = [&](int code) {
        // keep the context so the caller can read the message, but mark it
        *out = c;
        return code;
    };
    if (grid->nx < 11) return bad(fail(c, MIMADV_ERR_ARG, "nx must be >= 11 (MAXX >= 6)"));
    if (grid->ny != 1 && grid->ny < 11) return bad(fail(c, MIMADV_ERR_ARG, "ny must be 1 or >= 11"));
    if (grid->nz < 8) return bad(fail(c, MIMADV_ERR_ARG, "nz must be >= 8"));
    if (!(grid->dx > 0) || !(grid->dy > 0) || !(grid->dz > 0)) return bad(fail(c, MIMADV_ERR_ARG, "spacings must be positive"));
    if (!grid->fdz0 || !grid->dzw || !grid->dz1w || !grid->dz1) return bad(fail(c, MIMADV_ERR_ARG, "metric arrays must be given"));
    for (int b = 0; b < 6; ++b)
        if (grid->bcl[b] < 0 || grid->bcl[b] > 2) return bad(fail(c, MIMADV_ERR_ARG, "unknown boundary code"));
    int rc = check_opts(c, opts);
    if (rc) return bad(rc);
    apply_opts(c, opts);

    hipError_t e = hipGetDevice(&c->device);
    if (e != hipSuccess) return bad(hipfail(c, e, "hipGetDevice"));
    c->g.nx = grid->nx;
    c->g.ny = grid->ny;
    c->g.nz = grid->nz;
    c->g.dx = grid->dx;
    c->g.dy = grid->dy;
    c->g.dz = grid->dz;
    for (int b = 0; b < 6; ++b) c->g.bcl[b] = grid->bcl[b];
    int nz = grid->nz;
    std::vector<double> ppmc = ppm_coefficients(nz, grid->dz1w);
    std::vector<float> momz = momentum_z_table(nz, grid->dzw, grid->dz1w, grid->dz1);
    size_t nmomz = (momz.size() + 1) / 2;  // in doubles
    size_t nmet = (size_t)4 * nz + ppmc.size() + nmomz;
    e = hipMalloc(&c->d_metrics, nmet * sizeof(double));
    if (e != hipSuccess) return bad(fail(c, MIMADV_ERR_NOMEM, "metric allocation failed"));
    std::vector<double> h(nmet);
    std::memcpy(h.data() + 0 * nz, grid->fdz0, nz * sizeof(double));
    std::memcpy(h.data() + 1 * nz, grid->dzw, nz * sizeof(double));
    std::memcpy(h.data() + 2 * nz, grid->dz1w, nz * sizeof(double));
    std::memcpy(h.data() + 3 * nz, grid->dz1, nz * sizeof(double));
    std::memcpy(h.data() + 4 * nz, ppmc.data(), ppmc.size() * sizeof(double));
    std::memcpy(h.data() + 4 * nz + ppmc.size(), momz.data(), momz.size() * sizeof(float));
    e = hipMemcpy(c->d_metrics, h.data(), nmet * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) return bad(hipfail(c, e, "metric upload"));
    c->g.zm.fdz0 = c->d_metrics;
    c->g.zm.dzw = c->d_metrics + nz;
    c->g.zm.dz1w = c->d_metrics + 2 * nz;
    c->g.zm.dz1 = c->d_metrics + 3 * nz;
    c->g.zm.ppmc = c->d_metrics + 4 * nz;
    c->g.zm.momz = (const float *)(c->d_metrics + 4 * nz + ppmc.size());
    c->g.zm.dz = grid->dz;
    e = budget_scratch_setup(c->device, &c->budget_scratch);
    if (e != hipSuccess) return bad(hipfail(c, e, "hipGetSymbolAddress of the budget scratch"));
    e = hipMalloc(&c->d_red, (size_t)(2 + nz) * sizeof(double));
    if (e != hipSuccess) return bad(fail(c, MIMADV_ERR_NOMEM, "reduction allocation failed"));
    e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) return bad(hipfail(c, e, "hipStreamCreate"));
    c->stream = c->own_stream;
    e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e != hipSuccess) return bad(hipfail(c, e, "hipEventCreate"));
    *out = c;
    return MIMADV_OK;
}

void mimadv_destroy(mimadv_ctx *c)
{
    if (!c) return;
    if (c->own_stream) (void)hipStreamSynchronize(c->own_stream);
    if (c->stream && c->stream != c->own_stream) (void)hipStreamSynchronize(c->stream);
    if (c->d_metrics) (void)hipFree(c->d_metrics);
    if (c->d_budget) (void)hipFree(c->d_budget);
    if (c->d_red) (void)hipFree(c->d_red);
    if (c->stage) (void)hipFree(c->stage);
    if (c->work) (void)hipFree(c->work);
    if (c->bwork) (void)hipFree(c->bwork);
    if (c->adj) (void)hipFree(c->adj);
    if (c->d_den0) (void)hipFree(c->d_den0);
    if (c->d_w0) (void)hipFree(c->d_w0);
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    for (hipEvent_t e : c->prof_ev) (void)hipEventDestroy(e);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
}

const char *mimadv_last_error(const mimadv_ctx *c)
{
    if (!c) return "NULL context";
    return c->err.c_str();
}

int mimadv_set_stream(mimadv_ctx *c, void *stream)
{
    if (!c) return MIMADV_ERR_ARG;
    c->stream = stream == MIMADV_OWN_STREAM ? c->own_stream : (hipStream_t)stream;
    return MIMADV_OK;
}

int mimadv_synchronize(mimadv_ctx *c)
{
    if (!c) return MIMADV_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MIMADV_OK;
}

int mimadv_set_opts(mimadv_ctx *c, const mimadv_opts *o)
{
    if (!c) return MIMADV_ERR_ARG;
    int rc = check_opts(c, o);
    if (rc) return rc;
    apply_opts(c, o);
    return MIMADV_OK;
}

int mimadv_scalar_tendency(mimadv_ctx *c, double dt0, const double *u, const double *v, const double *w, int nf,
                           const double *const *q, double *const *xs, const int *adv_form, unsigned flags)
{
    if (!c) return MIMADV_ERR_ARG;
    if (nf < 0) return fail(c, MIMADV_ERR_ARG, "nf < 0");
    if (nf == 0) return MIMADV_OK;
    if (!q || !xs) return fail(c, MIMADV_ERR_ARG, "q/xs arrays are NULL");
    unsigned dirs = flags & MIMADV_DIR_ALL;
    if (dirs == 0) return fail(c, MIMADV_ERR_ARG, "no direction selected");
    bool three_d = c->g.ny > 1;
    if ((dirs & MIMADV_DIR_X) && !u) return fail(c, MIMADV_ERR_ARG, "u is NULL");
    if ((dirs & MIMADV_DIR_Y) && three_d && !v) return fail(c, MIMADV_ERR_ARG, "v is NULL");
    if ((dirs & MIMADV_DIR_Z) && !w) return fail(c, MIMADV_ERR_ARG, "w is NULL");
    const int line = line_scheme(c->sp.scheme, flags);
    if (uses_dt0(line) && !(std::isfinite(dt0))) return fail(c, MIMADV_ERR_ARG, "PPM and Lax-Wendroff need a finite dt0");
    for (int f = 0; f < nf; ++f) {
        if (!q[f] || !xs[f]) return fail(c, MIMADV_ERR_ARG, "field pointer is NULL");
    }
    SchemeParams sp = c->sp;
    sp.scheme = line;
    sp.dt0 = dt0;
    size_t n3 = c->n3();
    size_t bytes = n3 * sizeof(double);
    bool host = (flags & MIMADV_HOST_PTRS) != 0;
    const double *du = u, *dv = v, *dw = w;
    if (host) {
        int rc = ensure_stage(c, 5);
        if (rc) return rc;
        double *su = c->stage, *sv = c->stage + n3, *sw = c->stage + 2 * n3;
        if (u) HIPCHK(c, hipMemcpyAsync(su, u, bytes, hipMemcpyHostToDevice, c->stream));
        if (v && three_d) HIPCHK(c, hipMemcpyAsync(sv, v, bytes, hipMemcpyHostToDevice, c->stream));
        if (w) HIPCHK(c, hipMemcpyAsync(sw, w, bytes, hipMemcpyHostToDevice, c->stream));
        du = su;
        dv = sv;
        dw = sw;
    }
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    float kms_total = 0.0f;
    for (int f = 0; f < nf; ++f) {
        ScalarArgs a{};
        a.u = du;
        a.v = dv;
        a.w = dw;
        a.adv = adv_form ? adv_form[f] : 0;
        a.epi = (flags & MIMADV_OVERWRITE) ? EPI_OVERWRITE : EPI_ACCUM;
        a.dirmask = (int)dirs;
        a.fill = 0;
        if (host) {
            double *sq = c->stage + 3 * n3, *sx = c->stage + 4 * n3;
            HIPCHK(c, hipMemcpyAsync(sq, q[f], bytes, hipMemcpyHostToDevice, c->stream));
            // xs is staged in even for MIMADV_OVERWRITE: its ghost columns are
            // not written by the kernel and must come back unchanged
            HIPCHK(c, hipMemcpyAsync(sx, xs[f], bytes, hipMemcpyHostToDevice, c->stream));
            a.q = sq;
            a.xs = sx;
            HIPCHK(c, hipEventRecord(c->ev0, c->stream));
            HIPCHK(c, launch_scalar_fp64(c->g, sp, a, c->stream));
            HIPCHK(c, hipEventRecord(c->ev1, c->stream));
            HIPCHK(c, hipMemcpyAsync(xs[f], sx, bytes, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            float ms = 0.0f;
            HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
            kms_total += ms;
        } else {
            a.q = q[f];
            a.xs = xs[f];
            HIPCHK(c, launch_scalar_fp64(c->g, sp, a, c->stream));
        }
    }
    if (host) {
        c->last_ms = kms_total;
        c->timing_pending = false;
    } else {
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
        c->timing_pending = true;
    }
    return MIMADV_OK;
}

int mimadv_momentum_tendency(mimadv_ctx *c, int ntime, const double *mu, const double *mv, const double *mw,
                             const double *ut, const double *vt, const double *wt, double *ua, double *va, double *wa,
                             unsigned flags)
{
    if (!c) return MIMADV_ERR_ARG;
    bool three_d = c->g.ny > 1;
    if (!mu || !mw || !ut || !wt || !ua || !wa) return fail(c, MIMADV_ERR_ARG, "NULL momentum/velocity array");
    if (three_d && (!mv || !vt || !va)) return fail(c, MIMADV_ERR_ARG, "NULL v array in 3D");
    size_t n3 = c->n3(), bytes = n3 * sizeof(double);
    if (flags & MIMADV_HOST_PTRS) {
        int rc = ensure_stage(c, 9);
        if (rc) return rc;
        double *s = c->stage;
        const double *hin[6] = {mu, mv, mw, ut, vt, wt};
        for (int a = 0; a < 6; ++a)
            if (hin[a] && (three_d || (a != 1 && a != 4)))
                HIPCHK(c, hipMemcpyAsync(s + a * n3, hin[a], bytes, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipEventRecord(c->ev0, c->stream));
        HIPCHK(c, launch_momentum_fp64(c->g, c->opts.mom_ord, c->opts.lavisc, ntime, s, s + n3, s + 2 * n3, s + 3 * n3,
                                       s + 4 * n3, s + 5 * n3, s + 6 * n3, s + 7 * n3, s + 8 * n3, c->stream));
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
        HIPCHK(c, hipMemcpyAsync(ua, s + 6 * n3, bytes, hipMemcpyDeviceToHost, c->stream));
        if (three_d) HIPCHK(c, hipMemcpyAsync(va, s + 7 * n3, bytes, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(wa, s + 8 * n3, bytes, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        float ms = 0.0f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->last_ms = ms;
        return MIMADV_OK;
    }
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, launch_momentum_fp64(c->g, c->opts.mom_ord, c->opts.lavisc, ntime, mu, mv, mw, ut, vt, wt, ua, va, wa,
                                   c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    c->timing_pending = true;
    return MIMADV_OK;
}

int mimadv_fill_halo(mimadv_ctx *c, int nf, double *const *x, unsigned flags)
{
    if (!c) return MIMADV_ERR_ARG;
    if (nf < 0 || (nf > 0 && !x)) return fail(c, MIMADV_ERR_ARG, "bad field list");
    size_t bytes = c->n3() * sizeof(double);
    for (int f = 0; f < nf; ++f) {
        if (!x[f]) return fail(c, MIMADV_ERR_ARG, "field pointer is NULL");
        if (flags & MIMADV_HOST_PTRS) {
            int rc = ensure_stage(c, 1);
            if (rc) return rc;
            HIPCHK(c, hipMemcpyAsync(c->stage, x[f], bytes, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, launch_fill_scalar(c->g, c->stage, c->stream));
            HIPCHK(c, hipMemcpyAsync(x[f], c->stage, bytes, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        } else {
            HIPCHK(c, launch_fill_scalar(c->g, x[f], c->stream));
        }
    }
    return MIMADV_OK;
}

int mimadv_fill_winds(mimadv_ctx *c, double *u, double *v, double *w, unsigned flags)
{
    if (!c) return MIMADV_ERR_ARG;
    const bool three_d = c->g.ny > 1;
    if (!u || !w || (three_d && !v)) return fail(c, MIMADV_ERR_ARG, "NULL wind component");
    if (flags & MIMADV_HOST_PTRS) {
        size_t n3 = c->n3(), bytes = n3 * sizeof(double);
        int rc = ensure_stage(c, 3);
        if (rc) return rc;
        double *host[3] = {u, three_d ? v : nullptr, w};
        for (int n = 0; n < 3; ++n)
            if (host[n]) HIPCHK(c, hipMemcpyAsync(c->stage + n * n3, host[n], bytes, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, launch_fill_winds(c->g, c->stage, c->stage + n3, c->stage + 2 * n3, c->stream));
        for (int n = 0; n < 3; ++n)
            if (host[n]) HIPCHK(c, hipMemcpyAsync(host[n], c->stage + n * n3, bytes, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return MIMADV_OK;
    }
    HIPCHK(c, launch_fill_winds(c->g, u, v, w, c->stream));
    return MIMADV_OK;
}

int mimadv_integ_mom(mimadv_ctx *c, int sorder, int im, double dt, const double *z0, double zmax, double zdamp,
                     int sponge, const double *const *old, const double *const *a, double *const *tend,
                     double *const *x, unsigned flags)
{
    if (!c) return MIMADV_ERR_ARG;
    if (!old || !a || !x || (sorder > 1 && !tend)) return fail(c, MIMADV_ERR_ARG, "bad component list");
    if (im < 1) return fail(c, MIMADV_ERR_ARG, "stage index im is 1-based");
    if (sponge == 2 && !z0) return fail(c, MIMADV_ERR_ARG, "the implicit sponge needs the level heights z0");
    if (sponge == 2 && !(zmax > zdamp)) return fail(c, MIMADV_ERR_ARG, "the implicit sponge needs zmax > zdamp");
    const bool three_d = c->g.ny > 1;
    for (int n = 0; n < 3; ++n) {
        if (n == 1 && !three_d) continue;
        if (!old[n] || !a[n] || !x[n] || (sorder > 1 && !tend[n]))
            return fail(c, MIMADV_ERR_ARG, "NULL wind component");
    }
    const double tw = sponge_factor(c->g.nz, sorder, im, dt, z0, zmax, zdamp, sponge);
    if (!(flags & MIMADV_HOST_PTRS)) {
        HIPCHK(c, launch_integ_mom(c->g, sorder, im, dt, tw, old, a, sorder > 1 ? tend : nullptr, x, c->stream));
        return MIMADV_OK;
    }
    // host arrays: one component at a time through four staging arrays
    size_t n3 = c->n3(), bytes = n3 * sizeof(double);
    int rc = ensure_stage(c, 4);
    if (rc) return rc;
    double *st = c->stage;
    for (int n = 0; n < 3; ++n) {
        if (n == 1 && !three_d) continue;
        HIPCHK(c, hipMemcpyAsync(st, old[n], bytes, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(st + n3, a[n], bytes, hipMemcpyHostToDevice, c->stream));
        if (sorder > 1) HIPCHK(c, hipMemcpyAsync(st + 2 * n3, tend[n], bytes, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, launch_integ_mom_one(c->g, n, sorder, im, dt, n == 2 ? tw : 1.0, st, st + n3,
                                        sorder > 1 ? st + 2 * n3 : nullptr, st + 3 * n3, c->stream));
        if (sorder > 1) HIPCHK(c, hipMemcpyAsync(tend[n], st + 2 * n3, bytes, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(x[n], st + 3 * n3, bytes, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return MIMADV_OK;
}

int mimadv_integ_scal(mimadv_ctx *c, int sorder, int im, double dt, int nf, const double *const *xold,
                      const double *const *a, double *const *tend, double *const *x, unsigned flags)
{
    if (!c) return MIMADV_ERR_ARG;
    if (nf < 0 || (nf > 0 && (!xold || !a || !x || (sorder > 1 && !tend))))
        return fail(c, MIMADV_ERR_ARG, "bad field list");
    if (im < 1) return fail(c, MIMADV_ERR_ARG, "stage index im is 1-based");
    size_t n3 = c->n3(), bytes = n3 * sizeof(double);
    for (int f = 0; f < nf; ++f) {
        if (flags & MIMADV_HOST_PTRS) {
            int rc = ensure_stage(c, 4);
            if (rc) return rc;
            double *s = c->stage;
            HIPCHK(c, hipMemcpyAsync(s, xold[f], bytes, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(s + n3, a[f], bytes, hipMemcpyHostToDevice, c->stream));
            if (sorder > 1) HIPCHK(c, hipMemcpyAsync(s + 2 * n3, tend[f], bytes, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, launch_integ_scal((long)n3, sorder, im, dt, s, s + n3, s + 2 * n3, s + 3 * n3, c->stream));
            if (sorder > 1) HIPCHK(c, hipMemcpyAsync(tend[f], s + 2 * n3, bytes, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipMemcpyAsync(x[f], s + 3 * n3, bytes, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        } else {
            HIPCHK(c, launch_integ_scal((long)n3, sorder, im, dt, xold[f], a[f], sorder > 1 ? tend[f] : nullptr, x[f],
                                        c->stream));
        }
    }
    return MIMADV_OK;
}

int mimadv_totalav(mimadv_ctx *c, const double *x, double *mean, double *dzw_sum, unsigned flags)
{
    if (!c) return MIMADV_ERR_ARG;
    if (!x) return fail(c, MIMADV_ERR_ARG, "x is NULL");
    const double *dx = x;
    if (flags & MIMADV_HOST_PTRS) {
        int rc = ensure_stage(c, 1);
        if (rc) return rc;
        HIPCHK(c, hipMemcpyAsync(c->stage, x, c->n3() * sizeof(double), hipMemcpyHostToDevice, c->stream));
        dx = c->stage;
    }
    HIPCHK(c, launch_totalav(c->g, dx, c->d_red, c->stream));
    double h[2];
    HIPCHK(c, hipMemcpyAsync(h, c->d_red, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (mean) *mean = h[0];
    if (dzw_sum) *dzw_sum = h[1];
    return MIMADV_OK;
}

int mimadv_advance(mimadv_ctx *c, int integrator, int nsteps, double dt, const double *u, const double *v,
                   const double *w, int nf, double *const *q, const int *adv_form, unsigned flags)
{
    if (!c) return MIMADV_ERR_ARG;
    if (flags & MIMADV_HOST_PTRS) return fail(c, MIMADV_ERR_UNSUPPORTED, "mimadv_advance works on device arrays");
    if (integrator != MIMADV_INT_EULER && integrator != MIMADV_INT_HEUN && integrator != MIMADV_INT_SSPRK3)
        return fail(c, MIMADV_ERR_UNSUPPORTED, "integrator must be MIMADV_INT_EULER, MIMADV_INT_HEUN or MIMADV_INT_SSPRK3");
    for (int b = 0; b < 4; ++b)
        if (c->g.bcl[b] != MIMADV_BC_PER && !(c->g.ny == 1 && b >= 2))
            return fail(c, MIMADV_ERR_UNSUPPORTED, "mimadv_advance needs periodic lateral boundaries");
    if (c->g.bcl[4] == MIMADV_BC_PER || c->g.bcl[5] == MIMADV_BC_PER)
        return fail(c, MIMADV_ERR_UNSUPPORTED, "mimadv_advance needs non-periodic vertical boundaries");
    if (nsteps < 0 || nf < 0) return fail(c, MIMADV_ERR_ARG, "negative count");
    if (nf == 0 || nsteps == 0) return MIMADV_OK;
    if (!u || !w || (c->g.ny > 1 && !v) || !q) return fail(c, MIMADV_ERR_ARG, "NULL array");
    for (int f = 0; f < nf; ++f)
        if (!q[f]) return fail(c, MIMADV_ERR_ARG, "field pointer is NULL");
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    int rc = advance_steps(c, integrator, nsteps, dt, u, v, w, nf, q, adv_form,
                           flags & (MIMADV_UNLIMITED | MIMADV_POSITIVE));
    if (rc) return rc;
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->timing_pending = false;
    return MIMADV_OK;
}

int mimadv_set_tuning(mimadv_ctx *c, const mimadv_tuning *t)
{
    if (!c || !t) return MIMADV_ERR_ARG;
    if (t->group32 < 0 || t->group32 > FUSED_MAX32 || t->kchunk < 0 || t->blocks_per_cu < 0)
        return fail(c, MIMADV_ERR_ARG, "tuning values out of range");
    if (t->tile_y != 0 && t->tile_y != 2 && t->tile_y != 4 && t->tile_y != 8)
        return fail(c, MIMADV_ERR_ARG, "tile_y must be 2, 4 or 8");
    if (t->tile_x != 0 && t->tile_x != 64 && !(t->tile_x == 32 && (t->tile_y == 0 || t->tile_y == 8)) &&
        !(t->tile_x == 16 && t->tile_y == 0))
        return fail(c, MIMADV_ERR_ARG, "tile_x must be 64, 32 with tile_y 8, or 16 with tile_y 0");
    drop_graph(c);
    c->tuning = *t;
    return MIMADV_OK;
}

}  // extern "C"

namespace {

struct BatchGroup {
    std::vector<int> f64, f32;  // field indices
};

// FP64 fields ride with the first group; FP32 fields are cut into groups of
// at most group32 (default: all of them, up to the launch limit).
int plan_groups(mimadv_ctx *c, int nf, const int *prec, std::vector<BatchGroup> &groups)
{
    std::vector<int> i64, i32;
    for (int f = 0; f < nf; ++f) {
        if (prec[f] == MIMADV_F64) i64.push_back(f);
        else if (prec[f] == MIMADV_F32) i32.push_back(f);
        else return fail(c, MIMADV_ERR_ARG, "precision must be MIMADV_F32 or MIMADV_F64");
    }
    if ((int)i64.size() > FUSED_MAX64) return fail(c, MIMADV_ERR_UNSUPPORTED, "at most 2 FP64 fields per batch");
    int g32 = c->tuning.group32 > 0 ? std::min(c->tuning.group32, FUSED_MAX32) : FUSED_MAX32;
    groups.clear();
    size_t p = 0;
    do {
        BatchGroup b;
        if (groups.empty()) b.f64 = i64;
        for (int n = 0; n < g32 && p < i32.size(); ++n) b.f32.push_back(i32[p++]);
        groups.push_back(b);
    } while (p < i32.size());
    return MIMADV_OK;
}

// The batched kernels run the limited scheme of the context on device
// arrays; the line-kernel selection flags of mimadv_scalar_tendency have no
// batched counterpart and are refused rather than ignored.
int check_batch_flags(mimadv_ctx *c, unsigned flags)
{
    if (flags & MIMADV_HOST_PTRS) return fail(c, MIMADV_ERR_UNSUPPORTED, "the batched path works on device arrays");
    if (flags & (MIMADV_UNLIMITED | MIMADV_SORDER1))
        return fail(c, MIMADV_ERR_UNSUPPORTED, "MIMADV_UNLIMITED and MIMADV_SORDER1 are not implemented on the batched path");
    return MIMADV_OK;
}

int check_batch(mimadv_ctx *c, const float *u, const float *v, const float *w)
{
    if (c->sp.scheme != MIMADV_WENO && c->sp.scheme != MIMADV_MUSCL && c->sp.scheme != MIMADV_WENO5)
        return fail(c, MIMADV_ERR_UNSUPPORTED, "the batched path implements the WENO, WENO5 and MUSCL schemes");
    for (int b = 0; b < 4; ++b)
        if (c->g.bcl[b] != MIMADV_BC_PER && !(c->g.ny == 1 && b >= 2))
            return fail(c, MIMADV_ERR_UNSUPPORTED, "the batched path needs periodic lateral boundaries");
    if (c->g.bcl[4] != MIMADV_BC_NNF || c->g.bcl[5] != MIMADV_BC_NNF)
        return fail(c, MIMADV_ERR_UNSUPPORTED, "the batched path needs no-flux ('nnf') floor and lid");
    if (!u || !w || (c->g.ny > 1 && !v)) return fail(c, MIMADV_ERR_ARG, "NULL velocity");
    return MIMADV_OK;
}

// MIMADV_MOMENTUM: the first three fields are FP32 momenta
int check_momentum(mimadv_ctx *c, unsigned flags, int nf, const int *precision)
{
    if (!(flags & MIMADV_MOMENTUM)) return MIMADV_OK;
    if (nf < 3) return fail(c, MIMADV_ERR_ARG, "MIMADV_MOMENTUM needs the three momenta as the first fields");
    for (int n = 0; n < 3; ++n)
        if (precision[n] != MIMADV_F32) return fail(c, MIMADV_ERR_ARG, "the batched momenta are MIMADV_F32");
    return MIMADV_OK;
}

FusedArgs base_args(mimadv_ctx *c, const float *u, const float *v, const float *w)
{
    FusedArgs A{};
    A.nx = c->g.nx;
    A.ny = c->g.ny;
    A.nz = c->g.nz;
    A.mx = c->g.nx - 5;
    A.my = c->g.ny - 5;
    A.three_d = c->g.ny > 1;
    A.dx1 = 1.0 / c->g.dx;
    A.dy1 = 1.0 / c->g.dy;
    A.dz = c->g.dz;
    A.dz1w = c->g.zm.dz1w;
    A.dzw = c->g.zm.dzw;
    A.dz1 = c->g.zm.dz1;
    A.fdz0 = c->g.zm.fdz0;
    A.u = u;
    A.v = v;
    A.w = w;
    A.muscl_limiter = c->sp.muscl_limiter;
    A.lim_tol = c->sp.lim_tol;
    A.lim_tol_inv = 1.0 / c->sp.lim_tol;
    A.weno_bottom = c->sp.weno_bottom;
    A.kchunk = c->tuning.kchunk;
    A.tile_y = c->tuning.tile_y;
    A.tile_x = c->tuning.tile_x;
    A.dirmask = 7;
    return A;
}

// One launch per group.  in/in0/out give, per field index, the stencil
// input, the step-start state and the output.
enum { PROF_MOMENTUM = 0, PROF_F64 = 1, PROF_F32 = 2, PROF_KINDS = 3 };

// Brackets one launch with an event pair when the call profiles.
struct ProfScope {
    mimadv_ctx *c;
    size_t slot = 0;
    hipError_t err = hipSuccess;
    ProfScope(mimadv_ctx *ctx, int kind) : c(ctx)
    {
        if (!c->profiling) return;
        slot = c->prof_used;
        while (c->prof_ev.size() < 2 * (slot + 1)) {
            hipEvent_t e = nullptr;
            err = hipEventCreate(&e);
            if (err != hipSuccess) return;
            c->prof_ev.push_back(e);
        }
        c->prof_kind.resize(slot + 1);
        c->prof_kind[slot] = kind;
        c->prof_used = slot + 1;
        err = hipEventRecord(c->prof_ev[2 * slot], c->stream);
    }
    hipError_t end()
    {
        if (!c->profiling || err != hipSuccess) return err;
        return hipEventRecord(c->prof_ev[2 * slot + 1], c->stream);
    }
};

hipError_t launch_groups(mimadv_ctx *c, const FusedArgs &base, const std::vector<BatchGroup> &groups,
                         const std::vector<const void *> &in, const std::vector<const void *> &in0,
                         const std::vector<void *> &out)
{
    for (const BatchGroup &b : groups) {
        FusedArgs A = base;
        A.n64_active = (int)b.f64.size();
        A.n32_active = (int)b.f32.size();
        for (size_t n = 0; n < b.f64.size(); ++n) {
            int f = b.f64[n];
            A.q64[n] = (const double *)in[f];
            A.q0_64[n] = (const double *)in0[f];
            A.out64[n] = (double *)out[f];
        }
        for (size_t n = 0; n < b.f32.size(); ++n) {
            int f = b.f32[n];
            A.q32[n] = (const float *)in[f];
            A.q0_32[n] = (const float *)in0[f];
            A.out32[n] = (float *)out[f];
        }
        // the FP64 and the FP32 members are separate launches
        if (A.n64_active > 0) {
            FusedArgs B = A;
            B.n32_active = 0;
            ProfScope p(c, PROF_F64);
            if (p.err != hipSuccess) return p.err;
            hipError_t e = launch_fused(c->sp.scheme, B.n64_active, 0, B, c->stream, c->tuning.blocks_per_cu);
            if (e == hipSuccess) e = p.end();
            if (e != hipSuccess) return e;
        }
        if (A.n32_active > 0) {
            FusedArgs B = A;
            B.n64_active = 0;
            ProfScope p(c, PROF_F32);
            if (p.err != hipSuccess) return p.err;
            hipError_t e = launch_fused(c->sp.scheme, 0, fused_padded_count(B.n32_active), B, c->stream,
                                        c->tuning.blocks_per_cu);
            if (e == hipSuccess) e = p.end();
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

// Momentum launch of a batched stage, bracketed when profiling.
hipError_t launch_momentum_prof(mimadv_ctx *c, const MomBatchArgs &M)
{
    ProfScope p(c, PROF_MOMENTUM);
    if (p.err != hipSuccess) return p.err;
    hipError_t e = launch_momentum_batch(c->g, c->opts.mom_ord, c->opts.lavisc, M, c->stream);
    return e == hipSuccess ? p.end() : e;
}

// Start of a batched call: MIMADV_PROFILE turns the per-launch events on
// (eager launches only).
int begin_profile(mimadv_ctx *c, unsigned flags)
{
    c->profiling = false;
    if (!(flags & MIMADV_PROFILE)) return MIMADV_OK;
    if (flags & MIMADV_GRAPH) return fail(c, MIMADV_ERR_ARG, "MIMADV_PROFILE times eager launches: drop MIMADV_GRAPH");
    c->profiling = true;
    c->prof_used = 0;
    c->prof_kind.clear();
    return MIMADV_OK;
}

}  // namespace

extern "C" {

int mimadv_batch_tendency(mimadv_ctx *c, const float *u, const float *v, const float *w, int nf,
                          const int *precision, const void *const *q, void *const *xs, unsigned flags)
{
    if (!c) return MIMADV_ERR_ARG;
    int rc = check_batch_flags(c, flags);
    if (rc) return rc;
    rc = check_batch(c, u, v, w);
    if (rc) return rc;
    if (nf <= 0) return nf == 0 ? MIMADV_OK : fail(c, MIMADV_ERR_ARG, "nf < 0");
    if (!precision || !q || !xs) return fail(c, MIMADV_ERR_ARG, "NULL field list");
    const int nm = (flags & MIMADV_MOMENTUM) ? 3 : 0;
    rc = check_momentum(c, flags, nf, precision);
    if (rc) return rc;
    rc = begin_profile(c, flags);
    if (rc) return rc;
    if (nm && (flags & MIMADV_DIR_ALL) && (flags & MIMADV_DIR_ALL) != MIMADV_DIR_ALL)
        return fail(c, MIMADV_ERR_UNSUPPORTED, "the batched momentum tendency covers all directions");
    std::vector<BatchGroup> groups;
    rc = plan_groups(c, nf - nm, precision + nm, groups);
    if (rc) return rc;
    std::vector<const void *> in(nf - nm), in0(nf - nm);
    std::vector<void *> out(nf - nm);
    for (int f = 0; f < nf; ++f) {
        if (!q[f] || !xs[f]) return fail(c, MIMADV_ERR_ARG, "field pointer is NULL");
        if (f < nm) continue;
        in[f - nm] = q[f];
        in0[f - nm] = q[f];
        out[f - nm] = xs[f];
    }
    FusedArgs A = base_args(c, u, v, w);
    A.epi = 0;
    A.fill = 0;
    A.dirmask = (flags & MIMADV_DIR_ALL) ? (int)(flags & MIMADV_DIR_ALL) : 7;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    if (nf > nm) HIPCHK(c, launch_groups(c, A, groups, in, in0, out));
    if (nm) {
        MomBatchArgs M{};
        for (int n = 0; n < 3; ++n) {
            M.m[n] = (const float *)q[n];
            M.m0[n] = (const float *)q[n];
            M.out[n] = (float *)xs[n];
        }
        M.ut = u;
        M.vt = v ? v : u;
        M.wt = w;
        M.epi = 0;
        HIPCHK(c, launch_momentum_prof(c, M));
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    c->timing_pending = true;
    return MIMADV_OK;
}

int mimadv_batch_advance(mimadv_ctx *c, int integrator, int nsteps, double dt, const float *u, const float *v,
                         const float *w, int nf, const int *precision, void *const *q, unsigned flags)
{
    if (!c) return MIMADV_ERR_ARG;
    int rc = check_batch_flags(c, flags);
    if (rc) return rc;
    rc = check_batch(c, u, v, w);
    if (rc) return rc;
    if (integrator != MIMADV_INT_HEUN && integrator != MIMADV_INT_SSPRK3)
        return fail(c, MIMADV_ERR_UNSUPPORTED, "batched integrators: MIMADV_INT_HEUN, MIMADV_INT_SSPRK3");
    if (nsteps < 0 || nf < 0) return fail(c, MIMADV_ERR_ARG, "negative count");
    if (nsteps == 0 || nf == 0) return MIMADV_OK;
    if (!precision || !q) return fail(c, MIMADV_ERR_ARG, "NULL field list");
    const int nm = (flags & MIMADV_MOMENTUM) ? 3 : 0;
    rc = check_momentum(c, flags, nf, precision);
    if (rc) return rc;
    rc = begin_profile(c, flags);
    if (rc) return rc;
    std::vector<BatchGroup> groups;
    rc = plan_groups(c, nf - nm, precision + nm, groups);
    if (rc) return rc;
    size_t n3 = c->n3();
    // two stage buffers per field, in the field's precision
    std::vector<size_t> off(nf);
    size_t bytes = 0;
    for (int f = 0; f < nf; ++f) {
        if (!q[f]) return fail(c, MIMADV_ERR_ARG, "field pointer is NULL");
        size_t es = precision[f] == MIMADV_F64 ? 8 : 4;
        bytes = (bytes + 255) & ~(size_t)255;
        off[f] = bytes;
        bytes += 2 * n3 * es;
    }
    if (bytes > c->bwork_bytes) {
        if (c->bwork) (void)hipFree(c->bwork);
        c->bwork = nullptr;
        c->bwork_bytes = 0;
        drop_graph(c);
        if (hipMalloc(&c->bwork, bytes) != hipSuccess) return fail(c, MIMADV_ERR_NOMEM, "batch scratch allocation failed");
        c->bwork_bytes = bytes;
    }
    std::vector<void *> s1(nf), s2(nf);
    for (int f = 0; f < nf; ++f) {
        size_t es = precision[f] == MIMADV_F64 ? 8 : 4;
        s1[f] = c->bwork + off[f];
        s2[f] = c->bwork + off[f] + n3 * es;
        // stage buffers start as copies so their ghost columns are defined
        HIPCHK(c, hipMemcpyAsync(s1[f], q[f], n3 * es, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(s2[f], q[f], n3 * es, hipMemcpyDeviceToDevice, c->stream));
    }
    FusedArgs A = base_args(c, u, v, w);
    A.epi = 1;
    A.fill = 1;
    A.dt = dt;
    struct Stage {
        double c0, c1;
        int in, out;  // 0: q, 1: s1, 2: s2
    };
    std::vector<Stage> st;
    if (integrator == MIMADV_INT_HEUN) st = {{0.0, 1.0, 0, 1}, {0.5, 0.5, 1, 0}};
    else st = {{0.0, 1.0, 0, 1}, {0.75, 0.25, 1, 2}, {1.0 / 3.0, 2.0 / 3.0, 2, 0}};
    auto pick = [&](int which, int f) -> void * { return which == 0 ? q[f] : (which == 1 ? s1[f] : s2[f]); };
    auto one_step = [&]() -> hipError_t {
        for (const Stage &s : st) {
            FusedArgs B = A;
            B.c0 = s.c0;
            B.c1 = s.c1;
            std::vector<const void *> in(nf - nm), in0(nf - nm);
            std::vector<void *> out(nf - nm);
            for (int f = nm; f < nf; ++f) {
                in[f - nm] = pick(s.in, f);
                in0[f - nm] = q[f];
                out[f - nm] = pick(s.out, f);
            }
            if (nf > nm) {
                hipError_t e = launch_groups(c, B, groups, in, in0, out);
                if (e != hipSuccess) return e;
            }
            if (nm) {
                MomBatchArgs M{};
                for (int n = 0; n < 3; ++n) {
                    M.m[n] = (const float *)pick(s.in, n);
                    M.m0[n] = (const float *)q[n];
                    M.out[n] = (float *)pick(s.out, n);
                }
                M.ut = u;
                M.vt = v ? v : u;
                M.wt = w;
                M.c0 = s.c0;
                M.c1 = s.c1;
                M.dt = dt;
                M.epi = 1;
                M.fill = 1;
                hipError_t e = launch_momentum_prof(c, M);
                if (e != hipSuccess) return e;
            }
        }
        return hipSuccess;
    };
    bool use_graph = (flags & MIMADV_GRAPH) != 0;
    if (use_graph) {
        // the captured step depends on every pointer and scalar it bakes in
        std::vector<const void *> key = {u, v, w, (const void *)c->bwork};
        for (int f = 0; f < nf; ++f) key.push_back(q[f]);
        key.push_back((const void *)(size_t)integrator);
        key.push_back((const void *)(size_t)nf);
        key.push_back((const void *)(size_t)nm);
        for (int f = 0; f < nf; ++f) key.push_back((const void *)(size_t)precision[f]);
        double dtk = dt;
        size_t dbits;
        std::memcpy(&dbits, &dtk, sizeof dbits);
        key.push_back((const void *)dbits);
        if (!c->graph_exec || key != c->graph_key) {
            drop_graph(c);
            hipGraph_t graph = nullptr;
            // capture on the context's own stream (the caller's stream may
            // be the null stream, which cannot capture); the graph is then
            // launched on the caller's stream
            hipStream_t user = c->stream;
            c->stream = c->own_stream;
            hipError_t be = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
            if (be != hipSuccess) {
                c->stream = user;
                return hipfail(c, be, "hipStreamBeginCapture");
            }
            hipError_t le = one_step();
            hipError_t ce = hipStreamEndCapture(c->stream, &graph);
            c->stream = user;
            if (le != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                return hipfail(c, le, "launch during capture");
            }
            if (ce != hipSuccess) return hipfail(c, ce, "hipStreamEndCapture");
            hipError_t ie = hipGraphInstantiate(&c->graph_exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) return hipfail(c, ie, "hipGraphInstantiate");
            c->graph_key = key;
        }
    }
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    for (int s = 0; s < nsteps; ++s) {
        if (use_graph) HIPCHK(c, hipGraphLaunch(c->graph_exec, c->stream));
        else HIPCHK(c, one_step());
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    c->timing_pending = true;
    return MIMADV_OK;
}

int mimadv_batch_budget(mimadv_ctx *c, int nf, const int *precision, const void *const *q, double *out)
{
    if (!c) return MIMADV_ERR_ARG;
    if (nf < 0 || (nf > 0 && (!precision || !q || !out))) return fail(c, MIMADV_ERR_ARG, "bad field list");
    if (nf == 0) return MIMADV_OK;
    for (int f = 0; f < nf; ++f) {
        if (!q[f]) return fail(c, MIMADV_ERR_ARG, "field pointer is NULL");
        if (precision[f] != MIMADV_F32 && precision[f] != MIMADV_F64)
            return fail(c, MIMADV_ERR_ARG, "precision must be MIMADV_F32 or MIMADV_F64");
    }
    BudgetArgs a{};
    a.nx = c->g.nx;
    a.ny = c->g.ny;
    a.nz = c->g.nz;
    a.nb = budget_blocks(c->g);
    a.dzw = c->g.zm.dzw;
    const size_t need = (size_t)std::min(nf, BUDGET_MAXF) * (a.nb + 1);
    if (need > c->budget_elems) {
        if (c->d_budget) (void)hipFree(c->d_budget);
        c->d_budget = nullptr;
        c->budget_elems = 0;
        if (hipMalloc(&c->d_budget, need * sizeof(double)) != hipSuccess)
            return fail(c, MIMADV_ERR_NOMEM, "budget scratch allocation failed");
        c->budget_elems = need;
    }
    // one pass over up to BUDGET_MAXF fields, one copy back, one wait
    for (int f0 = 0; f0 < nf; f0 += BUDGET_MAXF) {
        const int n = std::min(BUDGET_MAXF, nf - f0);
        for (int f = 0; f < n; ++f) {
            a.q[f] = q[f0 + f];
            a.fp32[f] = precision[f0 + f] == MIMADV_F32;
        }
        a.part = c->d_budget + n;
        HIPCHK(c, launch_budgets(a, n, c->d_budget, c->stream));
        HIPCHK(c, hipMemcpyAsync(out + f0, c->d_budget, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return MIMADV_OK;
}

int mimadv_batch_kernel_ms(mimadv_ctx *c, double *ms, int *count)
{
    if (!c || !ms || !count) return MIMADV_ERR_ARG;
    for (int k = 0; k < PROF_KINDS; ++k) {
        ms[k] = 0.0;
        count[k] = 0;
    }
    if (c->prof_used == 0) return MIMADV_OK;
    HIPCHK(c, hipEventSynchronize(c->prof_ev[2 * c->prof_used - 1]));
    for (size_t s = 0; s < c->prof_used; ++s) {
        float t = 0.0f;
        HIPCHK(c, hipEventElapsedTime(&t, c->prof_ev[2 * s], c->prof_ev[2 * s + 1]));
        ms[c->prof_kind[s]] += t;
        count[c->prof_kind[s]] += 1;
    }
    return MIMADV_OK;
}

int mimadv_batch_budget_device(mimadv_ctx *c, int nf, const int *precision, const void *const *q, double *d_out)
{
    if (!c) return MIMADV_ERR_ARG;
    if (nf < 0 || (nf > 0 && (!precision || !q || !d_out))) return fail(c, MIMADV_ERR_ARG, "bad field list");
    if (nf > BUDGET_MAXF)
        return fail(c, MIMADV_ERR_ARG, "mimadv_batch_budget_device takes at most 32 fields per call (its static "
                                       "scratch holds the partials of 32 fields)");
    if (nf == 0) return MIMADV_OK;
    if (!c->budget_scratch) return fail(c, MIMADV_ERR_HIP, "budget scratch address unknown on this device");
    BudgetArgs a{};
    a.nx = c->g.nx;
    a.ny = c->g.ny;
    a.nz = c->g.nz;
    a.nb = budget_blocks(c->g);
    a.dzw = c->g.zm.dzw;
    for (int f = 0; f < nf; ++f) {
        if (!q[f]) return fail(c, MIMADV_ERR_ARG, "field pointer is NULL");
        if (precision[f] != MIMADV_F32 && precision[f] != MIMADV_F64)
            return fail(c, MIMADV_ERR_ARG, "precision must be MIMADV_F32 or MIMADV_F64");
        a.q[f] = q[f];
        a.fp32[f] = precision[f] == MIMADV_F32;
    }
    // every partial the call reads is written by the same call first:
    // nothing in the shared scratch needs clearing
    a.part = c->budget_scratch;
    HIPCHK(c, launch_budgets(a, nf, d_out, c->stream));
    return MIMADV_OK;
}

int mimadv_last_kernel_ms(mimadv_ctx *c, double *ms)
{
    if (!c || !ms) return MIMADV_ERR_ARG;
    if (c->timing_pending) {
        HIPCHK(c, hipEventSynchronize(c->ev1));
        float t = 0.0f;
        HIPCHK(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
        c->last_ms = t;
        c->timing_pending = false;
    }
    *ms = c->last_ms;
    return MIMADV_OK;
}

}  // extern "C"

namespace {

int check_adjoint(mimadv_ctx *c, const double *u, const double *v, const double *w)
{
    if (c->sp.scheme != MIMADV_WENO && c->sp.scheme != MIMADV_WENO5)
        return fail(c, MIMADV_ERR_UNSUPPORTED, "the adjoint implements the WENO and WENO5 schemes");
    if (c->sp.weno_bottom != MIMADV_WENO_BOTTOM_X0_IS_X1)
        return fail(c, MIMADV_ERR_UNSUPPORTED, "the adjoint implements the x(0) = x(1) floor convention");
    for (int b = 0; b < 4; ++b)
        if (c->g.bcl[b] != MIMADV_BC_PER && !(c->g.ny == 1 && b >= 2))
            return fail(c, MIMADV_ERR_UNSUPPORTED, "the adjoint needs periodic lateral boundaries");
    if (c->g.bcl[4] != MIMADV_BC_NNF || c->g.bcl[5] != MIMADV_BC_NNF)
        return fail(c, MIMADV_ERR_UNSUPPORTED, "the adjoint needs no-flux ('nnf') floor and lid");
    if (!u || !w || (c->g.ny > 1 && !v)) return fail(c, MIMADV_ERR_ARG, "NULL velocity");
    return MIMADV_OK;
}

AdjArgs adj_base(mimadv_ctx *c, const double *u, const double *v, const double *w, double *ub, double *vb,
                 double *wb)
{
    AdjArgs A{};
    A.nx = c->g.nx;
    A.ny = c->g.ny;
    A.nz = c->g.nz;
    A.dx1 = 1.0 / c->g.dx;
    A.dy1 = 1.0 / c->g.dy;
    A.dz1w = c->g.zm.dz1w;
    A.u = u;
    A.v = c->g.ny > 1 ? v : u;
    A.w = w;
    A.ub = ub;
    A.vb = c->g.ny > 1 ? vb : ub;
    A.wb = wb;
    A.kchunk = c->tuning.kchunk;
    A.weno5 = c->sp.scheme == MIMADV_WENO5 ? 1 : 0;
    return A;
}

}  // namespace

extern "C" {

int mimadv_adjoint_tendency(mimadv_ctx *c, const double *u, const double *v, const double *w, const double *q,
                            const double *lam, double *qbar, double *ubar, double *vbar, double *wbar)
{
    if (!c) return MIMADV_ERR_ARG;
    int rc = check_adjoint(c, u, v, w);
    if (rc) return rc;
    if (!q || !lam || !qbar || !ubar || !wbar || (c->g.ny > 1 && !vbar))
        return fail(c, MIMADV_ERR_ARG, "NULL array");
    if (qbar == lam || qbar == q) return fail(c, MIMADV_ERR_ARG, "qbar must not alias lam or q");
    AdjArgs A = adj_base(c, u, v, w, ubar, vbar, wbar);
    A.q = q;
    A.qbn = lam;
    A.out = qbar;
    A.ic = 0.0;
    A.ls = 1.0;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, launch_adjoint_stage(A, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    c->timing_pending = true;
    return MIMADV_OK;
}

}  // extern "C"

namespace {

// Backward sweep of nsteps steps with checkpoints every K steps: the state
// at step g*K comes from ext[g] when the caller recorded the forward, else
// from a first forward sweep from q0 into internal storage.
int adjoint_core(mimadv_ctx *c, int integrator, int nsteps, double dt, const double *u, const double *v,
                 const double *w, const double *q0, const double *const *ext, int K, double *qbar, double *ubar,
                 double *vbar, double *wbar)
{
    const size_t n3 = c->n3();
    const int S = integrator == MIMADV_INT_HEUN ? 2 : 3;
    const int nseg = nsteps > 0 ? (nsteps + K - 1) / K : 0;
    const int own_ck = ext ? 0 : nseg;
    // layout: P, T, ACC (adjoint arrays), own checkpoints, S*K stage states
    const size_t need = n3 * ((size_t)3 + own_ck + (size_t)S * K);
    if (need > c->adj_elems) {
        if (c->adj) (void)hipFree(c->adj);
        c->adj = nullptr;
        c->adj_elems = 0;
        if (hipMalloc(&c->adj, need * sizeof(double)) != hipSuccess)
            return fail(c, MIMADV_ERR_NOMEM, "adjoint trajectory allocation failed (reduce checkpoint)");
        c->adj_elems = need;
    }
    double *P = c->adj, *T = c->adj + n3, *ACC = c->adj + 2 * n3;
    double *ck = c->adj + 3 * n3, *st = ck + (size_t)own_ck * n3;
    auto ckpt = [&](int g) -> const double * { return ext ? ext[g] : ck + (size_t)g * n3; };
    auto X = [&](int n, int s) { return st + ((size_t)n * S + s) * n3; };  // stage s input of segment step n
    SchemeParams sp = c->sp;
    sp.dt0 = dt;

    // one step of the forward, exactly as mimadv_advance takes it, from X(n,0)
    auto fwd_step = [&](int n, double *qout) -> hipError_t {
        for (int s = 1; s <= S; ++s) {
            ScalarArgs a{};
            a.u = u;
            a.v = v;
            a.w = w;
            a.adv = 0;
            a.dirmask = 7;
            a.fill = 1;
            a.q0 = X(n, 0);
            a.q = X(n, s - 1);
            a.qn = s < S ? X(n, s) : qout;
            if (integrator == MIMADV_INT_HEUN) {
                a.epi = EPI_HEUN;
                a.tend = T;
                a.b_lm = s == 1 ? 0.0 : 1.0;
                a.adt = s == 1 ? 1.0 * dt : 0.5 * dt;
            } else {
                a.epi = EPI_SO;
                a.dt = dt;
                a.c0 = s == 1 ? 0.0 : (s == 2 ? 0.75 : 1.0 / 3.0);
                a.c1 = s == 1 ? 1.0 : (s == 2 ? 0.25 : 2.0 / 3.0);
            }
            hipError_t e = launch_scalar_fp64(c->g, sp, a, c->stream);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    };
    // replay segment g from its checkpoint; the state after its last step goes to qlast
    auto replay = [&](int g, double *qlast) -> hipError_t {
        int len = std::min(K, nsteps - g * K);
        hipError_t e = hipMemcpyAsync(X(0, 0), ckpt(g), n3 * sizeof(double), hipMemcpyDeviceToDevice, c->stream);
        for (int n = 0; n < len && e == hipSuccess; ++n) e = fwd_step(n, n + 1 < len ? X(n + 1, 0) : qlast);
        return e;
    };
    // Shu-Osher coefficients of the stages (Heun as its two-stage form)
    const double c0s[3] = {0.0, integrator == MIMADV_INT_HEUN ? 0.5 : 0.75, 1.0 / 3.0};
    const double c1s[3] = {1.0, integrator == MIMADV_INT_HEUN ? 0.5 : 0.25, 2.0 / 3.0};

    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, hipMemsetAsync(c->adj, 0, 3 * n3 * sizeof(double), c->stream));
    HIPCHK(c, hipMemcpyAsync(P, qbar, n3 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    if (!ext && nseg > 0) {
        HIPCHK(c, hipMemcpyAsync(ck, q0, n3 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        for (int g = 0; g + 1 < nseg; ++g) HIPCHK(c, replay(g, ck + (size_t)(g + 1) * n3));
    }
    AdjArgs base = adj_base(c, u, v, w, ubar, vbar, wbar);
    for (int g = nseg - 1; g >= 0; --g) {
        HIPCHK(c, replay(g, ACC));
        int len = std::min(K, nsteps - g * K);
        for (int n = len - 1; n >= 0; --n) {
            // P = adjoint of the step output; walk the stages backwards
            double *in = P, *tmp = T;
            for (int s = S - 1; s >= 0; --s) {
                AdjArgs A = base;
                A.q = X(n, s);
                A.qbn = in;
                A.ic = c1s[s];
                A.ls = c1s[s] * dt;
                A.c0 = c0s[s];
                if (s > 0) {
                    A.out = tmp;
                    A.accumulate = 0;
                    A.acc0 = ACC;
                    A.acc0_mode = s == S - 1 ? 1 : 2;
                } else {
                    A.out = ACC;  // the first stage reads the step input itself (c0 = 0)
                    A.accumulate = 1;
                    A.acc0_mode = 0;
                }
                HIPCHK(c, launch_adjoint_stage(A, c->stream));
                if (s > 0) std::swap(in, tmp);
            }
            std::swap(P, ACC);  // P: adjoint of the step input
        }
    }
    HIPCHK(c, launch_interior_copy(c->g, P, qbar, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    c->timing_pending = true;
    return MIMADV_OK;
}

int check_adjoint_call(mimadv_ctx *c, int integrator, int nsteps, const double *u, const double *v, const double *w,
                       double *qbar, double *ubar, double *vbar, double *wbar)
{
    int rc = check_adjoint(c, u, v, w);
    if (rc) return rc;
    if (integrator != MIMADV_INT_HEUN && integrator != MIMADV_INT_SSPRK3)
        return fail(c, MIMADV_ERR_UNSUPPORTED, "adjoint integrators: MIMADV_INT_HEUN, MIMADV_INT_SSPRK3");
    if (nsteps < 0) return fail(c, MIMADV_ERR_ARG, "negative step count");
    if (!qbar || !ubar || !wbar || (c->g.ny > 1 && !vbar)) return fail(c, MIMADV_ERR_ARG, "NULL array");
    return MIMADV_OK;
}

int checkpoint_interval(int checkpoint, int nsteps)
{
    return (checkpoint <= 0 || checkpoint >= nsteps) ? (nsteps > 0 ? nsteps : 1) : checkpoint;
}

}  // namespace

extern "C" {

int mimadv_adjoint_advance(mimadv_ctx *c, int integrator, int nsteps, double dt, const double *u, const double *v,
                           const double *w, const double *q0, double *qbar, double *ubar, double *vbar,
                           double *wbar, int checkpoint)
{
    if (!c) return MIMADV_ERR_ARG;
    int rc = check_adjoint_call(c, integrator, nsteps, u, v, w, qbar, ubar, vbar, wbar);
    if (rc) return rc;
    if (!q0) return fail(c, MIMADV_ERR_ARG, "q0 is NULL");
    return adjoint_core(c, integrator, nsteps, dt, u, v, w, q0, nullptr, checkpoint_interval(checkpoint, nsteps), qbar,
                        ubar, vbar, wbar);
}

int mimadv_advance_record(mimadv_ctx *c, int integrator, int nsteps, double dt, const double *u, const double *v,
                          const double *w, double *q, int checkpoint, double *const *ck)
{
    if (!c) return MIMADV_ERR_ARG;
    int rc = check_adjoint(c, u, v, w);
    if (rc) return rc;
    if (integrator != MIMADV_INT_HEUN && integrator != MIMADV_INT_SSPRK3)
        return fail(c, MIMADV_ERR_UNSUPPORTED, "integrators: MIMADV_INT_HEUN, MIMADV_INT_SSPRK3");
    if (nsteps < 0 || !q) return fail(c, MIMADV_ERR_ARG, "bad step count or NULL q");
    const int K = checkpoint_interval(checkpoint, nsteps);
    const int nseg = nsteps > 0 ? (nsteps + K - 1) / K : 0;
    if (nseg > 0 && !ck) return fail(c, MIMADV_ERR_ARG, "NULL checkpoint list");
    for (int g = 0; g < nseg; ++g)
        if (!ck[g]) return fail(c, MIMADV_ERR_ARG, "NULL checkpoint array");
    const size_t bytes = c->n3() * sizeof(double);
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    for (int g = 0; g < nseg; ++g) {
        HIPCHK(c, hipMemcpyAsync(ck[g], q, bytes, hipMemcpyDeviceToDevice, c->stream));
        rc = advance_steps(c, integrator, std::min(K, nsteps - g * K), dt, u, v, w, 1, &q, nullptr);
        if (rc) return rc;
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    c->timing_pending = true;
    return MIMADV_OK;
}

int mimadv_adjoint_advance_recorded(mimadv_ctx *c, int integrator, int nsteps, double dt, const double *u,
                                    const double *v, const double *w, int checkpoint, const double *const *ck,
                                    double *qbar, double *ubar, double *vbar, double *wbar)
{
    if (!c) return MIMADV_ERR_ARG;
    int rc = check_adjoint_call(c, integrator, nsteps, u, v, w, qbar, ubar, vbar, wbar);
    if (rc) return rc;
    const int K = checkpoint_interval(checkpoint, nsteps);
    const int nseg = nsteps > 0 ? (nsteps + K - 1) / K : 0;
    if (nseg > 0 && !ck) return fail(c, MIMADV_ERR_ARG, "NULL checkpoint list");
    for (int g = 0; g < nseg; ++g)
        if (!ck[g]) return fail(c, MIMADV_ERR_ARG, "NULL checkpoint array");
    return adjoint_core(c, integrator, nsteps, dt, u, v, w, nullptr, ck, K, qbar, ubar, vbar, wbar);
}

}  // extern "C"

extern "C" {

int mimadv_set_reference_density(mimadv_ctx *c, const double *den0, const double *avden0)
{
    if (!c) return MIMADV_ERR_ARG;
    if (!den0) return fail(c, MIMADV_ERR_ARG, "den0 is NULL");
    const int nz = c->g.nz;
    std::vector<double> h(2 * (size_t)nz), fdz0(nz);
    for (int k = 0; k < nz; ++k) {
        if (!(den0[k] != 0.0) || !std::isfinite(den0[k])) return fail(c, MIMADV_ERR_ARG, "den0 must be finite, non-zero");
        h[k] = den0[k];
    }
    if (avden0) {
        for (int k = 0; k < nz; ++k) h[nz + k] = avden0[k];
    } else {
        HIPCHK(c, hipMemcpy(fdz0.data(), c->g.zm.fdz0, nz * sizeof(double), hipMemcpyDeviceToHost));
        for (int k = 1; k < nz; ++k)
            h[nz + k] = (fdz0[k] * den0[k] + fdz0[k - 1] * den0[k - 1]) / (fdz0[k] + fdz0[k - 1]);
        h[nz] = den0[0];
    }
    if (!c->d_den0) {
        if (hipMalloc(&c->d_den0, 2 * (size_t)nz * sizeof(double)) != hipSuccess)
            return fail(c, MIMADV_ERR_NOMEM, "density profile allocation failed");
    }
    HIPCHK(c, hipMemcpyAsync(c->d_den0, h.data(), 2 * (size_t)nz * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MIMADV_OK;
}

int mimadv_advection_1d(mimadv_ctx *c, const double *w0, const double *dens, int nf, const double *const *q,
                        double *const *xs, unsigned flags)
{
    if (!c) return MIMADV_ERR_ARG;
    if (nf < 0) return fail(c, MIMADV_ERR_ARG, "nf < 0");
    if (nf == 0) return MIMADV_OK;
    if (!w0 || !q || !xs) return fail(c, MIMADV_ERR_ARG, "NULL w0 or field list");
    for (int f = 0; f < nf; ++f)
        if (!q[f] || !xs[f]) return fail(c, MIMADV_ERR_ARG, "field pointer is NULL");
    const bool den0 = (flags & MIMADV_DREF_DEN0) != 0;
    if (den0 && !c->d_den0) return fail(c, MIMADV_ERR_ARG, "no reference density: call mimadv_set_reference_density");
    const int nz = c->g.nz;
    if (!c->d_w0 && hipMalloc(&c->d_w0, (size_t)nz * sizeof(double)) != hipSuccess)
        return fail(c, MIMADV_ERR_NOMEM, "w0 allocation failed");
    HIPCHK(c, hipMemcpyAsync(c->d_w0, w0, (size_t)nz * sizeof(double), hipMemcpyHostToDevice, c->stream));
    const double *dprof = den0 ? c->d_den0 : nullptr;
    const double *d3 = den0 ? nullptr : dens;
    const bool host = (flags & MIMADV_HOST_PTRS) != 0;
    const size_t bytes = c->n3() * sizeof(double);
    if (host) {
        int rc = ensure_stage(c, 3);
        if (rc) return rc;
        double *sq = c->stage, *sx = c->stage + c->n3(), *sd = c->stage + 2 * c->n3();
        if (d3) {
            HIPCHK(c, hipMemcpyAsync(sd, d3, bytes, hipMemcpyHostToDevice, c->stream));
            d3 = sd;
        }
        for (int f = 0; f < nf; ++f) {
            HIPCHK(c, hipMemcpyAsync(sq, q[f], bytes, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(sx, xs[f], bytes, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, launch_advection_1d(c->g, c->d_w0, dprof, d3, sq, sx, c->stream));
            HIPCHK(c, hipMemcpyAsync(xs[f], sx, bytes, hipMemcpyDeviceToHost, c->stream));
        }
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return MIMADV_OK;
    }
    for (int f = 0; f < nf; ++f) HIPCHK(c, launch_advection_1d(c->g, c->d_w0, dprof, d3, q[f], xs[f], c->stream));
    return MIMADV_OK;
}

int mimadv_get_momentum(mimadv_ctx *c, double *u, double *v, double *w, int inverse, unsigned flags)
{
    if (!c) return MIMADV_ERR_ARG;
    if (!c->d_den0) return fail(c, MIMADV_ERR_ARG, "no reference density: call mimadv_set_reference_density");
    const bool three_d = c->g.ny > 1;
    if (!u || !w || (three_d && !v)) return fail(c, MIMADV_ERR_ARG, "NULL wind array");
    const double *p[3] = {c->d_den0, c->d_den0, c->d_den0 + c->g.nz};
    double *x[3] = {u, three_d ? v : nullptr, w};
    const size_t bytes = c->n3() * sizeof(double);
    const int div = inverse ? 1 : 0;
    for (int a = 0; a < 3; ++a) {
        if (!x[a]) continue;
        if (flags & MIMADV_HOST_PTRS) {
            int rc = ensure_stage(c, 1);
            if (rc) return rc;
            HIPCHK(c, hipMemcpyAsync(c->stage, x[a], bytes, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, launch_level_scale(c->g, p[a], c->stage, c->stage, div, 0.0, c->stream));
            HIPCHK(c, hipMemcpyAsync(x[a], c->stage, bytes, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        } else {
            HIPCHK(c, launch_level_scale(c->g, p[a], x[a], x[a], div, 0.0, c->stream));
        }
    }
    return MIMADV_OK;
}

int mimadv_get_advective(mimadv_ctx *c, double u0shift, double v0shift, const double *u, const double *v,
                         const double *w, double *ua, double *va, double *wa, unsigned flags)
{
    if (!c) return MIMADV_ERR_ARG;
    if (!c->d_den0) return fail(c, MIMADV_ERR_ARG, "no reference density: call mimadv_set_reference_density");
    const bool three_d = c->g.ny > 1;
    if (!u || !w || !ua || !wa || (three_d && (!v || !va))) return fail(c, MIMADV_ERR_ARG, "NULL wind array");
    const double *p[3] = {c->d_den0, c->d_den0, c->d_den0 + c->g.nz};
    const double *in[3] = {u, three_d ? v : nullptr, w};
    double *out[3] = {ua, three_d ? va : nullptr, wa};
    const double shift[3] = {u0shift, v0shift, 0.0};
    const size_t bytes = c->n3() * sizeof(double);
    for (int a = 0; a < 3; ++a) {
        if (!in[a]) continue;
        if (flags & MIMADV_HOST_PTRS) {
            int rc = ensure_stage(c, 1);
            if (rc) return rc;
            HIPCHK(c, hipMemcpyAsync(c->stage, in[a], bytes, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, launch_level_scale(c->g, p[a], c->stage, c->stage, 1, shift[a], c->stream));
            HIPCHK(c, hipMemcpyAsync(out[a], c->stage, bytes, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        } else {
            HIPCHK(c, launch_level_scale(c->g, p[a], in[a], out[a], 1, shift[a], c->stream));
        }
    }
    return MIMADV_OK;
}

int mimadv_get_conserved(mimadv_ctx *c, const double *dens, int nf, double *const *q, int inverse, unsigned flags)
{
    if (!c) return MIMADV_ERR_ARG;
    if (!dens) return fail(c, MIMADV_ERR_ARG, "dens is NULL");
    if (nf < 0 || (nf > 0 && !q)) return fail(c, MIMADV_ERR_ARG, "bad field list");
    const size_t n3 = c->n3(), bytes = n3 * sizeof(double);
    const double *dd = dens;
    if (flags & MIMADV_HOST_PTRS) {
        int rc = ensure_stage(c, 2);
        if (rc) return rc;
        HIPCHK(c, hipMemcpyAsync(c->stage + n3, dens, bytes, hipMemcpyHostToDevice, c->stream));
        dd = c->stage + n3;
    }
    for (int f = 0; f < nf; ++f) {
        if (!q[f]) return fail(c, MIMADV_ERR_ARG, "field pointer is NULL");
        if (flags & MIMADV_HOST_PTRS) {
            HIPCHK(c, hipMemcpyAsync(c->stage, q[f], bytes, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, launch_field_scale(c->g, c->stage, dd, inverse ? 1 : 0, c->stream));
            HIPCHK(c, hipMemcpyAsync(q[f], c->stage, bytes, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        } else {
            HIPCHK(c, launch_field_scale(c->g, q[f], dd, inverse ? 1 : 0, c->stream));
        }
    }
    return MIMADV_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// many independent subdomains in one launch
// ---------------------------------------------------------------------------

struct mimadv_multi {
    int ndom = 0;
    std::vector<mimadv_ctx *> ctx;
    std::vector<const float *> u, v, w;
    std::vector<std::vector<float *>> q;
    std::vector<int> tile_x, kchunk;
    std::vector<size_t> n3;
    float *work = nullptr;               // two stage arrays per field
    std::vector<std::vector<float *>> s1, s2;
    FusedArgs *d_table = nullptr;        // one table per stage kind (3 x ndom)
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
};

namespace {

int mfail(mimadv_multi *m, int code, const std::string &msg)
{
    if (m) m->err = msg;
    return code;
}

#define MHIPCHK(m, expr)                                                                        \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return mfail(m, MIMADV_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// cost model: face evaluations of the padded tile cover per interior cell
int model_tile_x(int mx, int my)
{
    static const int cand[3] = {64, 32, 16};
    int best = 64;
    double bestc = 1e300;
    for (int tx : cand) {
        const int ty = MULTI_NT / tx;
        const double cells = double((mx + tx - 1) / tx * tx) * double((my + ty - 1) / ty * ty);
        const double c = cells * (2.0 + 1.0 / tx + 1.0 / ty) / (double(mx) * my);
        if (c < bestc - 1e-12) {
            bestc = c;
            best = tx;
        }
    }
    return best;
}

// argument block of subdomain d for one stage: in/out buffers 0 q, 1 s1, 2 s2
FusedArgs multi_args(mimadv_multi *m, int d, double c0, double c1, double dt, int in, int out)
{
    mimadv_ctx *c = m->ctx[d];
    FusedArgs A = base_args(c, m->u[d], m->v[d], m->w[d]);
    A.epi = 1;
    A.fill = 1;
    A.c0 = c0;
    A.c1 = c1;
    A.dt = dt;
    A.scheme = c->sp.scheme;
    A.tile_x = m->tile_x[d];
    A.tile_y = MULTI_NT / m->tile_x[d];
    A.kchunk = m->kchunk[d];
    A.n64_active = 0;
    A.n32_active = (int)m->q[d].size();
    auto pick = [&](int which, int f) -> float * {
        return which == 0 ? m->q[d][f] : (which == 1 ? m->s1[d][f] : m->s2[d][f]);
    };
    for (int f = 0; f < A.n32_active; ++f) {
        A.q32[f] = pick(in, f);
        A.q0_32[f] = m->q[d][f];
        A.out32[f] = pick(out, f);
    }
    return A;
}

// number the work items of a table, upload it into table slot `slot`;
// returns the work count in *nwork
int multi_upload(mimadv_multi *m, std::vector<FusedArgs> &tab, int slot, int *nwork)
{
    int begin = 0;
    for (FusedArgs &A : tab) {
        multi_prepare(A, begin);
        begin += A.nwork;
    }
    MHIPCHK(m, hipMemcpyAsync(m->d_table + (size_t)slot * m->ndom, tab.data(), tab.size() * sizeof(FusedArgs),
                              hipMemcpyHostToDevice, m->stream));
    *nwork = begin;
    return MIMADV_OK;
}

}  // namespace

extern "C" {

int mimadv_multi_create(int ndom, mimadv_ctx *const *ctx, const float *const *u, const float *const *v,
                        const float *const *w, const int *nf, float *const *const *q, mimadv_multi **out)
{
    if (!out) return MIMADV_ERR_ARG;
    *out = nullptr;
    mimadv_multi *m = new (std::nothrow) mimadv_multi();
    if (!m) return MIMADV_ERR_NOMEM;
    *out = m;
    if (ndom <= 0 || !ctx || !u || !v || !w || !nf || !q) return mfail(m, MIMADV_ERR_ARG, "bad subdomain list");
    m->ndom = ndom;
    size_t total = 0;
    for (int d = 0; d < ndom; ++d) {
        mimadv_ctx *c = ctx[d];
        if (!c) return mfail(m, MIMADV_ERR_ARG, "NULL context");
        if (c->g.ny == 1) return mfail(m, MIMADV_ERR_UNSUPPORTED, "subdomains must be 3D");
        int rc = check_batch(c, u[d], v[d], w[d]);
        if (rc) return mfail(m, rc, "subdomain " + std::to_string(d) + ": " + c->err);
        if (nf[d] < 0 || nf[d] > FUSED_MAX32) return mfail(m, MIMADV_ERR_ARG, "1..18 fields per subdomain");
        if (nf[d] > 0 && !q[d]) return mfail(m, MIMADV_ERR_ARG, "NULL field list");
        m->ctx.push_back(c);
        m->u.push_back(u[d]);
        m->v.push_back(v[d]);
        m->w.push_back(w[d]);
        std::vector<float *> fl;
        for (int f = 0; f < nf[d]; ++f) {
            if (!q[d][f]) return mfail(m, MIMADV_ERR_ARG, "NULL field pointer");
            fl.push_back(q[d][f]);
        }
        m->q.push_back(fl);
        m->n3.push_back(c->n3());
        m->tile_x.push_back(model_tile_x(c->g.nx - 5, c->g.ny - 5));
        m->kchunk.push_back(0);
        total += 2 * c->n3() * (size_t)nf[d];
    }
    m->stream = ctx[0]->stream;
    if (total > 0 && hipMalloc(&m->work, total * sizeof(float)) != hipSuccess)
        return mfail(m, MIMADV_ERR_NOMEM, "stage array allocation failed");
    size_t off = 0;
    for (int d = 0; d < ndom; ++d) {
        std::vector<float *> a, b;
        for (size_t f = 0; f < m->q[d].size(); ++f) {
            a.push_back(m->work + off);
            off += m->n3[d];
            b.push_back(m->work + off);
            off += m->n3[d];
            // stage arrays start as copies so that their ghost columns are defined
            MHIPCHK(m, hipMemcpyAsync(a.back(), m->q[d][f], m->n3[d] * sizeof(float), hipMemcpyDeviceToDevice,
                                      m->stream));
            MHIPCHK(m, hipMemcpyAsync(b.back(), m->q[d][f], m->n3[d] * sizeof(float), hipMemcpyDeviceToDevice,
                                      m->stream));
        }
        m->s1.push_back(a);
        m->s2.push_back(b);
    }
    if (hipMalloc(&m->d_table, 3 * (size_t)ndom * sizeof(FusedArgs)) != hipSuccess)
        return mfail(m, MIMADV_ERR_NOMEM, "table allocation failed");
    MHIPCHK(m, hipEventCreate(&m->ev0));
    MHIPCHK(m, hipEventCreate(&m->ev1));
    MHIPCHK(m, hipStreamSynchronize(m->stream));
    return MIMADV_OK;
}

void mimadv_multi_destroy(mimadv_multi *m)
{
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    if (m->work) (void)hipFree(m->work);
    if (m->d_table) (void)hipFree(m->d_table);
    if (m->ev0) (void)hipEventDestroy(m->ev0);
    if (m->ev1) (void)hipEventDestroy(m->ev1);
    delete m;
}

const char *mimadv_multi_last_error(const mimadv_multi *m)
{
    return m ? m->err.c_str() : "NULL plan";
}

int mimadv_multi_set_tiles(mimadv_multi *m, const int *tile_x, const int *kchunk)
{
    if (!m) return MIMADV_ERR_ARG;
    for (int d = 0; d < m->ndom; ++d) {
        if (tile_x && !multi_tile_ok(tile_x[d], MULTI_NT / (tile_x[d] > 0 ? tile_x[d] : 1)))
            return mfail(m, MIMADV_ERR_ARG, "tile_x must be 16, 32 or 64");
        if (kchunk && kchunk[d] < 0) return mfail(m, MIMADV_ERR_ARG, "kchunk < 0");
    }
    for (int d = 0; d < m->ndom; ++d) {
        if (tile_x) m->tile_x[d] = tile_x[d];
        if (kchunk) m->kchunk[d] = kchunk[d];
    }
    return MIMADV_OK;
}

int mimadv_multi_get_tiles(const mimadv_multi *m, int *tile_x, int *tile_y, int *kchunk)
{
    if (!m) return MIMADV_ERR_ARG;
    for (int d = 0; d < m->ndom; ++d) {
        if (tile_x) tile_x[d] = m->tile_x[d];
        if (tile_y) tile_y[d] = MULTI_NT / m->tile_x[d];
        if (kchunk) kchunk[d] = m->kchunk[d] > 0 ? std::min(m->kchunk[d], MULTI_KC) : MULTI_KC;
    }
    return MIMADV_OK;
}

int mimadv_multi_autotune(mimadv_multi *m, int reps)
{
    if (!m) return MIMADV_ERR_ARG;
    if (reps < 1) reps = 1;
    static const int cand_x[3] = {64, 32, 16};
    static const int cand_k[2] = {4, MULTI_KC};
    // A candidate is timed at saturation: the subdomain's work list is
    // repeated until the launch holds at least `target` work items (the
    // copies write the same values to the same scratch stage array), so a
    // small subdomain is not rewarded for merely creating more work items.
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
        ncu = 256;
    const int target = 8 * ncu;
    const int max_copies = 256;
    FusedArgs *d_tune = nullptr;
    if (hipMalloc(&d_tune, (size_t)max_copies * sizeof(FusedArgs)) != hipSuccess)
        return mfail(m, MIMADV_ERR_NOMEM, "autotune table allocation failed");
    struct Key {
        int mx, my, nz, nf, scheme;
        bool operator==(const Key &o) const
        {
            return mx == o.mx && my == o.my && nz == o.nz && nf == o.nf && scheme == o.scheme;
        }
    };
    std::vector<std::pair<Key, std::pair<int, int>>> memo;
    int status = MIMADV_OK;
    for (int d = 0; d < m->ndom && status == MIMADV_OK; ++d) {
        mimadv_ctx *c = m->ctx[d];
        Key key{c->g.nx - 5, c->g.ny - 5, c->g.nz, (int)m->q[d].size(), c->sp.scheme};
        bool found = false;
        for (auto &e : memo)
            if (e.first == key) {
                m->tile_x[d] = e.second.first;
                m->kchunk[d] = e.second.second;
                found = true;
                break;
            }
        if (found || m->q[d].empty()) continue;
        double best = 1e300;
        int bx = m->tile_x[d], bk = m->kchunk[d];
        for (int tx : cand_x) {
            for (int kc : cand_k) {
                m->tile_x[d] = tx;
                m->kchunk[d] = kc;
                // a middle stage into s2 from q: the fields are only read
                FusedArgs one = multi_args(m, d, 0.75, 0.25, 1e-3, 0, 2);
                multi_prepare(one, 0);
                const int copies = std::min(max_copies, std::max(1, (target + one.nwork - 1) / one.nwork));
                std::vector<FusedArgs> tab(copies, one);
                for (int r = 0; r < copies; ++r) tab[r].work_begin = r * one.nwork;
                const int nw = copies * one.nwork;
                hipError_t e = hipMemcpy(d_tune, tab.data(), tab.size() * sizeof(FusedArgs), hipMemcpyHostToDevice);
                if (e == hipSuccess) e = launch_multi(d_tune, copies, nw, m->stream);  // warm-up
                if (e == hipSuccess) e = hipEventRecord(m->ev0, m->stream);
                for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_multi(d_tune, copies, nw, m->stream);
                if (e == hipSuccess) e = hipEventRecord(m->ev1, m->stream);
                if (e == hipSuccess) e = hipEventSynchronize(m->ev1);
                float ms = 0.0f;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, m->ev0, m->ev1);
                if (e != hipSuccess) {
                    status = mfail(m, MIMADV_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
                    break;
                }
                const double per = (double)ms / copies;
                if (per < best) {
                    best = per;
                    bx = tx;
                    bk = kc;
                }
            }
            if (status != MIMADV_OK) break;
        }
        m->tile_x[d] = bx;
        m->kchunk[d] = bk;
        memo.push_back({key, {bx, bk}});
    }
    (void)hipFree(d_tune);
    return status;
}

int mimadv_multi_advance(mimadv_multi *m, int integrator, int nsteps, const double *dt)
{
    if (!m) return MIMADV_ERR_ARG;
    if (integrator != MIMADV_INT_HEUN && integrator != MIMADV_INT_SSPRK3)
        return mfail(m, MIMADV_ERR_UNSUPPORTED, "integrators: MIMADV_INT_HEUN, MIMADV_INT_SSPRK3");
    if (nsteps < 0 || !dt) return mfail(m, MIMADV_ERR_ARG, "bad step count or dt");
    struct Stage {
        double c0, c1;
        int in, out;
    };
    std::vector<Stage> st;
    if (integrator == MIMADV_INT_HEUN) st = {{0.0, 1.0, 0, 1}, {0.5, 0.5, 1, 0}};
    else st = {{0.0, 1.0, 0, 1}, {0.75, 0.25, 1, 2}, {1.0 / 3.0, 2.0 / 3.0, 2, 0}};
    std::vector<std::vector<FusedArgs>> tabs(st.size());
    std::vector<int> nwork(st.size(), 0);
    for (size_t s = 0; s < st.size(); ++s) {
        for (int d = 0; d < m->ndom; ++d)
            if (!m->q[d].empty()) tabs[s].push_back(multi_args(m, d, st[s].c0, st[s].c1, dt[d], st[s].in, st[s].out));
        int rc = multi_upload(m, tabs[s], (int)s, &nwork[s]);
        if (rc) return rc;
    }
    MHIPCHK(m, hipEventRecord(m->ev0, m->stream));
    for (int n = 0; n < nsteps; ++n)
        for (size_t s = 0; s < st.size(); ++s)
            MHIPCHK(m, launch_multi(m->d_table + s * m->ndom, (int)tabs[s].size(), nwork[s], m->stream));
    MHIPCHK(m, hipEventRecord(m->ev1, m->stream));
    // the host tables must outlive their (pageable, staged) upload
    MHIPCHK(m, hipStreamSynchronize(m->stream));
    return MIMADV_OK;
}

int mimadv_multi_last_ms(mimadv_multi *m, double *ms)
{
    if (!m || !ms) return MIMADV_ERR_ARG;
    MHIPCHK(m, hipEventSynchronize(m->ev1));
    float t = 0.0f;
    MHIPCHK(m, hipEventElapsedTime(&t, m->ev0, m->ev1));
    *ms = t;
    return MIMADV_OK;
}

}  // extern "C"
