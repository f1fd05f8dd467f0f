// mimadv_internal.hpp -- shared host/device declarations of libmimadv.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "schemes_fp64.hpp"

namespace mimadv {

// Grid description held on the device side of a context.
struct GridDev {
    int nx, ny, nz;
    double dx, dy, dz;
    int bcl[6];
    ZMetrics zm;  // device pointers
};

// Epilogue of the scalar kernel.
enum Epilogue : int {
    EPI_ACCUM = 0,      // xs = ((xs - X) - Y) - Z
    EPI_OVERWRITE = 1,  // xs = ((0 - X) - Y) - Z
    EPI_HEUN = 2,       // a = ((0 - X) - Y) - Z; tend = a + b*tend; qn = q0 + adt*tend
    EPI_SO = 3,         // a as above; qn = c0*q0 + c1*(q + dt*a)  (Shu-Osher stage)
};

struct ScalarArgs {
    const double *q;  // stencil input
    const double *u, *v, *w;
    double *xs;        // ACCUM / OVERWRITE output
    const double *q0;  // stage: state at the start of the step
    double *tend;      // HEUN accumulator
    double *qn;        // stage output (may alias q0, never q)
    double b_lm, adt;  // HEUN coefficients
    double c0, c1, dt; // SO coefficients
    int adv;
    int epi;
    int dirmask;
    int fill;  // write periodic ghost images of every interior value (stage epilogues)
};

// Adjoint of one Shu-Osher stage of the 'weno' flux-form tendency (adjoint.hip):
//   out (=|+=) ic*qbn + (dL/dq)^T (ls*qbn);  acc0 (=|+=) c0*qbn;
//   ub/vb/wb += (dL/d(u,v,w))^T (ls*qbn).  Interior cells only.
struct AdjArgs {
    int nx, ny, nz;
    double dx1, dy1;
    const double *dz1w;  // device metric
    const double *u, *v, *w;
    const double *q;    // linearisation state (the stage input)
    const double *qbn;  // adjoint of the stage output; never aliases out or acc0
    double *out;
    double *acc0;
    double *ub, *vb, *wb;
    double ic, ls, c0;
    int accumulate;  // out += instead of out =
    int acc0_mode;   // 0: untouched, 1: acc0 = c0*qbn, 2: acc0 += c0*qbn
    int kchunk;      // levels per workgroup (0: default)
    int weno5;       // the WENO5-JS extension instead of the reference 'weno'
    int ntx, nty, nwork;  // set by the launcher
};

// Launchers (scalar_fp64.hip, momentum_fp64.hip, util.hip, adjoint.hip).
hipError_t launch_adjoint_stage(const AdjArgs &a, hipStream_t s);
// dst = src on interior cells, 0 on ghost cells
hipError_t launch_interior_copy(const GridDev &g, const double *src, double *dst, hipStream_t s);
hipError_t launch_scalar_fp64(const GridDev &g, const SchemeParams &sp, const ScalarArgs &a, hipStream_t s);

// One positivity-preserving Shu-Osher stage of the flux-form 'weno'
// tendency: qn = c0*q0 + c1*(q + dt*L~(q)), L~ from face fluxes scaled by
// their donor cell's theta (scratch fx, fy, fz, th: nx*ny*nz each).
struct PosArgs {
    const double *q, *u, *v, *w, *q0;
    double *qn;
    double c0, c1, dt, dx1, dy1;
    double *fx, *fy, *fz, *th;
};
hipError_t launch_positive_stage(const GridDev &g, const SchemeParams &sp, const PosArgs &a, hipStream_t s);
// column-marching kernels (weno_march.hip): 'weno', WENO5, MUSCL, QUICK,
// Lax-Wendroff and the unlimited forms; kchunk 0 = default
hipError_t launch_march(const GridDev &g, const SchemeParams &sp, const ScalarArgs &a, int kchunk, hipStream_t s);
hipError_t launch_momentum_fp64(const GridDev &g, int mom_ord, int lavisc, int ntime, const double *mu,
                                const double *mv, const double *mw, const double *ut, const double *vt,
                                const double *wt, double *ua, double *va, double *wa, hipStream_t s);

// Batched momentum stage (FP32 storage): m = (rho0 u, rho0 v, rho0 w) stage
// input, m0 the step-start state, out the result; ut/vt/wt the advecting
// velocities.  epi 0: tendency, 1: Shu-Osher stage c0*m0 + c1*(m + dt*L).
struct MomBatchArgs {
    const float *m[3];
    const float *m0[3];
    float *out[3];
    const float *ut, *vt, *wt;
    double c0, c1, dt;
    int epi, fill;
};
hipError_t launch_momentum_batch(const GridDev &g, int mom_ord, int lavisc, const MomBatchArgs &a, hipStream_t s);
hipError_t launch_fill_scalar(const GridDev &g, double *x, hipStream_t s);
// lbcu_p4 of a single-tile run (u, v, w; v unused in 2D)
hipError_t launch_fill_winds(const GridDev &g, double *u, double *v, double *w, hipStream_t s);
// integ_mom of one component (tw = the sponge factor for w, 1 otherwise)
hipError_t launch_integ_mom_one(const GridDev &g, int comp, int sorder, int im, double dt, double tw,
                                const double *old, const double *a, double *tend, double *x, hipStream_t s);
// integ_mom on (u, v, w) with the sponge factor tw applied to w
hipError_t launch_integ_mom(const GridDev &g, int sorder, int im, double dt, double tw, const double *const *old,
                            const double *const *a, double *const *tend, double *const *x, hipStream_t s);
// the reference's implicit sponge factor of integ_mom (host, FP64)
double sponge_factor(int nz, int sorder, int im, double dt, const double *z0, double zmax, double zdamp, int sponge);
// xs -= dref*w0*grad_1d(q/dref) on interior columns; dref = d3 (3D) if not
// NULL, else dprof (levels) if not NULL, else 1 (util.hip)
hipError_t launch_advection_1d(const GridDev &g, const double *w0, const double *dprof, const double *d3,
                               const double *q, double *xs, hipStream_t s);
hipError_t launch_integ_scal(long n, int sorder, int im, double dt, const double *xold, const double *a,
                             double *tend, double *x, hipStream_t s);
hipError_t launch_totalav(const GridDev &g, const double *x, double *d_out2, hipStream_t s);
// out = in*p(k) or in/p(k), minus shift; p a level profile (in may alias out)
hipError_t launch_level_scale(const GridDev &g, const double *prof, const double *in, double *out, int divide,
                              double shift, hipStream_t s);
// q = q*d or q/d elementwise
hipError_t launch_field_scale(const GridDev &g, double *q, const double *d, int divide, hipStream_t s);
hipError_t launch_wsum(const GridDev &g, const void *x, bool fp32, double *d_part, double *d_out, hipStream_t s);

// Domain budgets (dzw-weighted interior sums) of up to BUDGET_MAXF fields in
// two launches: per-block partials reduced with wavefront shuffles, then one
// wave per field (deterministic order).  part holds nf*nb doubles.
constexpr int BUDGET_MAXF = 32;
struct BudgetArgs {
    int nx, ny, nz, nb;
    const void *q[BUDGET_MAXF];
    int fp32[BUDGET_MAXF];
    const double *dzw;
    double *part;
};
constexpr int BUDGET_MAXNB = 1024;  // first-pass blocks per field
constexpr size_t BUDGET_SCRATCH = (size_t)BUDGET_MAXF * (BUDGET_MAXNB + 1);
constexpr int BUDGET_MAX_DEVICES = 64;
int budget_blocks(const GridDev &g);
// Device address of the budget scratch on `device` (made current by the
// caller), looked up once per device and kept.
hipError_t budget_scratch_setup(int device, double **out);
hipError_t launch_budgets(const BudgetArgs &a, int nf, double *d_out, hipStream_t s);

// Host-side PPM per-level coefficient table (12 doubles per level).
std::vector<double> ppm_coefficients(int nz, const double *dz1w);

// Level coefficients of the FP32 batched momentum operator, MOMZ_N floats
// per level k (1-based, entry k-1), from the reference's metrics
// (windadv.f90:656-688, 1077-1121):
//   [0] wa = dzw(k+1)/(dzw(k+1)+dzw(k)), [1] wb = dzw(k)/(dzw(k+1)+dzw(k))
//       (the dzw-weighted face average between levels k and k+1);
//   [2] dz1w(k); [3] dz1(k);
//   [4] 1/(dzw(k-1)+dzw(k)+dzw(k+1)); [5] 1/(dzw(k)+dzw(k-1));
//   [6] (level nz only) dzw(nz-1)^2/(dzw(nz)+dzw(nz-1))^2, the 'nnf' lid
//       factor of advf(nz-1).
constexpr int MOMZ_N = 8;
std::vector<float> momentum_z_table(int nz, const double *dzw, const double *dz1w, const double *dz1);

}  // namespace mimadv
