// fused.hpp -- arguments of the batched mixed-precision advection kernel.
#pragma once
#include <hip/hip_runtime.h>

namespace mimadv {

constexpr int FUSED_MAX64 = 2;
constexpr int FUSED_MAX32 = 18;

struct FusedArgs {
    int nx, ny, nz, mx, my, three_d;
    double dx1, dy1, dz;
    const double *dz1w, *dzw, *dz1, *fdz0;  // device metrics
    const float *u, *v, *w;                 // advecting velocity (FP32)
    const double *q64[FUSED_MAX64];         // stencil input
    const double *q0_64[FUSED_MAX64];       // state at the start of the step
    double *out64[FUSED_MAX64];             // tendency or next stage state
    const float *q32[FUSED_MAX32];
    const float *q0_32[FUSED_MAX32];
    float *out32[FUSED_MAX32];
    int n64_active, n32_active;  // fields beyond these are padding
    int epi;                     // 0: out = -div; 1: Shu-Osher stage
    double c0, c1, dt;           // out = c0*q0 + c1*(q + dt*(-div))
    int fill;                    // write periodic ghost images
    int muscl_limiter;
    double lim_tol, lim_tol_inv;
    int weno_bottom;
    int ntx, nty, kchunk, nwork, xcd_remap;
    int tile_y;  // rows of the tile in 3D (2, 4, 8; 0: see tile_x)
    int tile_x;  // tile width: batched launches 64, 32 (with tile_y 8 or 0) or
                 // 16 (16x16); both 0: 32x8; multi-domain launches
                 // tile_x * tile_y = 256
    // multi-domain launches only
    int scheme;        // 0 WENO, 2 MUSCL
    int work_begin;    // first work id of this subdomain in the launch
    int dirmask;  // directions included in the divergence (bit0 x, bit1 y, bit2 z)
};

// FP32 field count of one launch (-1 if above FUSED_MAX32).
int fused_padded_count(int n32);
hipError_t launch_fused(int scheme, int n64, int n32, const FusedArgs &A, hipStream_t s, int blocks_per_cu);

// Multi-domain launch: table[0..ndom-1] on the device, one entry per
// subdomain (FP32 fields only, 3D, tile_x * tile_y = 256, ntx/nty/kchunk/
// work_begin filled in by multi_prepare), nwork the total work count.
constexpr int MULTI_NT = 256;
constexpr int MULTI_KC = 6;  // levels of a multi-domain work item (maximum and default)
bool multi_tile_ok(int tile_x, int tile_y);
void multi_prepare(FusedArgs &A, int work_begin);  // sets ntx, nty, nwork; kchunk clamped
hipError_t launch_multi(const FusedArgs *d_table, int ndom, int nwork, hipStream_t s);

}  // namespace mimadv
