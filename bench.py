#!/usr/bin/env python3
"""Benchmark of the MI355X advection operator (one JSON line on stdout).

Workload (default, --workload fused): the fused multi-field batch on one
GPU -- 512x512x160 interior cells, dx = dy = 50 m, dz stretched linearly
from 20 m to 80 m, periodic laterally and no-flux at floor and lid,
divergence-free turbulent FP32 wind (curl of a seeded vector potential,
k^-5/3, 3 m/s RMS) and 20 prognostic fields: the momenta rho0*(u, v, w) in
FP32 (momentum operator of qnadv_uvw), theta and qv in FP64, and 15
positive-definite FP32 species (log-normal, mu in [-20, -6], sigma 1.5,
30 % exact zeros); reference 'weno' scheme, SSP-RK3, every step replayed
from a hipGraph, all state resident in HBM.  A "step" is one full time
step (three stages of all 20 fields); K steps are timed.

value: cell.species-updates/s = interior cells x fields x stages x K / time.
roofline: the dominant kernel (the FP32 species launch of a stage) against
the 8 TB/s HBM peak: algorithmic bytes per launch over its average
duration, from HIP events the library records around every launch of a
second, eager run of K steps on its stream (MIMADV_PROFILE).
cpu_baseline: the C restatement of the reference (the oracle, FP64) on the
host cores, on a bounded sample of the same 20-field workload.
sub_records: the plain per-field FP64 forward at 256x256x128 (with its
scheme variants), the trajectory adjoint of 200 SSP-RK3 steps (cost ratio)
and the 48 uneven subdomains, each measured the same way.

Other workloads alone: --workload plain | adjoint | multi.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch (per stage for the batched workloads) of the dominant
# kernels, measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of
# this script, each counter scaled by what it reads for a 256 MiB device copy
# (--copy-probe 256: FETCH_SIZE counts half the bytes read, WRITE_SIZE all
# bytes written).  Reported only for the exact configuration measured.
HBM_TRAFFIC = {
    ("plain", "256,256,128", "weno", "heun", 2): 551.2e6,
    ("adjoint", "256,256,128", "weno", "ssprk3", 1): 1078.2e6,
    ("fused", "f32", "512,512,160", "weno", "ssprk3", 20): 13.542e9,
}
SCHEME_ID = {"weno": 0, "ppm": 1, "muscl": 2, "weno5": 3, "quick": 4, "lw": 5}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_case(maxx, maxy, nz, seed, rms):
    from fdd3260_amd import cases, grid as gm

    g = gm.stretched(maxx, maxy, nz, dx=50.0, dy=50.0, dz_bottom=20.0, dz_top=80.0)
    u, v, w = cases.turbulent_wind(g, seed=seed, rms=rms)
    q = cases.lognormal(g, seed=seed + 1, mu=-9.0, sigma=1.0)
    th = cases.theta_profile(g, seed=seed + 2, noise=0.1)
    dt = cases.cfl_dt(g, u, v, w, 0.5)
    return g, u, v, w, [q, th], dt


def stage_bytes(integrator, ncell):
    """Algorithmic HBM bytes of each stage launch of one FP64 field: every
    array the stage touches, read or written once, plus the 3 velocity
    components."""
    vel = 3 * 8
    if integrator == "heun":
        # stage 1: read q (also q0), write tend, write q*; stage 2: read q*,
        # q0, tend, write q
        per = [(1 + 2) * 8 + vel, (3 + 1) * 8 + vel]
    else:
        # Shu-Osher stages: stage 1 reads q (= q0) writes q1; later stages
        # read the stage state and q0, write the next state
        per = [(1 + 1) * 8 + vel, (2 + 1) * 8 + vel, (2 + 1) * 8 + vel]
    return [p * ncell for p in per]


def cpu_baseline(g, u, v, w, fields, dt, scheme, limiter, integrator, steps, threads):
    from oracle.oracle import Oracle

    ora = Oracle(g, scal_adv=scheme, muscllimiter=limiter)
    ora.set_threads(threads)
    qs = [f.copy() for f in fields]
    t0 = time.perf_counter()
    for _ in range(steps):
        for q in qs:
            (ora.heun_step if integrator == "heun" else ora.ssprk3_step)(dt, u, v, w, q)
    el = time.perf_counter() - t0
    nst = 2 if integrator == "heun" else 3
    ncell = g.maxx * max(g.maxy, 1) * g.nz
    return ncell * len(fields) * nst * steps / el, el


def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model


def fused_cpu_baseline(g, u, v, w, fields, dt, steps, threads):
    """The FP64 oracle on the same 20-field workload: `steps` SSP-RK3 steps
    of every field (momenta with qnadv_uvw, the rest with the scalar
    'weno'), on `threads` host threads; plus one stage of one species on a
    single thread.  Returns (rate, seconds, rate on one thread, its seconds)."""
    from oracle.oracle import Oracle

    ora = Oracle(g, scal_adv="weno")
    ora.set_threads(threads)
    ncell = g.maxx * max(g.maxy, 1) * g.nz
    state = [f.copy() for f in fields]
    t0 = time.perf_counter()
    for _ in range(steps):
        m0 = state[:3]
        cur = [m.copy() for m in m0]
        for c0, c1 in ((0.0, 1.0), (0.75, 0.25), (1.0 / 3.0, 2.0 / 3.0)):
            L = ora.qnadv_uvw(0, *cur, u, v, w)
            cur = [c0 * a0 + c1 * (x - dt * l) for a0, x, l in zip(m0, cur, L)]
            for x in cur:
                ora.fill_scalar(x)
            del L
        state[:3] = cur
        for q in state[3:]:
            ora.ssprk3_step(dt, u, v, w, q)
    el = time.perf_counter() - t0
    rate = ncell * len(fields) * 3 * steps / el
    ora.set_threads(1)
    xs = np.zeros(g.shape)
    t0 = time.perf_counter()
    ora.scalar_tendency(dt, u, v, w, fields[5], xs)
    el1 = time.perf_counter() - t0
    return rate, el, ncell / el1, el1


def fused(args):
    """Batched mixed-precision run (512x512x160 by default): the three
    momenta rho0*(u, v, w) in FP32 (advanced with the momentum operator),
    theta and qv in FP64, 15 positive-definite species in FP32 (mass/number
    log-normal, mu in [-20, -6], sigma 1.5, 30 % exact zeros), FP32 velocity
    at 3 m/s RMS."""
    import torch

    from fdd3260_amd import Advection, _lib, cases, grid as gm

    size = args.size or "512,512,160"
    integ = args.integrator or "ssprk3"
    steps = args.steps if args.steps is not None else 50
    maxx, maxy, nz = (int(s) for s in size.split(","))
    t0 = time.time()
    g = gm.stretched(maxx, maxy, nz, dx=50.0, dy=50.0, dz_bottom=20.0, dz_top=80.0)
    u, v, w = cases.turbulent_wind(g, seed=2, rms=3.0)
    u, v, w = (a.astype(np.float32).astype(np.float64) for a in (u, v, w))
    dt = cases.cfl_dt(g, u, v, w, 0.5)
    host, prec = cases.config3_fields(g, u, v, w)
    dev = torch.device("cuda:0")
    du, dv, dw = (torch.from_numpy(a.astype(np.float32)).to(dev) for a in (u, v, w))
    init = [torch.from_numpy(np.ascontiguousarray(h.astype(np.float32 if p == "f32" else np.float64))).to(dev)
            for h, p in zip(host, prec)]
    fields = [f.clone() for f in init]
    log(f"fused case built in {time.time() - t0:.1f}s, dt = {dt:.4f}s")
    scheme = args.scheme if args.scheme in ("weno", "muscl", "weno5") else "weno"
    adv = Advection(g, scal_adv=scheme, muscllimiter=args.limiter)
    adv.set_tuning(group32=args.group32, kchunk=args.kchunk, blocks_per_cu=args.blocks_per_cu, tile_y=args.tile_y,
                   tile_x=args.tile_x)
    code = _lib.INT_HEUN if integ == "heun" else _lib.INT_SSPRK3
    nst = 2 if integ == "heun" else 3
    graph = not args.no_graph
    if args.warmup > 0:
        adv.batch_advance(code, args.warmup, dt, du, dv, dw, fields, graph=graph, momentum=True)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    adv.batch_advance(code, steps, dt, du, dv, dw, fields, graph=graph, momentum=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    log(f"fused: {steps} steps in {wall:.3f}s")
    # the same steps again, eager, with an event pair around every launch
    for f, i0 in zip(fields, init):
        f.copy_(i0)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    adv.batch_advance(code, steps, dt, du, dv, dw, fields, momentum=True, profile=True)
    torch.cuda.synchronize()
    wall_eager = time.perf_counter() - t_start
    kinds = adv.batch_kernel_ms()
    for f in fields:
        if not bool(torch.isfinite(f).all()):
            raise RuntimeError("non-finite state after the timed steps")
    ncell = maxx * maxy * nz
    nf = len(fields)
    updates = ncell * nf * nst * steps
    # algorithmic bytes per stage and launch kind: stage 1 reads q and writes
    # q1; later stages read the stage state and q0 and write; the launch
    # reads the 3 FP32 velocities once
    groups = {"momentum": [0, 1, 2], "f64": [3, 4], "f32": list(range(5, nf))}

    def kind_bytes(idx):
        fb = sum(fields[n].element_size() for n in idx)
        per = [(2 * fb + 12) * ncell] + [(3 * fb + 12) * ncell] * (nst - 1)
        return sum(per) / nst

    per_kind = {}
    for k, idx in groups.items():
        ms, cnt = kinds[k]
        if cnt == 0:
            continue
        avg_s = ms * 1e-3 / cnt
        ab = kind_bytes(idx)
        per_kind[k] = {"launches": cnt, "avg_launch_us": avg_s * 1e6, "alg_bytes_per_launch": ab,
                       "achieved_GBps": ab / avg_s / 1e9, "frac": ab / avg_s / 1e9 / HBM_PEAK_GBS}
    dom = max(per_kind, key=lambda k: per_kind[k]["avg_launch_us"])
    stage_us = sum(v["avg_launch_us"] for v in per_kind.values())
    stage_bytes_all = sum(kind_bytes(idx) for idx in groups.values())
    sid = SCHEME_ID[scheme]
    tile = ("32,8" if args.tile_x == 32 or (args.tile_x == 0 and args.tile_y == 0)
            else "16,16" if args.tile_x == 16 else f"64,{args.tile_y or 4}")
    kernel_names = {"f32": f"fused_kernel<{sid},{tile},float> (15 FP32 species)",
                    "f64": f"fused_kernel<{sid},{tile},double> (theta, qv)",
                    "momentum": "momentum_batch_kernel<32,8> (3 FP32 momenta)"}
    cpu = None
    cpu_steps = args.cpu_steps if args.cpu_steps is not None else 1
    if cpu_steps > 0:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"cpu baseline: {cpu_steps} SSP-RK3 step(s) of the 20 fields with {threads} threads")
        rate, el, rate1, el1 = fused_cpu_baseline(g, u, v, w, host, dt, cpu_steps, threads)
        cpu = {
            "value": rate,
            "unit": "cell*species-updates/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{cpu_steps} SSP-RK3 step(s) of the same {maxx}x{maxy}x{nz} 20-field workload in FP64 "
                      f"(momenta with qnadv_uvw, 17 scalars with weno), {el:.1f}s on {threads} threads; one "
                      f"stage of one species on 1 thread: {rate1:.3e} updates/s ({el1:.1f}s); C restatement "
                      f"of the reference, gcc 11.4 -O2 -ffp-contract=off, OpenMP over lines; host CPU "
                      f"{cpu_info()}",
            "value_1core": rate1,
        }
    return {
        "metric": "cell*species-updates/s and % HBM roofline at 1 GPU",
        "value": updates / wall,
        "unit": "cell*species-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded turbulent wind, rho0 u/v/w momenta, log-normal species with 30% zeros, theta, qv)",
        "config": {"workload": f"fused multi-field batch {maxx}x{maxy}x{nz}, {nf} fields (3 FP32 momenta + 2 FP64 "
                               f"+ {nf - 5} FP32 species), {adv.scal_adv}, {integ}, hipGraph={graph}",
                   "cells": ncell, "fields": nf, "stages_per_step": nst, "dt_s": dt,
                   "eager_profiled_value": updates / wall_eager,
                   "stage_us": stage_us, "stage_alg_bytes": stage_bytes_all,
                   "stage_frac": stage_bytes_all / (stage_us * 1e-6) / 1e9 / HBM_PEAK_GBS},
        "roofline": {"bound": "hbm", "kernel": kernel_names[dom],
                     "achieved": per_kind[dom]["achieved_GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": per_kind[dom]["frac"],
                     "traffic": HBM_TRAFFIC.get(("fused", dom, size, scheme, integ, nf)),
                     "alg_bytes_per_launch": per_kind[dom]["alg_bytes_per_launch"],
                     "avg_launch_us": per_kind[dom]["avg_launch_us"],
                     "kernels": {kernel_names[k]: v for k, v in per_kind.items()}},
        "cpu_baseline": cpu,
    }


def adjoint(args):
    """Trajectory adjoint: J = sum over the top half of rho0*q(T)^2 after
    `steps` SSP-RK3 'weno' steps of the log-normal scalar of the plain
    workload; gradient with respect to (q0, u, v, w) in FP64, checkpoints
    every `checkpoint` steps.  value = wall(adjoint incl. recomputation of
    every segment) / wall(forward); the forward that records the
    checkpoints is timed beside the plain one.  Correctness at this size is
    checked in the test-suite (dot product against the CPU tangent-linear
    model)."""
    import numpy as np
    import torch

    from fdd3260_amd import Advection, _lib, cases

    size = args.size or "256,256,128"
    maxx, maxy, nz = (int(s) for s in size.split(","))
    t0 = time.time()
    g, u, v, w, fields, dt = build_case(maxx, maxy, nz, seed=2, rms=2.0)
    q0 = fields[0]
    dev = torch.device("cuda:0")
    U, V, W, Q0 = (torch.from_numpy(a).to(dev) for a in (u, v, w, q0))
    rho0 = 1.2 * np.exp(-g.zc() / 8000.0)
    top = np.zeros(g.shape)
    I = g.interior()
    top[nz // 2:] = rho0[nz // 2:, None, None]
    mask = np.zeros(g.shape)
    mask[I] = 1.0
    TOP = torch.from_numpy(top * mask).to(dev)
    log(f"adjoint case built in {time.time() - t0:.1f}s, dt = {dt:.4f}s")
    scheme = args.scheme if args.scheme in ("weno", "weno5") else "weno"
    adv = Advection(g, scal_adv=scheme)
    adv.set_stream(torch.cuda.current_stream(dev).cuda_stream)
    code, ck = _lib.INT_SSPRK3, args.checkpoint
    nsteps = args.steps if args.steps is not None else 200

    def forward(q0_, u_, v_, w_):
        q = q0_.clone()
        adv.advance(code, nsteps, dt, u_, v_, w_, [q])
        return q

    def J(*a):
        q = forward(*a)
        return float((TOP * q * q).sum())

    def gradient():
        qT = forward(Q0, U, V, W)
        qb = (2.0 * TOP * qT).contiguous()
        ub, vb, wb = (torch.zeros_like(Q0) for _ in range(3))
        adv.adjoint_advance(code, nsteps, dt, U, V, W, Q0, qb, ub, vb, wb, checkpoint=ck)
        return qb, ub, vb, wb


    for _ in range(max(1, args.warmup // 5)):
        forward(Q0, U, V, W)
        gradient()
    torch.cuda.synchronize()
    ts = time.perf_counter()
    forward(Q0, U, V, W)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - ts
    # the forward that records its checkpoints (what a caller that needs the
    # gradient runs), then the adjoint from those checkpoints: per segment a
    # replay of its steps into stage storage and the backward sweep
    nseg = -(-nsteps // ck) if 0 < ck < nsteps else 1
    cks = [torch.empty_like(Q0) for _ in range(nseg)]
    qT = Q0.clone()
    torch.cuda.synchronize()
    ts = time.perf_counter()
    adv.advance_record(code, nsteps, dt, U, V, W, qT, ck, cks)
    torch.cuda.synchronize()
    t_rec = time.perf_counter() - ts
    qb = (2.0 * TOP * qT).contiguous()
    ub, vb, wb = (torch.zeros_like(Q0) for _ in range(3))
    torch.cuda.synchronize()
    ts = time.perf_counter()
    adv.adjoint_advance_recorded(code, nsteps, dt, U, V, W, ck, cks, qb, ub, vb, wb)
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - ts
    grads = (qb, ub, vb, wb)
    # the self-contained variant (its own forward sweep for the checkpoints)
    qb2 = (2.0 * TOP * qT).contiguous()
    z3 = [torch.zeros_like(Q0) for _ in range(3)]
    torch.cuda.synchronize()
    ts = time.perf_counter()
    adv.adjoint_advance(code, nsteps, dt, U, V, W, Q0, qb2, *z3, checkpoint=ck)
    torch.cuda.synchronize()
    t_adj_self = time.perf_counter() - ts
    if not torch.equal(qb, qb2):
        raise RuntimeError("recorded and self-contained adjoints differ")
    # self-contained with the whole trajectory held in HBM (checkpoint 0: one
    # segment, every stage state of the run stored; 42 GB at 256x256x128 and
    # 200 SSP-RK3 steps): one forward replay and the backward sweep, no
    # checkpoint sweep
    qb3 = (2.0 * TOP * qT).contiguous()
    for z in z3:
        z.zero_()
    torch.cuda.synchronize()
    ts = time.perf_counter()
    adv.adjoint_advance(code, nsteps, dt, U, V, W, Q0, qb3, *z3, checkpoint=0)
    torch.cuda.synchronize()
    t_adj_full = time.perf_counter() - ts
    if not torch.equal(qb, qb3):
        raise RuntimeError("whole-trajectory and checkpointed adjoints differ")
    full_bytes = (4 + 3 * nsteps) * Q0.numel() * 8  # 3 adjoint arrays, q0, the stage states
    del qb2, qb3, z3
    # the stage-adjoint kernel alone, timed with HIP events on the library stream
    lam = (2.0 * TOP * qT).contiguous()
    o = [torch.zeros_like(Q0) for _ in range(4)]
    reps, kt = 20, 0.0
    for _ in range(reps):
        adv.adjoint_tendency(U, V, W, qT, lam, *o)
        kt += adv.last_kernel_ms()
    adj_kernel_s = kt * 1e-3 / reps
    ncell = maxx * max(maxy, 1) * nz
    # algorithmic bytes of one stage adjoint per cell: q, adjoint in, 3 winds,
    # adjoint out, the q0 accumulator read+write, 3 wind adjoints read+write
    adj_bytes = (8 + 8 + 24 + 8 + 16 + 48) * ncell
    achieved = adj_bytes / adj_kernel_s / 1e9
    q_grad_norm = float(torch.linalg.vector_norm(qb))
    # the same gradient for the smooth theta field: the exact derivative of
    # the scheme on the rough log-normal q grows without bound over a long
    # trajectory (limiter_weno and the eps-regularised weights have huge
    # local derivatives near small values), on theta it stays bounded
    TH = torch.from_numpy(fields[1]).to(dev)
    cth = [torch.empty_like(Q0) for _ in range(nseg)]
    thT = TH.clone()
    adv.advance_record(code, nsteps, dt, U, V, W, thT, ck, cth)
    tb = (2.0 * TOP * thT).contiguous()
    gu, gv, gw = (torch.zeros_like(Q0) for _ in range(3))
    adv.adjoint_advance_recorded(code, nsteps, dt, U, V, W, ck, cth, tb, gu, gv, gw)
    theta_grad_norm = float(torch.linalg.vector_norm(tb))
    del cth, thT, tb, gu, gv, gw
    nseg = -(-nsteps // ck) if 0 < ck < nsteps else 1
    kk = ck if 0 < ck < nsteps else nsteps
    out = {
        "metric": "adjoint/forward cost ratio",
        "value": t_adj / t_fwd,
        "unit": "x",
        "n_gpus": 1,
        "steps": nsteps,
        "warmup": args.warmup,
        "ms_per_step": t_adj * 1e3 / nsteps,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded turbulent wind, log-normal scalar)",
        "config": {"workload": f"trajectory adjoint {maxx}x{max(maxy, 1)}x{nz}, {scheme}, ssprk3, {nsteps} steps, "
                               f"checkpoint every {ck or nsteps} steps, J = sum(top half rho0 q(T)^2), "
                               f"gradient wrt (q0, u, v, w)",
                   "cells": ncell, "forward_s": t_fwd, "adjoint_s": t_adj, "dt_s": dt,
                   "forward_recording_s": t_rec, "adjoint_self_contained_s": t_adj_self,
                   "self_contained_ratio": t_adj_self / t_fwd,
                   "adjoint_whole_trajectory_s": t_adj_full,
                   "whole_trajectory_ratio": t_adj_full / t_fwd,
                   "whole_trajectory_bytes": full_bytes,
                   "device_arrays": 3 + nseg + 3 * kk},
        "roofline": {"bound": "hbm", "kernel": ("adjoint_kernel<32,8,weno5>" if scheme == "weno5" else
                                                "adjoint_kernel<64,8>") + " (one stage adjoint)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": HBM_TRAFFIC.get(("adjoint", size, scheme, "ssprk3", 1)),
                     "alg_bytes_per_launch": adj_bytes, "avg_launch_us": adj_kernel_s * 1e6},
        "gradient_norms": {"q0_lognormal": q_grad_norm, "q0_theta": theta_grad_norm},
        "cpu_baseline": None,
    }
    adv.close()  # releases the trajectory storage before the next workload
    return out


def multi(args):
    """Uneven independent subdomains: 48 periodic subdomains with seeded
    sizes (MAXX, MAXY in [48, 192], nz in [48, 96]), 4-16 FP32 species each
    (log-normal), wind RMS in [1, 4] m/s, scheme weno (weno5 with --scheme
    weno5) or muscl (MC) per subdomain, SSP-RK3, one launch per stage for the whole batch with tile
    shapes from the autotuner.  Also timed: the same batch as one batched
    launch per subdomain and stage."""
    import numpy as np
    import torch

    from fdd3260_amd import Advection, MultiDomain, _lib, cases, grid as gm

    steps = args.steps if args.steps is not None else 20
    ndom = args.ndom
    t0 = time.time()
    dev = torch.device("cuda:0")
    advs, vels, fields, dts, cells, nspec = [], [], [], [], [], []
    weno = "weno5" if args.scheme == "weno5" else "weno"
    for d in cases.config4_draw(seed=4, ndom=ndom, weno=weno):
        g = gm.stretched(d["maxx"], d["maxy"], d["nz"])
        u, v, w = cases.turbulent_wind(g, seed=d["wind_seed"], rms=d["rms"])
        dts.append(cases.cfl_dt(g, u, v, w, 0.5))
        vels.append(tuple(torch.from_numpy(a.astype(np.float32)).to(dev) for a in (u, v, w)))
        fl = []
        for sd, mu in zip(d["species_seeds"], d["mu"]):
            q = cases.lognormal(g, seed=sd, mu=mu, sigma=1.5)
            fl.append(torch.from_numpy(q.astype(np.float32)).to(dev))
        fields.append(fl)
        advs.append(Advection(g, scal_adv=d["scheme"], muscllimiter="MC"))
        cells.append(d["maxx"] * d["maxy"] * d["nz"])
        nspec.append(d["nspec"])
    torch.cuda.synchronize()
    log(f"multi case built in {time.time() - t0:.1f}s")
    code, nst = _lib.INT_SSPRK3, 3
    m = MultiDomain(advs, vels, fields)
    t_at = time.perf_counter()
    if not args.no_autotune:
        m.autotune(reps=2)
    t_at = time.perf_counter() - t_at
    tiles = m.tiles()
    if args.warmup > 0:
        m.advance(code, args.warmup, dts)
    torch.cuda.synchronize()
    ts = time.perf_counter()
    m.advance(code, steps, dts)
    torch.cuda.synchronize()
    wall = time.perf_counter() - ts
    kern_ms = m.last_ms()
    # the same work as one batched launch per subdomain and stage
    torch.cuda.synchronize()
    ts = time.perf_counter()
    for a, (u, v, w), fl, dt in zip(advs, vels, fields, dts):
        a.set_stream(torch.cuda.current_stream(dev).cuda_stream)
        a.batch_advance(code, steps, dt, u, v, w, fl)
    torch.cuda.synchronize()
    wall_seq = time.perf_counter() - ts
    upd = sum(c * n for c, n in zip(cells, nspec)) * nst * steps
    stage_bytes = sum(c * (n * 12 + 12) for c, n in zip(cells, nspec))  # q, q0, out per species + 3 velocities
    stage_s = kern_ms * 1e-3 / (nst * steps)
    achieved = stage_bytes / stage_s / 1e9
    shapes = {}
    for tx, ty, kc in tiles:
        shapes[f"{tx}x{ty}/k{kc}"] = shapes.get(f"{tx}x{ty}/k{kc}", 0) + 1
    out = {
        "metric": "cell*species-updates/s and % HBM roofline at 1 GPU",
        "value": upd / wall,
        "unit": "cell*species-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded subdomain sizes, turbulent winds, log-normal species)",
        "config": {"workload": f"{ndom} uneven periodic subdomains, MAXX/MAXY 48-192, nz 48-96, 4-16 FP32 species, "
                               f"{weno}/muscl, ssprk3, one launch per stage",
                   "cells": sum(cells), "species_cells": sum(c * n for c, n in zip(cells, nspec)),
                   "autotune_s": t_at, "tile_choices": shapes,
                   "per_subdomain_launches_updates_per_s": upd / wall_seq,
                   "single_launch_speedup": wall_seq / wall},
        "roofline": {"bound": "hbm", "kernel": "multi_kernel (one stage of all subdomains)", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "alg_bytes_per_stage": stage_bytes, "avg_stage_us": stage_s * 1e6},
        "cpu_baseline": None,
    }
    m.close()
    for a in advs:
        a.close()
    return out


def copy_probe(mib):
    """A device-to-device copy of `mib` MiB: reads and writes exactly that
    many bytes, the reference for the HBM byte counters."""
    import torch

    x = torch.empty(mib * 2 ** 20 // 4, dtype=torch.float32, device="cuda:0")
    x.fill_(1.0)
    torch.cuda.synchronize()
    y = x.clone()
    torch.cuda.synchronize()
    del x, y


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: 50 fused, 100 plain, "
                                                               "200 adjoint, 20 multi)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--scheme", default="weno", choices=["weno", "ppm", "muscl", "weno5", "quick", "lw"])
    ap.add_argument("--limiter", default="MC")
    ap.add_argument("--integrator", default=None, choices=["heun", "ssprk3"],
                    help="default: ssprk3 for fused, heun (the reference's) for plain")
    ap.add_argument("--size", default=None, help="MAXX,MAXY,NZ (default: 512,512,160 fused, 256,256,128 plain "
                                                 "and adjoint)")
    ap.add_argument("--cpu-steps", type=int, default=None,
                    help="steps of the CPU-baseline sample (0: skip; default 1 fused, 6 plain)")
    ap.add_argument("--no-subrecords", action="store_true",
                    help="fused workload: skip the sub-records (plain, adjoint, multi-domain, and the "
                         "WENO5 batch, adjoint and multi-domain)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: all host cores (at most 16)")
    ap.add_argument("--workload", default="fused", choices=["plain", "fused", "adjoint", "multi"])
    ap.add_argument("--ndom", type=int, default=48, help="multi workload: number of subdomains")
    ap.add_argument("--no-autotune", action="store_true", help="multi workload: cost-model tiles only")
    ap.add_argument("--checkpoint", type=int, default=20, help="adjoint workload: steps between checkpoints")
    ap.add_argument("--group32", type=int, default=0)
    ap.add_argument("--kchunk", type=int, default=0)
    ap.add_argument("--tile-y", type=int, default=0, help="fused workload: rows of a 64-wide column tile (2, 4, 8; "
                                                       "default: 32x8 tiles)")
    ap.add_argument("--tile-x", type=int, default=0, help="fused workload: columns of the tile (64, 16, or 32 with "
                                                       "--tile-y 8)")
    ap.add_argument("--blocks-per-cu", type=int, default=0,
                    help="fused workload: persistent workgroups per CU (0: the occupancy limit)")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-variants", action="store_true", help="plain workload: skip the scheme variants")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="after the run, copy this many MiB on the device (a known-size kernel that scales "
                         "hardware byte counters)")
    args = ap.parse_args()

    if args.copy_probe > 0:
        import atexit

        atexit.register(copy_probe, args.copy_probe)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the advection operator is built and measured on a single "
              "MI355X; runs over several GPUs are not part of this build.", file=sys.stderr)
        sys.exit(2)

    if args.workload == "adjoint":
        out = adjoint(args)
    elif args.workload == "multi":
        out = multi(args)
    elif args.workload == "plain":
        out = plain(args)
    else:
        out = fused(args)
        if not args.no_subrecords:
            # the other benchmark workloads, each measured the same way, with
            # their own default sizes and steps
            sub = {}
            # and the WENO5 variants of the batch, the uneven subdomains and
            # the adjoint (the scheme those workloads are specified with;
            # the reference scheme 'weno' is the headline)
            for name, fn, scheme in (("plain_fp64_256x256x128", plain, "weno"),
                                     ("adjoint_200_steps", adjoint, "weno"),
                                     ("multi_48_subdomains", multi, "weno"),
                                     ("fused_weno5_512x512x160", fused, "weno5"),
                                     ("adjoint_weno5_200_steps", adjoint, "weno5"),
                                     ("multi_weno5_48_subdomains", multi, "weno5")):
                sa = argparse.Namespace(**vars(args))
                sa.size = None
                sa.steps = None
                sa.integrator = None
                sa.scheme = scheme
                sa.cpu_steps = 0 if name != "plain_fp64_256x256x128" else None
                log(f"sub-record {name}")
                r = fn(sa)
                sub[name] = {k: r[k] for k in ("metric", "value", "unit", "steps", "ms_per_step", "dtype", "config",
                                                "roofline", "cpu_baseline") if k in r}
            out["sub_records"] = sub
    print(json.dumps(out), flush=True)


def plain(args):
    """The 3D plain per-field FP64 forward (256x256x128 by default): the
    log-normal scalar and theta, the reference 'weno' with Heun stages (the
    reference's integrator), one launch per field and stage, and the other
    schemes as variants."""
    import torch

    from fdd3260_amd import Advection, _lib

    size = args.size or "256,256,128"
    integ = args.integrator or "heun"
    steps = args.steps if args.steps is not None else 100
    maxx, maxy, nz = (int(s) for s in size.split(","))
    t0 = time.time()
    g, u, v, w, fields, dt = build_case(maxx, maxy, nz, seed=2, rms=2.0)
    log(f"case built in {time.time() - t0:.1f}s, dt = {dt:.4f}s")
    dev = torch.device("cuda:0")
    du, dv, dw = (torch.from_numpy(a).to(dev) for a in (u, v, w))
    dq = [torch.from_numpy(f).to(dev) for f in fields]
    adv = Advection(g, scal_adv=args.scheme, muscllimiter=args.limiter)
    code = _lib.INT_HEUN if integ == "heun" else _lib.INT_SSPRK3
    nst = 2 if integ == "heun" else 3

    if args.warmup > 0:
        adv.advance(code, args.warmup, dt, du, dv, dw, dq)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    adv.advance(code, steps, dt, du, dv, dw, dq)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    kern_ms = adv.last_kernel_ms()
    for f in dq:
        if not bool(torch.isfinite(f).all()):
            raise RuntimeError("non-finite state after the timed steps")

    ncell = maxx * max(maxy, 1) * nz
    nfield = len(fields)
    updates = ncell * nfield * nst * steps
    value = updates / wall
    launches = nfield * nst * steps
    avg_launch_s = kern_ms * 1e-3 / launches
    sb = stage_bytes(integ, ncell)
    alg_bytes = sum(sb) / len(sb)
    achieved = alg_bytes / avg_launch_s / 1e9

    # the other scheme variants of the same workload (and the reference scheme
    # with SSP-RK3), timed the same way over a short run
    variants = {}
    if not args.no_variants and args.scheme == "weno" and integ == "heun":
        vsteps = max(1, min(steps, 20))
        for name, scheme, lim, integ in (("weno/ssprk3", "weno", "MC", "ssprk3"), ("ppm/heun", "ppm", "MC", "heun"),
                                         ("muscl-MC/heun", "muscl", "MC", "heun"),
                                         ("weno5/ssprk3", "weno5", "MC", "ssprk3"), ("quick/heun", "quick", "MC", "heun"),
                                         ("lw/heun", "lw", "MC", "heun")):
            va = Advection(g, scal_adv=scheme, muscllimiter=lim)
            vq = [torch.from_numpy(f).to(dev) for f in fields]
            vcode = _lib.INT_HEUN if integ == "heun" else _lib.INT_SSPRK3
            vn = 2 if integ == "heun" else 3
            va.advance(vcode, 2, dt, du, dv, dw, vq)
            torch.cuda.synchronize()
            ts = time.perf_counter()
            va.advance(vcode, vsteps, dt, du, dv, dw, vq)
            torch.cuda.synchronize()
            vw = time.perf_counter() - ts
            vk = va.last_kernel_ms() * 1e-3 / (nfield * vn * vsteps)
            vb = sum(stage_bytes(integ, ncell)) / vn
            variants[name] = {"value": ncell * nfield * vn * vsteps / vw, "avg_launch_us": vk * 1e6,
                              "frac": vb / vk / 1e9 / HBM_PEAK_GBS, "steps": vsteps}
            va.close()
            del vq

    cpu = None
    cpu_steps = args.cpu_steps if args.cpu_steps is not None else 6
    if cpu_steps > 0:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"cpu baseline: {cpu_steps} steps with {threads} threads")
        rate, el = cpu_baseline(g, u, v, w, fields, dt, args.scheme, args.limiter, integ,
                                cpu_steps, threads)
        cpu = {
            "value": rate,
            "unit": "cell*species-updates/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{cpu_steps} {integ} steps of the same {maxx}x{max(maxy, 1)}x{nz} "
                      f"{nfield}-field workload, C restatement (gcc -O2, no FMA contraction, OpenMP over "
                      f"lines), {el:.1f}s",
        }

    out = {
        "metric": "cell*species-updates/s and % HBM roofline at 1 GPU",
        "value": value,
        "unit": "cell*species-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded turbulent wind, log-normal scalar, theta profile)",
        "config": {
            "workload": f"3D plain per-field forward {maxx}x{max(maxy, 1)}x{nz}, stretched dz 20-80 m, "
                        f"{args.scheme}{'/' + args.limiter if args.scheme == 'muscl' else ''}, "
                        f"{integ}, {nfield} fields, FP64",
            "cells": ncell,
            "fields": nfield,
            "stages_per_step": nst,
            "dt_s": dt,
            "variants": variants,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": {"weno": "weno_march_kernel<32,8>", "weno5": "weno5_march_kernel<16,16>",
                       "muscl": "line_march_kernel<2,32,8>", "quick": "line_march_kernel<4,32,8>",
                       "lw": "line_march_kernel<5,32,8>", "ppm": "scalar_fp64_kernel<1,64,4>"}[args.scheme],
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": HBM_TRAFFIC.get(("plain", size, args.scheme, integ, nfield)),
            "alg_bytes_per_launch": alg_bytes,
            "avg_launch_us": avg_launch_s * 1e6,
        },
        "cpu_baseline": cpu,
    }
    adv.close()
    return out




if __name__ == "__main__":
    main()
