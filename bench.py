#!/usr/bin/env python3
"""Benchmark of the MI355X compressible FV step (one JSON line on stdout).

Workload (the single-GPU configuration the metric is quoted on): 3-D periodic
256^3 ideal gas (gamma 1.4) on [0, 2pi]^3, Taylor-Green velocity at Mach 0.3,
p = 1/gamma, rho = 1 + 0.1 * sum of 8 seeded Fourier modes (SplitMix64 seed
0x5EED); plain method = WCNS (characteristic WENO5-Z) + HLLC everywhere
(FluxMode::Weno) with the positivity limiter, SSP-RK3, dt = 0.5 / CFL rate
recomputed on the device every step.

A "step" is one SSP-RK3 step (3 stages) over the whole grid with the state
resident in HBM.  value = cells * steps * 3 / seconds (cell-updates per second
per RK stage).  The roofline entry is that of the dominant kernel (the fused
axis pass), timed with HIP events on the solver's stream inside the timed
region.  The CPU baseline is the oracle (bit-faithful restatement of the
reference CPU solver) timed for one step of the same 256^3 workload on the
host cores.

The "amr" entry is the batched block-AMR workload: 3-D shock-bubble on
[0,2]x[0,1]x[0,1] (Mach 1.22 shock at x = 0.5 into rho = 1, p = 1; sphere of
radius 0.25 at (1, 0.5, 0.5) with rho = 0.138), 8^3-cell blocks, 3 levels,
wavelet threshold 1e-2, regrid every 4 steps, Hybrid flux mode, CFL 0.4,
outflow x / reflecting y, z boundaries (about 2,900 blocks).  value = leaf
cells * steps * 3 / seconds, regrids included, for the batched path (one
persistent launch per phase, graph-replayed steps) and for per-block launches
of the same kernels, whose states are checked bit-identical.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N = 256
CFL = 0.5
# Algorithmic FP64 work of the WCNS (Weno-mode) residual, counted from the
# reference formulas with sparse eigenvectors, div/sqrt counted as 1 FLOP:
# 3,500 FLOP per cell per RK stage in 3-D, i.e. 3,500/3 per cell per axis pass.
# tile shapes of the axis pass (T cells, P pencils, threads) by cfg id, as
# the library's table
TILE_SHAPES = [(32, 16, 640), (64, 8, 576), (128, 4, 576), (256, 2, 576), (27, 8, 256), (59, 4, 256), (59, 8, 512),
               (27, 16, 512), (123, 4, 512), (59, 16, 1024)]
FLOP_PER_CELL_STAGE = 3500.0
FLOP_PER_CELL_STAGE_CENTRAL = 1600.0  # KEEP-PE central flux on every face, 3-D
BYTES_PER_CELL_STAGE = 120.0  # read U^(s), read U^n, write U^(s+1) (5 x f64 each)
PEAK_FP64 = 78.6e12
PEAK_FP32 = 157.3e12
PEAK_HBM = 8.0e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=N)
    ap.add_argument("--fast", action="store_true",
                    help="FMA-contracted variant (differs from the reference arithmetic by rounding)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-repeats", type=int, default=4, help="CPU-baseline repeats (median)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default tile shapes (short profiled runs)")
    ap.add_argument("--tiles", type=str, default=None,
                    help="fixed Weno tile shapes per axis, e.g. 6,6,5 (indices into TILE_SHAPES; no autotuning)")
    ap.add_argument("--hybrid-tiles", type=str, default=None, help="fixed Hybrid tile shapes per axis")
    ap.add_argument("--no-f32", action="store_true", help="skip the FP32 lines (256^3 Euler, FiveEq probe, AMR)")
    ap.add_argument("--f32-tiles", type=str, default=None, help="fixed FP32 tile shapes per axis")
    ap.add_argument("--no-amr", action="store_true", help="skip the block-AMR workload")
    ap.add_argument("--amr-only", action="store_true", help="only the block-AMR workload (profiling)")
    ap.add_argument("--amr-steps", type=int, default=100)
    ap.add_argument("--no-amr-group", action="store_true", help="skip the uneven-block AMR group")
    ap.add_argument("--amr-group-only", action="store_true", help="only the uneven-block AMR group")
    ap.add_argument("--amr-group-steps", type=int, default=20)
    ap.add_argument("--amr-per-block-steps", type=int, default=3)
    ap.add_argument("--amr-no-graph", action="store_true",
                    help="batched AMR steps launched eagerly, not replayed as graphs (for counter collection)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint workload")
    ap.add_argument("--adjoint-only", action="store_true", help="only the adjoint workload (profiling)")
    ap.add_argument("--adj-n", type=int, default=512)
    ap.add_argument("--adj-tile", type=int, default=0, help="interior cells per adjoint tile at most (0: library default)")
    ap.add_argument("--adj-steps", type=int, default=400)
    ap.add_argument("--adj-checkpoint", type=int, default=20)
    ap.add_argument("--adj-fd-cells", type=int, default=64, help="central-difference checks (seeded cells)")
    ap.add_argument("--adj-fd-modes", type=int, default=16, help="central-difference checks along density modes")
    ap.add_argument("--adj-fd-sweep", type=int, default=16, help="cells also checked at h = 1e-5 and 1e-6 rho")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the ensemble workload")
    ap.add_argument("--no-fiveeq", action="store_true", help="skip the two-phase FiveEq workload")
    ap.add_argument("--no-hybrid", action="store_true", help="skip the Hybrid-mode 256^3 line")
    ap.add_argument("--fiveeq-only", action="store_true", help="only the FiveEq workload (profiling)")
    ap.add_argument("--fe-n", type=int, default=256)
    ap.add_argument("--fe-steps", type=int, default=10)
    ap.add_argument("--fe-tiles", type=str, default=None, help="fixed FiveEq tile shapes per axis (no autotuning)")
    ap.add_argument("--ensemble-only", action="store_true", help="only the ensemble workload (profiling)")
    ap.add_argument("--ens-n", type=int, default=256, help="instances")
    ap.add_argument("--ens-steps", type=int, default=300)
    ap.add_argument("--ens-checkpoint", type=int, default=20)
    ap.add_argument("--ens-fd", type=int, default=2, help="instances checked by central differences in Mach")
    return ap.parse_args()


def ensemble_bench(em, args):
    """256 2-D shock-bubble instances (seeded Mach U[1.1,2], radius U[0.1,0.3],
    density ratio U[0.1,0.5], side uniform in {96, 128, ..., 384}) advanced
    together; forward with checkpoints and the adjoint of the bubble-region
    mean pressure with respect to Mach."""
    rng = np.random.default_rng(0x5EED + 2)
    bc = [["outflow", "outflow"], ["outflow", "outflow"], ["periodic", "periodic"]]
    inst = []
    for _ in range(args.ens_n):
        n = int(rng.choice(np.arange(96, 385, 32)))
        mach, r, dens = rng.uniform(1.1, 2.0), rng.uniform(0.1, 0.3), rng.uniform(0.1, 0.5)
        u, du = em.shock_bubble_2d(n, mach, r, dens)
        inst.append((n, mach, r, dens, u, du))
    cells = sum(n * n for n, *_ in inst)
    sizes = [(n, n, 1) for n, *_ in inst]
    states = [x[4] for x in inst]

    def bubble(u, n, r):
        x = (np.arange(n) + 0.5) / n
        X, Y = np.meshgrid(x, x, indexing="ij")
        m = (X - 0.5) ** 2 + (Y - 0.5) ** 2 < r * r
        q = u[6:6 + n, 6:6 + n, 0]
        vel = q[..., 1:4] / q[..., :1]
        ke = 0.5 * np.sum(vel * vel, -1)
        p = 0.4 * (q[..., 4] - q[..., 0] * ke)
        g = np.zeros_like(u)
        gi = g[6:6 + n, 6:6 + n, 0]
        gi[..., 0] = 0.4 * ke
        gi[..., 1:4] = -0.4 * vel
        gi[..., 4] = 0.4
        gi *= (m / m.sum())[..., None]
        return float(p[m].mean()), g

    # plain batched stepping vs one launch per instance (same kernels, the
    # tile shape the autotuner picks for the batched launch)
    timing = {}
    tiles = None
    for batched in (True, False):
        ens = em.HipEnsemble(sizes, states, bc=bc, batched=batched)
        if tiles is None:
            tiles = ens.autotune(3)
        else:
            ens.set_tiles(tiles)
        ens.run_steps(2)
        k = 20 if batched else 5
        t0 = time.perf_counter()
        ens.run_steps(k)
        timing[batched] = (time.perf_counter() - t0) / k
        del ens
    ens = em.HipEnsemble(sizes, states, bc=bc)
    ens.set_tiles(tiles)

    def objective():
        djs = []
        for i, (n, _, r, *_) in enumerate(inst):
            djs.append(bubble(ens.download(i), n, r)[1])
        return djs

    # warm-up pair of the same length: the timed pair allocates nothing
    ens.forward(args.ens_steps, args.ens_checkpoint)
    ens.backward(objective())
    for i, u in enumerate(states):
        ens.upload(i, u)
    t0 = time.perf_counter()
    ens.forward(args.ens_steps, args.ens_checkpoint)
    t_fwd = time.perf_counter() - t0
    djs = objective()
    t0 = time.perf_counter()
    du0 = ens.backward(djs)
    t_bwd = time.perf_counter() - t0
    grads = [float(np.sum(du0[i] * x[5])) for i, x in enumerate(inst)]
    errs = []
    for i in range(min(args.ens_fd, len(inst))):
        n, mach, r, dens, _, _ = inst[i]
        dts = ens.dts(i)
        h = 1e-6
        jj = []
        for sg in (1.0, -1.0):
            up, _ = em.shock_bubble_2d(n, mach + sg * h, r, dens)
            g1 = em.HipSolver((n, n, 1), (0, 0, 0), (1, 1, 1), up, bc=bc, mode="weno", cfl=0.5)
            g1.forward(args.ens_steps, args.ens_steps, dts)
            jj.append(bubble(g1.download(), n, r)[0])
        fd = (jj[0] - jj[1]) / (2 * h)
        errs.append(abs(fd - grads[i]) / max(abs(grads[i]), 1e-12))
    return {
        "workload": f"{args.ens_n} 2-D shock-bubble instances, sides 96..384, Weno, CFL 0.5, {args.ens_steps} steps, "
                    "J = bubble-region mean p, dJ/dMach",
        "cells": cells,
        "value": cells * 3 / timing[True],
        "unit": "cell-updates/s per RK stage",
        "tiles": tiles,
        "per_instance_launches": {"value": cells * 3 / timing[False]},
        "speedup_vs_per_instance": timing[False] / timing[True],
        "forward_s": t_fwd,
        "backward_s": t_bwd,
        "cost_ratio": t_bwd / t_fwd,
        "checkpoint_every": args.ens_checkpoint,
        "dJ_dMach_first": grads[:4],
        "fd": {"instances": len(errs), "h": 1e-6, "max_rel_err": max(errs) if errs else None},
    }


def _splitmix64(state):
    state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return state, z ^ (z >> 31)


def adjoint_modes(seed=0x5EED + 1):
    """The 16 seeded Fourier modes (k, phase) of the adjoint workload (|k| <= 4,
    SplitMix64)."""
    s = seed
    modes = []
    while len(modes) < 16:
        s, a = _splitmix64(s)
        s, b = _splitmix64(s)
        k = (int(a % 9) - 4, int(b % 9) - 4)
        if 0 < k[0] ** 2 + k[1] ** 2 <= 16:
            s, c = _splitmix64(s)
            modes.append((k, 2 * np.pi * (c >> 11) * 2.0 ** -53))
    return modes


def adjoint_case(n, seed=0x5EED + 1):
    """2-D four-quadrant Riemann problem (split (0.8, 0.8), outflow) on [0,1]^2
    with rho scaled by 1 + 0.01 * sum of 16 seeded Fourier modes (|k| <= 4,
    SplitMix64); CPU layout with 6 ghost layers."""
    modes = adjoint_modes(seed)
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    pert = sum(np.cos(2 * np.pi * (k[0] * X + k[1] * Y) + ph) for k, ph in modes)
    e, nn = X > 0.8, Y > 0.8
    quad = [(1.5, 0.0, 0.0, 1.5), (0.5323, 1.206, 0.0, 0.3), (0.138, 1.206, 1.206, 0.029), (0.5323, 0.0, 1.206, 0.3)]
    sel = np.where(e & nn, 0, np.where(nn, 1, np.where(~e, 2, 3)))
    w = np.array(quad)[sel]
    rho = w[..., 0] * (1.0 + 0.01 * pert)
    u = np.zeros((n + 12, n + 12, 1, 5))
    ui = u[6:6 + n, 6:6 + n, 0]
    ui[..., 0] = rho
    ui[..., 1] = rho * w[..., 1]
    ui[..., 2] = rho * w[..., 2]
    ui[..., 4] = w[..., 3] / 0.4 + 0.5 * rho * (w[..., 1] ** 2 + w[..., 2] ** 2)
    return u


def pressure_window(u, n, lo=0.6, hi=0.8):
    """J = sum of p dx dy over cells with centre in [lo, hi]^2, and dJ/du."""
    x = (np.arange(n) + 0.5) / n
    m = (x >= lo) & (x <= hi)
    mask = m[:, None] & m[None, :]
    q = u[6:6 + n, 6:6 + n, 0]
    vel = q[..., 1:4] / q[..., :1]
    ke = 0.5 * np.sum(vel * vel, axis=-1)
    p = 0.4 * (q[..., 4] - q[..., 0] * ke)
    dxdy = 1.0 / (n * n)
    J = float(np.sum(p[mask]) * dxdy)
    g = np.zeros_like(u)
    gi = g[6:6 + n, 6:6 + n, 0]
    gi[..., 0] = 0.4 * ke
    gi[..., 1:4] = -0.4 * vel
    gi[..., 4] = 0.4
    gi *= (mask * dxdy)[..., None]
    return J, g


def adjoint_bench(em, args):
    """Adjoint/forward cost of dJ/drho_0 (J = integral of p over [0.6,0.8]^2 at
    the last step) through args.adj_steps SSP-RK3 steps with checkpoints, and
    central finite differences of the forward run at seeded cells."""
    n = args.adj_n
    u0 = adjoint_case(n)
    bc = [["outflow", "outflow"], ["outflow", "outflow"], ["periodic", "periodic"]]
    mk = lambda u: em.HipSolver((n, n, 1), (0, 0, 0), (1, 1, 1), u, bc=bc, mode="weno", cfl=0.5)
    g = mk(u0)
    if args.adj_tile:
        g.set_adjoint_tile(args.adj_tile)
    # warm-up: one full forward/backward pair of the same length, so that the
    # timed pair allocates nothing (checkpoints, dt history, segment stores,
    # adjoint work arrays) and replays already-instantiated graphs.  Both
    # timed runs take and leave their data in device memory: the backward's
    # cotangents are device arrays (em_solver_backward_device), as the
    # forward's state is, so no host layout conversion is timed.
    g.forward(args.adj_steps, args.adj_checkpoint)
    J, dj = pressure_window(g.download(), n)
    dj_d = em.DeviceArray.from_host(g.to_soa(dj))
    out_d = em.DeviceArray(dj_d.n)
    g.backward_device(dj_d, out_d)
    g.upload(u0)
    g.synchronize()
    t0 = time.perf_counter()
    dts = g.forward(args.adj_steps, args.adj_checkpoint)  # CFL dt computed and recorded on the device
    g.synchronize()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    g.backward_device(dj_d, out_d)
    t_bwd = time.perf_counter() - t0
    du0 = g.from_soa(out_d.to_host())
    host_api_same = bool(np.array_equal(g.backward(dj), du0))  # the host-array entry point, untimed
    # the same pair with every step's u_n and stage states kept on the device
    # (checkpoint_every = 0: 120 B per cell and step), so that the backward
    # recomputes nothing; warm-up pair first, as above
    g.upload(u0)
    g.forward(args.adj_steps, 0)
    g.backward_device(dj_d, out_d)
    g.upload(u0)
    g.synchronize()
    t0 = time.perf_counter()
    dts_k = g.forward(args.adj_steps, 0)
    g.synchronize()
    t_fwd_k = time.perf_counter() - t0
    t0 = time.perf_counter()
    g.backward_device(dj_d, out_d)
    t_bwd_k = time.perf_counter() - t0
    du0_k = g.from_soa(out_d.to_host())
    kept = {"forward_s": t_fwd_k, "backward_s": t_bwd_k, "cost_ratio": t_bwd_k / t_fwd_k,
            "device_bytes": 120 * n * n * args.adj_steps,
            "gradient_identical_to_checkpointed": bool(np.array_equal(du0_k, du0) and np.array_equal(dts_k, dts))}
    q0 = u0[6:6 + n, 6:6 + n, 0]
    vel = q0[..., 1:4] / q0[..., :1]
    dirn = np.concatenate([np.ones_like(q0[..., :1]), vel, 0.5 * np.sum(vel * vel, axis=-1, keepdims=True)], -1)
    grad_rho = np.sum(du0[6:6 + n, 6:6 + n, 0] * dirn, axis=-1)
    scale = float(np.abs(grad_rho).max())
    f = mk(u0)

    def jfd(du, h):
        js = []
        for sg in (1.0, -1.0):
            f.upload(u0 + sg * h * du)
            f.forward(args.adj_steps, args.adj_steps, dts)
            js.append(pressure_window(f.download(), n)[0])
        return (js[0] - js[1]) / (2 * h)

    # (1) per cell, h = 1e-7 rho: seeded cells drawn from the gradient's
    # support (|dJ/drho_0| >= 1e-2 max; elsewhere the adjoint is ~0 and a
    # difference quotient is rounding noise of J, ~1e-9 at this h)
    rng = np.random.default_rng(0x5EED)
    support = np.argwhere(np.abs(grad_rho) >= 1e-2 * scale)
    cells = support[rng.choice(len(support), size=min(args.adj_fd_cells, len(support)), replace=False)]
    errs, best = [], []
    for q, (i, j) in enumerate(cells):
        du = np.zeros_like(u0)
        du[6 + i, 6 + j, 0] = dirn[i, j]
        ad = float(grad_rho[i, j])
        fd = jfd(du, 1e-7 * q0[i, j, 0])
        errs.append(abs(fd - ad) / abs(ad))
        if q < args.adj_fd_sweep:
            # the same cell at h = 1e-5 and 1e-6 rho: truncation falls like h^2
            # and the rounding of J over h grows like 1/h, so the best h
            # shows the adjoint's agreement (tests/test_oracle_adjoint.py
            # holds the full step-size study on a reduced case)
            e = [errs[-1]] + [abs(jfd(du, h * q0[i, j, 0]) - ad) / abs(ad) for h in (1e-5, 1e-6)]
            best.append(min(e))
    # (2) along each seeded density mode (at fixed velocity and pressure):
    # dJ/da_m for rho = rho_q (1 + 0.01 sum_m a_m cos(k_m . x + phi_m)),
    # a sum over the whole support, with h = 1e-5 in a_m
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    rho_q = q0[..., 0] / (1.0 + 0.01 * sum(np.cos(2 * np.pi * (k[0] * X + k[1] * Y) + ph)
                                             for k, ph in adjoint_modes()))
    merrs, mgrads = [], []
    for k, ph in adjoint_modes()[:args.adj_fd_modes]:
        du = np.zeros_like(u0)
        du[6:6 + n, 6:6 + n, 0] = (0.01 * rho_q * np.cos(2 * np.pi * (k[0] * X + k[1] * Y) + ph))[..., None] * dirn
        ad = float(np.sum(du0 * du))
        mgrads.append(ad)
        merrs.append((ad, jfd(du, 1e-5)))
    mscale = max((abs(a) for a, _ in merrs), default=0.0)
    merr = [abs(fd - a) / max(abs(a), 1e-2 * mscale) for a, fd in merrs]
    # normwise: the largest |FD - adjoint| over the largest |dJ/da|; the FD
    # error at h = 1e-5 is ~1e-11 absolute on every mode (rounding below this
    # h, kinks of the limiter and WENO branches above it)
    mnorm = max((abs(fd - a) for a, fd in merrs), default=0.0) / mscale if mscale > 0 else None
    return {
        "workload": f"2-D {n}^2 four-quadrant Riemann + 0.01 x 16 seeded Fourier modes on rho, Weno, CFL 0.5, "
                    f"{args.adj_steps} steps, J = integral of p over [0.6,0.8]^2, dJ/drho_0",
        "forward_s": t_fwd,
        "backward_s": t_bwd,
        "cost_ratio": t_bwd / t_fwd,
        "checkpoint_every": args.adj_checkpoint,
        "backward_entry_point": "em_solver_backward_device (cotangents in device memory)",
        "host_array_entry_point_identical": host_api_same,
        "kept_states": kept,
        "J": J,
        "grad_max_abs": scale,
        "fd_cells": {"cells": len(errs), "h": "1e-7 rho", "drawn_from": "|dJ/drho_0| >= 1e-2 max",
                     "max_rel_err": max(errs) if errs else None,
                     "median_rel_err": float(np.median(errs)) if errs else None,
                     "best_of_h": {"cells": len(best), "h": "1e-5, 1e-6, 1e-7 rho",
                                   "max_rel_err": max(best) if best else None,
                                   "median_rel_err": float(np.median(best)) if best else None}},
        "fd_modes": {"modes": len(merr), "h": 1e-5, "dJ_da": mgrads, "normwise_rel_err": mnorm,
                     "median_rel_err": float(np.median(merr)) if merr else None,
                     "max_rel_err": max(merr) if merr else None, "rel_err_floor": "1e-2 x max |dJ/da|"},
    }


AMR_BC = [["outflow", "outflow"], ["reflect", "reflect"], ["reflect", "reflect"]]
AMR_CASE = dict(base=(8, 4, 4), b=8, levels=3, lo=(0.0, 0.0, 0.0), hi=(2.0, 1.0, 1.0),
                params=[1.22, 0.5, 1.0, 0.5, 0.5, 0.25, 0.138], eps=1e-2, regrid_every=4, mode="hybrid", cfl=0.4)


AMR_GROUP = [dict(b=8, base=(8, 4, 4)), dict(b=16, base=(4, 2, 2)), dict(b=32, base=(2, 1, 1))]


def amr_group_bench(em, args):
    """Hierarchies of uneven block size in one batched step: the 3-D
    shock-bubble as three hierarchies with 8^3, 16^3 and 32^3 blocks (3
    levels each), stepped as a group (one flux-pass launch per axis and
    stage over the leaves of all three) and, for comparison, one after
    another with their own batched, graph-replayed steps; the states must
    agree bit for bit."""
    c = AMR_CASE
    kw = dict(eps=c["eps"], regrid_every=c["regrid_every"], bc=AMR_BC, mode=c["mode"], cfl=c["cfl"])
    hiers = []
    for m in AMR_GROUP:
        root = em.HipAmr(m["base"], m["b"], c["levels"], c["lo"], c["hi"], case="shock_bubble", params=c["params"],
                         **kw)
        hiers.append((m, root.blocks() + (root.download(),)))
        del root
    make = [lambda m=m, h=h: em.HipAmr(m["base"], m["b"], c["levels"], c["lo"], c["hi"], blocks=h, **kw)
            for m, h in hiers]
    steps = args.amr_group_steps
    solo = [f() for f in make]
    for s in solo:
        s.autotune(3)
        s.run_steps(1)  # graph capture outside the timed region
    cells = sum(s.leaf_count() * s.b ** 3 for s in solo)
    el_solo = 0.0
    for s in solo:
        s.synchronize()
        t0 = time.perf_counter()
        s.run_steps(steps)
        s.synchronize()
        el_solo += time.perf_counter() - t0
    grp_members = [f() for f in make]
    grp = em.HipAmrGroup(grp_members)
    grp.run_steps(1)
    for g in grp_members:
        g.synchronize()
    t0 = time.perf_counter()
    grp.run_steps(steps)
    for g in grp_members:
        g.synchronize()
    el = time.perf_counter() - t0
    same = all(np.array_equal(a.download(), b.download()) for a, b in zip(grp_members, solo))
    return {
        "workload": "3-D shock-bubble as three hierarchies with 8^3, 16^3 and 32^3 blocks, 3 levels, Hybrid, CFL 0.4, "
                    "regrid every 4; one flux-pass launch per axis and stage over the leaves of all three",
        "members": [{"b": s.b, "leaves": s.leaf_count(), "cells": s.leaf_count() * s.b ** 3} for s in solo],
        "value": cells * steps * 3 / el,
        "unit": "cell-updates/s per RK stage",
        "ms_per_step": el / steps * 1e3,
        "steps": steps,
        "one_after_another": {"value": cells * steps * 3 / el_solo, "ms_per_step": el_solo / steps * 1e3},
        "identical_to_members_alone": bool(same),
    }


def amr_bench(em, args):
    """Batched vs per-block block-AMR stepping on the 3-D shock-bubble."""
    c = AMR_CASE
    kw = dict(eps=c["eps"], regrid_every=c["regrid_every"], bc=AMR_BC, mode=c["mode"], cfl=c["cfl"])
    root = em.HipAmr(c["base"], c["b"], c["levels"], c["lo"], c["hi"], case="shock_bubble", params=c["params"], **kw)
    hier = root.blocks() + (root.download(),)
    nblocks, nleaves = root.counts()
    del root
    bat = em.HipAmr(c["base"], c["b"], c["levels"], c["lo"], c["hi"], blocks=hier, **kw)
    if args.amr_no_graph:
        bat.set_graph(False)
    shapes = bat.autotune(3)
    k = args.amr_per_block_steps
    same, el_per, cells_per = None, None, None
    if k > 0:  # 0: no per-block comparison (counter runs)
        per = em.HipAmr(c["base"], c["b"], c["levels"], c["lo"], c["hi"], blocks=hier, batched=False, **kw)
        per.set_shapes(shapes)
        bat.run_steps(k)
        per.synchronize()
        t0 = time.perf_counter()
        per.run_steps(k)
        per.synchronize()
        el_per = time.perf_counter() - t0
        same = bool(np.array_equal(bat.download(), per.download()))
        cells_per = per.leaf_count() * c["b"] ** 3
        del per
    bat.synchronize()
    leaves0 = bat.leaf_count()
    c0 = bat.counters()[0]
    t0 = time.perf_counter()
    bat.run_steps(args.amr_steps)
    bat.synchronize()
    el = time.perf_counter() - t0
    leaves1 = bat.leaf_count()
    b = c["b"]
    faces = 0.5 * (leaves0 + leaves1) * 3 * (b + 1) * b * b * 3 * args.amr_steps
    share = (bat.counters()[0] - c0) / faces
    cells = 0.5 * (leaves0 + leaves1) * c["b"] ** 3
    bat.set_profile(True)
    bat.run_steps(16)
    phase = [v / 16 for v in bat.profile()]
    value = cells * args.amr_steps * 3 / el
    value_per = cells_per * k * 3 / el_per if k > 0 else None
    # algorithmic FLOPs of the flux phase (shock-path share at the Weno
    # figure, the rest central) over its measured time per step
    flop = share * FLOP_PER_CELL_STAGE + (1.0 - share) * FLOP_PER_CELL_STAGE_CENTRAL
    achieved = cells * 3 * flop / (phase[2] * 1e-3)
    cpu = None if args.no_cpu_baseline else amr_cpu_baseline(nleaves * c["b"] ** 3, args.cpu_repeats)
    f32 = None
    if not args.no_f32:
        # the same hierarchy and steps on the FP32 pool (the reference's
        # GpuAmr<f32>, the precision of its own 3-D bubble runs)
        leaves_f64 = bat.leaf_count()
        del bat
        g32 = em.HipAmr(c["base"], c["b"], c["levels"], c["lo"], c["hi"], blocks=hier, real="f32", **kw)
        shapes32 = g32.autotune(3)
        g32.synchronize()
        l0 = g32.leaf_count()
        t0 = time.perf_counter()
        g32.run_steps(args.amr_steps)
        g32.synchronize()
        el32 = time.perf_counter() - t0
        l1 = g32.leaf_count()
        f32 = {
            "value": 0.5 * (l0 + l1) * c["b"] ** 3 * args.amr_steps * 3 / el32,
            "unit": "cell-updates/s per RK stage",
            "ms_per_step": el32 / args.amr_steps * 1e3,
            "pass_shapes": shapes32,
            "leaves_after": l1,
            "leaves_after_f64": leaves_f64,
        }
        del g32
    return {
        "workload": "3-D shock-bubble block AMR, 8^3 blocks, 3 levels, Hybrid, CFL 0.4, regrid every 4",
        "blocks": nblocks,
        "leaves": nleaves,
        "value": value,
        "unit": "cell-updates/s per RK stage",
        "ms_per_step": el / args.amr_steps * 1e3,
        "steps": args.amr_steps,
        "per_block": {"value": value_per, "ms_per_step": el_per / k * 1e3, "steps": k} if k > 0 else None,
        "pass_shapes": shapes,
        "speedup_vs_per_block": value / value_per if k > 0 else None,
        "identical_to_per_block": same,
        "phase_ms_per_step": dict(zip(["restrict", "ghosts", "fluxes", "reflux_update", "regrid"], phase)),
        "shock_path_face_share": share,
        "roofline": {"bound": "mfma", "roof": "fp64-valu", "kernel": "amr_pass (flux phase, 3 axes)",
                     "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64, "flop_per_cell_stage": flop},
        "cpu_baseline": cpu,
        "f32": f32,
    }


def amr_cpu_baseline(cells, repeats):
    """Oracle block AMR (restated reference CPU Amr) on the host cores: one
    global SSP-RK3 step of the same hierarchy."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle  # noqa: E402  (CPU-baseline leg only)

    c = AMR_CASE
    threads = _oracle.load().emref_max_threads()
    runs = []
    for _ in range(repeats):
        o = _oracle.OracleAmr("shock_bubble", c["params"], c["base"], c["b"], c["levels"], c["lo"], c["hi"],
                              eps=c["eps"], regrid_every=c["regrid_every"], bc=AMR_BC, mode=c["mode"], cfl=c["cfl"])
        t0 = time.perf_counter()
        o.step(o.dt_cfl())
        runs.append(time.perf_counter() - t0)
    el = float(np.median(runs))
    rec = {"value": cells * 3 / el, "unit": "cell-updates/s per RK stage", "cores": int(threads), "kind": "port",
           "sample": f"1 global SSP-RK3 step of the same hierarchy, OpenMP over leaves, median of {repeats} runs "
                     f"({el:.2f} s each)", "repeats_s": runs}
    rec.update(host_info())
    return rec


def profiled_traffic(variant, mode, d, shape):
    """HBM bytes (read + write) per launch of the axis pass along d in flux
    mode `mode` (0 Hybrid, 1 Weno) with tile shape (T, P, threads), from the
    committed rocprofv3 summary (128/64/32-byte read requests and WRITE_SIZE);
    None when the summary has no row for this kernel."""
    path = os.path.join(ROOT, "profiles", "axis_pass_256cubed.md")
    if not os.path.exists(path):
        return None
    sig = f"`em::{variant}::axis_pass<{d}, {mode}, {shape[0]}, {shape[1]}, {shape[2]}>`"
    for line in open(path):
        cells = [c.strip() for c in line.split("|")]
        # | kernel | read requests | read | write | read + write | algorithmic | ratio |
        if len(cells) >= 8 and cells[1] == sig and cells[5].endswith(" GB"):
            return float(cells[5].split()[0]) * 1e9
    return None


def host_info():
    """CPU model and core counts of the host the CPU baseline runs on."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def cpu_baseline(em, n, u, repeats):
    """Oracle (restated reference CPU solver, bit-identical to it) on the
    host cores, by the reference's cpu-bench protocol (median of alternating
    repeats, 1 thread and all threads, hash check):
    * headline: one SSP-RK3 step (dt_cfl + 3 stages) of the same n^3 Weno
      workload on all threads, median of `repeats` runs;
    * 1 thread vs all threads, alternating, on a 64^3 sample of the same
      workload (same initial-condition construction; the per-cell work of a
      periodic 3-D Weno step does not depend on the size);
    * the 64^3 sample's state after its step must hash equal to the GPU's
      (precise variant) after the same step."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle  # noqa: E402  (CPU-baseline leg only)

    lib = _oracle.load()
    threads = int(lib.emref_max_threads())
    hi = (2 * np.pi,) * 3

    def one_step(nn, uu, nthreads):
        lib.emref_set_threads(nthreads)
        o = _oracle.OracleSolver((nn, nn, nn), (0, 0, 0), hi)
        o.set_u(uu)
        o.set_opts(mode="weno", cfl=CFL)
        t0 = time.perf_counter()
        dt = o.dt_cfl()
        o.step(dt)
        el = time.perf_counter() - t0
        return el, o, dt

    runs = [one_step(n, u, threads)[0] for _ in range(repeats)]
    m = 64
    us = em.case_init(1, (m, m, m), (0, 0, 0), hi, 6, [0.3, float(0x5EED)])
    t1, tall = [], []
    for _ in range(repeats):  # alternating
        t1.append(one_step(m, us, 1)[0])
        el, o, dt = one_step(m, us, threads)
        tall.append(el)
    lib.emref_set_threads(threads)
    g = em.HipSolver((m, m, m), (0, 0, 0), hi, us, mode="weno", cfl=CFL, precise=True)
    g.step(dt)
    same = bool(np.array_equal(g.interior(g.download()), o.interior(o.get_u())))
    med = float(np.median(runs))
    rec = {
        "value": n ** 3 * 3 / med,
        "unit": "cell-updates/s per RK stage",
        "cores": threads,
        "kind": "port",
        "sample": f"1 SSP-RK3 step (dt_cfl + 3 stages) of the same {n}^3 Weno workload, OpenMP over pencils, "
                  f"median of {repeats} runs ({med:.2f} s each)",
        "repeats_s": runs,
        "threads_1_vs_all": {
            "sample": f"1 SSP-RK3 step of the {m}^3 version of the workload, {repeats} alternating repeats",
            "value_1_thread": m ** 3 * 3 / float(np.median(t1)),
            "value_all_threads": m ** 3 * 3 / float(np.median(tall)),
            "threads": threads,
        },
        "state_identical_to_gpu": same,
    }
    rec.update(host_info())
    return rec


def fiveeq_bench(em, args):
    """The reference's `interface` probe at n^3 (examples/cpu_validate.rs:100-
    146): a water drop (stiffened gas) in air moving at 0.1 through p = 1e5,
    Kapila five-equation model, Hybrid mode, 3-D periodic; fixed dt = half the
    CFL dt of the initial state, steps timed after a warm-up step.  Then the
    same steps on the FP32 variant (the reference's GpuSolver<f32> with
    GpuSystem::FiveEq) at the same dt, with the relative L1 difference of its
    mixture density from the FP64 run's."""
    n = args.fe_n
    x = (np.arange(n) + 0.5) / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    r = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2)
    a = np.clip(0.5 * (1.0 + np.tanh((0.25 - r) * n)), 1e-6, 1.0 - 1e-6)
    (g1, pi1, _, _), (g2, pi2, _, _) = em.fiveeq.WATER_AIR
    a1, a2, vx = a * 1000.0, (1.0 - a) * 1.2, 0.1
    rho = a1 + a2
    e = 1e5 * (a / (g1 - 1.0) + (1.0 - a) / (g2 - 1.0)) + g1 * pi1 * a / (g1 - 1.0) + 0.5 * rho * vx * vx
    u = np.zeros((n + 12, n + 12, n + 12, 7))
    u[6:-6, 6:-6, 6:-6] = np.stack([a1, a2, rho * vx, 0 * a, 0 * a, e, a], axis=-1)
    del X, Y, Z, r, a, a1, a2, rho, e

    def run(real, dt=None):
        g = em.HipFiveEq((n, n, n), (0, 0, 0), (1, 1, 1), u, mode="hybrid", real=real)
        tiles = pick_tiles(g, args, args.fe_tiles)
        dt = 0.5 * g.dt_cfl() if dt is None else dt
        g.step(dt)
        g.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.fe_steps):
            g.step(dt)
        g.synchronize()
        return g, time.perf_counter() - t0, tiles, dt

    g, el, tiles, dt = run("f64")
    shock = g.counters()[0] / (3 * (args.fe_steps + 1) * 3 * (n + 1) * n * n)
    rec = {
        "workload": f"3-D periodic {n}^3 water drop in air, Kapila five-equation model (7 variables), Hybrid, "
                    "the reference's interface probe at this size",
        "value": n ** 3 * args.fe_steps * 3 / el,
        "unit": "cell-updates/s per RK stage",
        "ms_per_step": el / args.fe_steps * 1e3,
        "steps": args.fe_steps,
        "tiles": tiles,
        "dt": dt,
        "shock_path_face_fraction": shock,
    }
    if not args.no_f32:
        rho64 = g.interior(g.download())[..., :2].sum(axis=-1)
        g.close()
        g32, el32, tiles32, _ = run("f32", dt)
        rho32 = g32.interior(g32.download())[..., :2].sum(axis=-1)
        g32.close()
        rec["f32"] = {
            "value": n ** 3 * args.fe_steps * 3 / el32,
            "unit": "cell-updates/s per RK stage",
            "ms_per_step": el32 / args.fe_steps * 1e3,
            "tiles": tiles32,
            "density_rel_l1_vs_f64": float(np.abs(rho32 - rho64).sum() / np.abs(rho64).sum()),
        }
    return rec


def f32_bench(em, args, n, u):
    """The same 256^3 Weno workload on the FP32 variant of the solver (the
    reference's GpuSolver<f32>: state and arithmetic in FP32, multiply-adds
    contracted), with its difference from the FP64 state after the same
    number of steps.  FP32 VALU peak: 157.3 TFLOP/s."""
    lo, hi = (0.0, 0.0, 0.0), (2 * np.pi,) * 3
    g = em.HipSolver((n, n, n), lo, hi, u, mode="weno", cfl=CFL, real="f32")
    tiles = pick_tiles(g, args, args.f32_tiles)
    g.upload(u)
    g.run_steps(args.warmup)
    g.synchronize()
    t0 = time.perf_counter()
    pass_ms = g.run_steps(args.steps, timed=True)
    g.synchronize()
    el = time.perf_counter() - t0
    value = n ** 3 * args.steps * 3 / el
    dom = int(np.argmax(pass_ms))
    achieved = FLOP_PER_CELL_STAGE / 3.0 * n ** 3 / (pass_ms[dom] * 1e-3)
    return {
        "workload": f"3-D periodic {n}^3 (same state), FluxMode::Weno, CFL {CFL}, FP32 state and arithmetic",
        "value": value,
        "unit": "cell-updates/s per RK stage",
        "dtype": "f32",
        "ms_per_step": el / args.steps * 1e3,
        "tiles": tiles,
        "pass_ms": pass_ms,
        "roofline": {"bound": "mfma", "roof": "fp32-valu", "kernel": f"axis_pass_f32 d={dom}",
                     "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32, "hbm_fraction": value * BYTES_PER_CELL_STAGE / 2 / PEAK_HBM},
    }


def pick_tiles(g, args, fixed):
    """Tile shape per axis: the given indices, the defaults, or the
    autotuner's pick (a few timed passes of every shape)."""
    if fixed:
        tiles = [int(v) for v in fixed.split(",")]
        g.set_tiles(tiles)
        return tiles
    return None if args.no_autotune else g.autotune(3)


def hybrid_bench(em, args, n, u):
    """The same 256^3 workload in the reference's default mode (Hybrid: the
    KEEP-PE central flux on smooth faces, WCNS where the sensor flags, a blend
    between).  Algorithmic FLOPs per cell and stage: the shock-path share s
    of the faces at the Weno figure, the rest at the central one."""
    lo, hi = (0.0, 0.0, 0.0), (2 * np.pi,) * 3
    g = em.HipSolver((n, n, n), lo, hi, u, mode="hybrid", cfl=CFL, precise=not args.fast)
    tiles = pick_tiles(g, args, args.hybrid_tiles)
    g.upload(u)
    g.run_steps(args.warmup)
    c0 = g.counters()[0]
    g.synchronize()
    t0 = time.perf_counter()
    pass_ms = g.run_steps(args.steps, timed=True)
    g.synchronize()
    el = time.perf_counter() - t0
    share = (g.counters()[0] - c0) / (args.steps * 3 * 3 * (n + 1) * n * n)
    value = n ** 3 * args.steps * 3 / el
    flop = share * FLOP_PER_CELL_STAGE + (1.0 - share) * FLOP_PER_CELL_STAGE_CENTRAL
    dom = int(np.argmax(pass_ms))
    achieved = flop / 3.0 * n ** 3 / (pass_ms[dom] * 1e-3)
    return {
        "workload": f"3-D periodic {n}^3 (same state), FluxMode::Hybrid (tol 1e-3), CFL {CFL}",
        "value": value,
        "unit": "cell-updates/s per RK stage",
        "ms_per_step": el / args.steps * 1e3,
        "tiles": tiles,
        "pass_ms": pass_ms,
        "shock_path_face_share": share,
        "roofline": {"bound": "mfma", "roof": "fp64-valu", "kernel": f"axis_pass d={dom}", "achieved": achieved / 1e12,
                     "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64,
                     "traffic": profiled_traffic("fast" if args.fast else "precise", 0, dom,
                                                 TILE_SHAPES[(tiles or [6, 6, 6])[dom]]),
                     "flop_per_cell_stage": flop, "hbm_fraction": value * BYTES_PER_CELL_STAGE / PEAK_HBM},
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is a design note only)", file=sys.stderr)
        sys.exit(2)
    import emergence_amd as em

    if args.amr_only:
        print(json.dumps({"amr": amr_bench(em, args)}), flush=True)
        return
    if args.amr_group_only:
        print(json.dumps({"amr_group": amr_group_bench(em, args)}), flush=True)
        return
    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_bench(em, args)}), flush=True)
        return
    if args.ensemble_only:
        print(json.dumps({"ensemble": ensemble_bench(em, args)}), flush=True)
        return
    if args.fiveeq_only:
        print(json.dumps({"fiveeq": fiveeq_bench(em, args)}), flush=True)
        return
    n = args.n
    lo, hi = (0.0, 0.0, 0.0), (2 * np.pi,) * 3
    u = em.case_init(1, (n, n, n), lo, hi, 6, [0.3, float(0x5EED)])
    g = em.HipSolver((n, n, n), lo, hi, u, mode="weno", cfl=CFL, precise=not args.fast)
    tiles = pick_tiles(g, args, args.tiles)
    g.upload(u)
    if args.warmup:
        g.run_steps(args.warmup)
    g.synchronize()
    t0 = time.perf_counter()
    pass_ms = g.run_steps(args.steps, timed=True)
    g.synchronize()
    el = time.perf_counter() - t0
    cells = n ** 3
    value = cells * args.steps * 3 / el
    ms_per_step = el / args.steps * 1e3
    # dominant kernel: the slowest axis pass
    dom = int(np.argmax(pass_ms))
    flop_launch = FLOP_PER_CELL_STAGE / 3.0 * cells
    achieved = flop_launch / (pass_ms[dom] * 1e-3)
    rec = {
        "metric": "FP64 cell-updates/s per RK stage (256³, batched AMR); adjoint/forward cost ratio",
        "value": value,
        "unit": "cell-updates/s per RK stage",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Fourier-mode density, Taylor-Green velocity)",
        "config": {
            "workload": f"3-D periodic {n}^3 ideal gas, WCNS+HLLC+positivity (FluxMode::Weno), SSP-RK3, CFL {CFL}",
            "cells": cells,
            "variant": "fast" if args.fast else "precise",
            "tiles": tiles,
            "pass_ms": pass_ms,
        },
        "roofline": {
            # compute roof: MI355X's FP64 peak, 78.6 TFLOP/s for vector and
            # matrix FP64 alike; the pass runs FP64 VALU code, no MFMA
            "bound": "mfma",
            "roof": "fp64-valu (the compute roof: 78.6 TFLOP/s FP64; the pass issues VALU FP64, no MFMA)",
            "kernel": f"axis_pass d={dom}",
            "achieved": achieved / 1e12,
            "peak": PEAK_FP64 / 1e12,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64,
            "traffic": profiled_traffic("fast" if args.fast else "precise", 1, dom, TILE_SHAPES[(tiles or [6, 6, 6])[dom]]),
            "hbm_fraction": value * BYTES_PER_CELL_STAGE / PEAK_HBM,
            "fp64_fraction": value * FLOP_PER_CELL_STAGE / PEAK_FP64,
        },
    }
    if not args.no_cpu_baseline:
        rec["cpu_baseline"] = cpu_baseline(em, n, u, args.cpu_repeats)
    del g
    if not args.no_hybrid:
        rec["hybrid"] = hybrid_bench(em, args, n, u)
    if not args.no_f32:
        rec["f32"] = f32_bench(em, args, n, u)
    if not args.no_amr:
        rec["amr"] = amr_bench(em, args)
    if not args.no_amr_group:
        rec["amr_group"] = amr_group_bench(em, args)
    if not args.no_adjoint:
        rec["adjoint"] = adjoint_bench(em, args)
        rec["adjoint_forward_cost_ratio"] = rec["adjoint"]["cost_ratio"]  # checkpoints every --adj-checkpoint steps
        rec["adjoint_forward_cost_ratio_kept_states"] = rec["adjoint"]["kept_states"]["cost_ratio"]
    if not args.no_ensemble:
        rec["ensemble"] = ensemble_bench(em, args)
    if not args.no_fiveeq:
        rec["fiveeq"] = fiveeq_bench(em, args)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
