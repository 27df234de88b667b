#!/usr/bin/env python3
"""Throughput of the HelioCubed MHD RHS on one MI355X: cell-updates/s per RK stage.

One step = one classical RK4 step (4 stage preparations + 4 RHS evaluations
+ the stage accumulate) of the whole six-panel shell, inputs resident in HBM.
Prints one JSON line. The CPU baseline is the FP64 oracle (oracle/) timed on
this host's cores on one RK4 step of the same grid.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Workloads: grid, operator options, and the algorithmic bytes per unit of
# the kernels that can dominate (unit = face for the face-flux kernels, cell
# for the others).
WORKLOADS = {
    # 6 panels x 128 radial x 64x64 angular, fourth-order standard operator +
    # Roe 8-wave, radial wind with density noise and Alfvenic perturbations
    "shell-6x128x64x64-o4-roe": dict(n_rad=128, n_ang=64, limiter=1, riemann=2, order=4, config=3),
    # same grid, Berta WENOZ (Jameson detector, WENO3 fallback) + Roe, with a
    # seeded low-beta cone (p = 1e-4 |B|^2/2 within 15 degrees of an axis)
    "shell-6x128x64x64-berta-roe": dict(n_rad=128, n_ang=64, limiter=1, riemann=2, order=4, config=4, op=2),
    # 6 panels x 256 radial x 128x128 angular (25.2M cells, 768 boxes of 32^3),
    # Berta WENOZ + Roe; four-sector B_r with a magnetized blob moving at v_r = 2
    "shell-6x256x128x128-berta-roe": dict(n_rad=256, n_ang=128, limiter=1, riemann=2, order=4, config=5, op=2),
    "shell-6x128x64x64-o2-minmod-roe": dict(n_rad=128, n_ang=64, limiter=1, riemann=2, order=2, config=0),
    "shell-6x32x32x32-o2-minmod-rusanov": dict(n_rad=32, n_ang=32, limiter=1, riemann=1, order=2, config=0),
    # three-level AMR box table in one launch per kernel: base 6 x 64 x 48x48 in
    # 32x16x16 boxes, level 1 (2x) 16^3 boxes in a 30-degree wedge, level 2 (4x)
    # 8^3 boxes at a shock front inside the wedge; 652 boxes, RHS of every box
    "amr-3level-652boxes-o4-roe": dict(n_rad=64, n_ang=48, limiter=1, riemann=2, order=4, config=3, amr=True),
}


def amr_table():
    """Levels (n_rad, n_ang) and the (level, (block, lo, hi)) box table of the
    three-level hierarchy: the wedge is the central 32x32 angular cells of
    panel 1 at level 1 (30 degrees of the 90-degree panel), the shock front
    the radial band [96, 160) of level 2 inside it."""
    levels = [(64, 48), (128, 96), (256, 192)]
    table = []
    for p in range(6):
        for k in range(0, 48, 16):
            for j in range(0, 48, 16):
                for i in range(0, 64, 32):
                    table.append((0, (p, (i, j, k), (i + 31, j + 15, k + 15))))
    for k in range(32, 64, 16):
        for j in range(32, 64, 16):
            for i in range(0, 128, 16):
                table.append((1, (1, (i, j, k), (i + 15, j + 15, k + 15))))
    for k in range(64, 128, 8):
        for j in range(64, 128, 8):
            for i in range(96, 160, 8):
                table.append((2, (1, (i, j, k), (i + 7, j + 7, k + 7))))
    return levels, table
# Algorithmic HBM bytes per unit for each kernel (8 FP64 components = 64 B):
#   face flux: read the 8-component cell state once + write the 8-component face flux;
#   divergence: read 3 face fluxes + write K + read-modify-write U;
#   prepare: read U0 and K, write W.
ALGO_BYTES = {
    # fourth-order stage: read U_stage (as JU and U0+K), read-modify-write U, write the next stage
    "k4_rhs": 320, "k4_prepare": 320,
    # fused RHS: read the 8-component stage state, write K, read-modify-write U
    "k_rhs_fused": 64 + 64 + 128,
    "k_face_flux_0": 128, "k_face_flux_1": 128, "k_face_flux_2": 128,
    "k_divergence": 64 * 3 + 64 + 128,
    "k_prepare_valid": 64 * 3,
}


def pmc_summary(kernel, workload):
    """HBM traffic (bytes) and FP64 work (flop) per launch of `kernel` on
    `workload` from the committed rocprofv3 PMC summary (FETCH_SIZE and
    WRITE_SIZE scaled by the copy-kernel calibration, FP64 VALU instruction
    counts), or (None, None) when no summary exists for the pair."""
    import re
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_pmc.txt")
    if not os.path.exists(path):
        return None, None, None
    txt = open(path).read()
    t = re.search(r"HBM traffic per stage: ([0-9.]+) GB", txt)
    f = re.search(r"FP64 work per stage: ([0-9.]+) GFLOP", txt)
    return (float(t.group(1)) * 1e9 if t else None), (float(f.group(1)) * 1e9 if f else None), \
        os.path.relpath(path, ROOT)


def units_per_launch(name, nr, na):
    if name == "k_face_flux_0":
        return 6 * (nr + 1) * na * na
    if name in ("k_face_flux_1", "k_face_flux_2"):
        return 6 * nr * (na + 1) * na
    return 6 * nr * na * na


def cpu_baseline(wl, ju, inner, dt):
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    threads = min(threads, 16)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc  # the oracle is the timed CPU baseline only

    cells = 6 * wl["n_rad"] * wl["n_ang"] ** 2
    if wl["order"] == 4:
        O = orc.Oracle4(wl["n_rad"], wl["n_ang"], limiter=wl["limiter"], spatial_operator=wl.get("op", 1))
        t0 = time.perf_counter()
        O.rhs(O.prepare(ju, inner))
        t = time.perf_counter() - t0
        return {"value": cells / t, "unit": "cell-updates/s per RK stage", "cores": threads, "kind": "port",
                "sample": f"1 RK stage (prepare + RHS) of the same shell on the CPU FP64 oracle, "
                          f"{threads} OpenMP threads, {t:.2f} s"}
    O = orc.Oracle(wl["n_rad"], wl["n_ang"], limiter=wl["limiter"], riemann=wl["riemann"])
    O.rk4(ju, inner, dt, 1)  # first call builds the OpenMP pool and caches
    t0 = time.perf_counter()
    O.rk4(ju, inner, dt, 1)
    t = time.perf_counter() - t0
    return {"value": cells * 4 / t, "unit": "cell-updates/s per RK stage", "cores": threads,
            "kind": "port",
            "sample": f"1 RK4 step of the same shell on the CPU FP64 oracle, {threads} OpenMP threads, {t:.2f} s"}


def bench_amr(args, wl):
    """One step = one RHS evaluation of every box of the three-level table
    (hc_boxop_apply), inputs resident in HBM; timed with HIP events on the
    launch stream."""
    import numpy as np
    import torch
    import heliocubed_amd as hc

    levels, table = amr_table()
    shells = [hc.Shell(nr, na) for nr, na in levels]
    opts = hc.Options(spatial_order=4, limiter=wl["limiter"], riemann_solver=wl["riemann"],
                      precision=1 if args.precision == "mixed" else 0)
    boxes = [b for _, b in table]
    lev = [l for l, _ in table]
    U = np.concatenate([hc.init_native_boxes(shells[l], opts, [b], config=wl["config"], seed=20261203, quad=1)
                        for l, b in table])
    op = hc.BoxOp(shells, opts, boxes, levels=lev)
    dev = torch.device("cuda", 0)
    Ud = torch.from_numpy(U).to(dev)
    R = torch.zeros(op.count(1), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        op.apply(Ud, R)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(args.steps):
        op.apply(Ud, R)
    stop.record(stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ev_ms = start.elapsed_time(stop) / args.steps
    if not torch.isfinite(R).all():
        raise RuntimeError("non-finite RHS")
    graph_ms = None
    if not args.only_timed:
        # the same evaluation captured once in a hipGraph (torch's graph
        # capture of the library's launches on the current stream) and
        # replayed, outside the timed region above; one replay is checked
        # against an eager evaluation bit for bit
        Rg = torch.zeros_like(R)
        graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            with torch.cuda.graph(graph, stream=side):
                op.apply(Ud, Rg)
        torch.cuda.current_stream().wait_stream(side)
        Rg.zero_()
        graph.replay()
        Re = torch.zeros_like(R)
        op.apply(Ud, Re)
        torch.cuda.synchronize()
        if not torch.equal(Rg, Re):
            raise RuntimeError("graph replay of the box table differs from the eager evaluation")
        g0, g1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        g0.record()
        for _ in range(args.steps):
            graph.replay()
        g1.record()
        torch.cuda.synchronize()
        graph_ms = g0.elapsed_time(g1) / args.steps
        del graph, Rg, Re
    cells = sum((hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1) for _, lo, hi in boxes)
    value = cells * args.steps / elapsed
    algo = 128 * cells   # RHS only: read the 8-component state, write the 8-component RHS
    achieved = algo / (ev_ms / 1e3) / 1e9
    line = {
        "metric": f"cell-updates/s per RK stage ({args.precision})", "value": value, "unit": "cell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64 state, f32 Riemann dissipation",
        "data": "synthetic seeded radial wind with density noise and Alfvenic perturbations on every box",
        "config": {"workload": args.workload, "cells": cells, "boxes": len(boxes),
                   "levels": [list(x) for x in levels], "box_sizes": "32x16x16, 16^3, 8^3",
                   "operator": "standard", "limiter": wl["limiter"], "riemann_solver": wl["riemann"],
                   "precision": args.precision},
        "roofline": {"bound": "hbm", "kernel": "box-table RHS (all launches of one evaluation)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "avg_launch_us": ev_ms * 1e3},
        "graph_ms_per_step": graph_ms,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="shell-6x128x64x64-o4-roe", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only-timed", action="store_true",
                    help="run the warmup and timed steps only (no graph replay, autotuner, CFL or mixed-error "
                         "runs and no CPU baseline): the command profiled under rocprofv3")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the operator shards by boxes across GPUs, which this build does not run)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import heliocubed_amd as hc

    wl = WORKLOADS[args.workload]
    if wl.get("amr"):
        bench_amr(args, wl)
        return
    shell = hc.Shell(wl["n_rad"], wl["n_ang"])
    opts = hc.Options(spatial_order=wl["order"], limiter=wl["limiter"], riemann_solver=wl["riemann"],
                      precision=1 if args.precision == "mixed" else 0, spatial_operator=wl.get("op", 1))
    if wl["order"] == 4:
        # cell averages by a 2^3 Gauss rule (midpoint values on the 25M-cell shell,
        # where the host-side averaging would dominate the run)
        ju, inner = hc.init_level4(shell, opts, wl["config"], 20261203, 1 if wl["n_rad"] > 128 else 2)
    else:
        ju, inner = hc.init_state(shell, opts, wl["config"], 20261203)
    dt = 0.25 * (shell.inner_radius * np.pi / 2 / shell.n_ang) / 1.6
    dev = torch.device("cuda", 0)
    U = torch.from_numpy(ju).to(dev)
    G = torch.from_numpy(inner).to(dev)
    L = hc.Level(shell, opts)
    U_init = U.clone() if args.precision == "mixed" else None

    for _ in range(args.warmup):
        L.rk4_advance(U, G, dt, 1)
    torch.cuda.synchronize()
    L.profile(True)
    t0 = time.perf_counter()
    L.rk4_advance(U, G, dt, args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    prof = L.profile_read()
    L.profile(False)
    if not torch.isfinite(U).all():
        raise RuntimeError("non-finite state after the timed steps")
    graph_ms = tuned = dt_cfl = cfl_ms = None
    if not args.only_timed:
        # the same steps replayed from the captured hipGraph of one RK4 step
        # (per-kernel events are off in this mode), outside the timed region above
        L.rk4_advance(U, G, dt, 1)
        torch.cuda.synchronize()
        tg = time.perf_counter()
        L.rk4_advance(U, G, dt, args.steps)
        torch.cuda.synchronize()
        graph_ms = (time.perf_counter() - tg) / args.steps * 1e3
    if wl["order"] == 4 and not args.only_timed:
        # box-table tile picked by the autotuner (results are identical for
        # every tile); timed like the steps above, outside the timed region
        tile, tune_ms = L.autotune(U, G)
        L.rk4_advance(U, G, dt, 1)
        torch.cuda.synchronize()
        ta = time.perf_counter()
        L.rk4_advance(U, G, dt, args.steps)
        torch.cuda.synchronize()
        ta = time.perf_counter() - ta
        tuned = {"tile": tile, "rhs_ms_per_stage": tune_ms, "ms_per_step": ta / args.steps * 1e3,
                 "value": 6 * wl["n_rad"] * wl["n_ang"] ** 2 * 4 * args.steps / ta}
    if not args.only_timed:
        # level CFL time step (the reference's dtCFL, taken once per step by
        # the driver), timed on its own with the host read-back of dt
        L.dt_cfl(U)
        tc = time.perf_counter()
        for _ in range(10):
            dt_cfl = L.dt_cfl(U)
        cfl_ms = (time.perf_counter() - tc) / 10 * 1e3
    aux = None
    if not args.only_timed:
        # the level's other per-step kernels, each timed with HIP events on the
        # launch stream over 10 calls: conserved-state check (with its host
        # read-back) and inner boundary ghosts from a boundary-file strip
        g = L.ghost()
        n = wl["n_ang"] + 2 * g
        strip = torch.zeros((6, 8, n, n), dtype=torch.float64, device=dev)
        strip[:, 0] = 500.0
        strip[:, 1] = 4e7
        strip[:, 4] = 5e3
        strip[:, 5] = 1e3
        ghosts = torch.empty((6, 8, n, n, g), dtype=torch.float64, device=dev)
        L.inner_ghosts_from_bc(strip, ghosts)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(10):
            L.inner_ghosts_from_bc(strip, ghosts)
        e1.record()
        torch.cuda.synchronize()
        bc_ms = e0.elapsed_time(e1) / 10
        ts = time.perf_counter()
        for _ in range(10):
            ok, _first = L.valid_state(U)
        vs_ms = (time.perf_counter() - ts) / 10 * 1e3
        aux = {"valid_state_ms": vs_ms, "inner_ghosts_from_bc_ms": bc_ms, "state_valid": ok}
        del strip, ghosts

    mixed_err = None
    if U_init is not None and not args.only_timed:
        # mixed precision against FP64 from the same initial state, after 1
        # and after 100 RK4 steps: per-variable relative L-infinity and L2
        # error of the state (outside the timed regions)
        import dataclasses
        L64 = hc.Level(shell, dataclasses.replace(opts, precision=0))
        Um, U64 = U_init.clone(), U_init.clone()
        mixed_err = {}
        done = 0
        for n in (1, 100):
            L.rk4_advance(Um, G, dt, n - done)
            L64.rk4_advance(U64, G, dt, n - done)
            done = n
            torch.cuda.synchronize()
            d = (Um - U64).abs()
            linf = (d.amax(dim=(0, 2, 3, 4)) / U64.abs().amax(dim=(0, 2, 3, 4))).cpu().tolist()
            l2 = (d.square().sum(dim=(0, 2, 3, 4)).sqrt() / U64.square().sum(dim=(0, 2, 3, 4)).sqrt()).cpu().tolist()
            mixed_err[f"after_{n}_steps"] = {"rel_linf": linf, "rel_l2": l2}
        # sensitivity baseline: FP64 from the same state perturbed by a
        # relative 1e-9 (the size of the mixed-precision error after one
        # step), against the unperturbed FP64 state after 100 steps
        gen = torch.Generator(device=dev)
        gen.manual_seed(20261203)
        noise = 2 * torch.rand(U_init.shape, generator=gen, device=dev, dtype=torch.float64) - 1
        Up = U_init * (1 + 1e-9 * noise)
        del noise
        L64.rk4_advance(Up, G, dt, 100)
        torch.cuda.synchronize()
        d = (Up - U64).abs()
        mixed_err["fp64_perturbed_1e-9_after_100_steps"] = {
            "rel_linf": (d.amax(dim=(0, 2, 3, 4)) / U64.abs().amax(dim=(0, 2, 3, 4))).cpu().tolist(),
            "rel_l2": (d.square().sum(dim=(0, 2, 3, 4)).sqrt() / U64.square().sum(dim=(0, 2, 3, 4)).sqrt()).cpu().tolist()}
        del L64, Um, U64, Up, d

    cells = 6 * wl["n_rad"] * wl["n_ang"] ** 2
    value = cells * 4 * args.steps / elapsed
    dominant = max((k for k in prof if k in ALGO_BYTES), key=lambda k: prof[k][0])
    ms, launches = prof[dominant]
    avg_s = ms / launches / 1e3
    algo = ALGO_BYTES[dominant] * units_per_launch(dominant, wl["n_rad"], wl["n_ang"])
    achieved = algo / avg_s / 1e9
    traffic, flop, pmc_path = pmc_summary(dominant, args.workload) if args.precision == "fp64" else (None, None, None)
    line = {
        "metric": f"cell-updates/s per RK stage ({args.precision})",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else (
            "f64 state, f32 reconstruction weights and Riemann dissipation" if wl.get("op", 1) == 2
            else "f64 state, f32 slopes and Riemann dissipation"),
        "data": ("synthetic seeded radial-wind shell with four-sector B_r and a magnetized blob"
                 if wl["config"] == 5 else
                 "synthetic seeded radial-wind shell with a Gaussian pressure pulse"
                 + (", lognormal density noise and Alfvenic perturbations" if wl["config"] in (3, 4) else "")
                 + (", low-beta cone" if wl["config"] == 4 else "")),
        "config": {"workload": args.workload, "cells": cells, "panels": 6, "n_rad": wl["n_rad"],
                   "n_ang": wl["n_ang"], "spatial_order": wl["order"],
                   "operator": "berta-wenoz" if wl.get("op", 1) == 2 else "standard",
                   "limiter": wl["limiter"], "riemann_solver": wl["riemann"], "precision": args.precision,
                   "boxes": "32^3 box table" if wl["order"] == 4 else "panel arrays"},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "avg_launch_us": avg_s * 1e6},
        "kernels_ms_per_step": {k: v[0] / args.steps for k, v in prof.items()},
        "graph_ms_per_step": graph_ms,
        "dt_cfl": {"value": dt_cfl, "ms": cfl_ms},
        "level_aux_kernels": aux,
        "autotuned_tile": tuned,
    }
    if traffic is not None:
        line["roofline"]["traffic_source"] = pmc_path
        line["roofline"]["traffic_hbm_gbs"] = traffic / avg_s / 1e9
    if flop is not None:
        line["valu_fp64"] = {"achieved": flop / avg_s / 1e12, "peak": 78.6, "unit": "TFLOP/s",
                             "frac": flop / avg_s / 1e12 / 78.6, "flop_per_launch": flop, "source": pmc_path}
    if mixed_err is not None:
        line["mixed_vs_fp64"] = mixed_err
    if not args.no_cpu_baseline and not args.only_timed:
        line["cpu_baseline"] = cpu_baseline(wl, ju, inner, dt)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
