#!/usr/bin/env python3
"""Throughput of the HelioCubed MHD RHS on one MI355X: cell-updates/s per RK stage.

One step = one classical RK4 step (4 stage preparations + 4 RHS evaluations
+ the stage accumulate) of the whole six-panel shell, inputs resident in HBM.
The default workload is the 25M-cell shell (6 x 256 x 128 x 128, Berta WENOZ +
Roe 8-wave): the batched mixed-precision level is `value`, the plain FP64 level
is reported beside it, both timed from the captured hipGraph of one step (no
per-kernel events inside the timed region), with

* the per-variable relative L-infinity / L2 error of one mixed-precision stage
  RHS against the FP64 one over the whole shell, and against the FP64 CPU
  oracle on a sample of interior boxes;
* the roofline of the stage RHS (a separate pass with per-kernel HIP events);
* the reference's calling pattern (one BoxOp call per box, FP64) on a sample
  of boxes, scaled to the whole table;
* the CPU baseline: the FP64 oracle's box operator on a sample of the same
  boxes, one box per OpenMP thread at a time.

Prints one JSON line.
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
FP64_PEAK_TFS = 78.6
METRIC = "cell-updates/s per RK stage, mixed vs FP64 at stated L-inf error; % HBM peak"

# Workloads: grid, operator options and the seeded state (config numbers of
# hc_init_level4 / hc_init_state).
WORKLOADS = {
    # 6 x 256 radial x 128x128 angular (25.2M cells, 768 boxes of 32^3), Berta
    # WENOZ + Roe; four-sector B_r with a magnetized blob moving at v_r = 2
    "shell-6x256x128x128-berta-roe": dict(n_rad=256, n_ang=128, limiter=1, riemann=2, order=4, config=5, op=2),
    # 6 x 128 radial x 64x64 angular, fourth-order standard operator + Roe,
    # radial wind with density noise and Alfvenic perturbations
    "shell-6x128x64x64-o4-roe": dict(n_rad=128, n_ang=64, limiter=1, riemann=2, order=4, config=3),
    # same grid, Berta WENOZ (Jameson detector, WENO3 fallback) + Roe, with a
    # seeded low-beta cone (p = 1e-4 |B|^2/2 within 15 degrees of an axis)
    "shell-6x128x64x64-berta-roe": dict(n_rad=128, n_ang=64, limiter=1, riemann=2, order=4, config=4, op=2),
    "shell-6x128x64x64-o2-minmod-roe": dict(n_rad=128, n_ang=64, limiter=1, riemann=2, order=2, config=0),
    "shell-6x32x32x32-o2-minmod-rusanov": dict(n_rad=32, n_ang=32, limiter=1, riemann=1, order=2, config=0),
    # three-level AMR box table in one launch per kernel: base 6 x 64 x 48x48 in
    # 32x16x16 boxes, level 1 (2x) 16^3 boxes in a 30-degree wedge, level 2 (4x)
    # 8^3 boxes at a shock front inside the wedge; 652 boxes, RHS of every box
    "amr-3level-652boxes-o4-roe": dict(n_rad=64, n_ang=48, limiter=1, riemann=2, order=4, config=3, amr=True),
    # three-level AMR hierarchy advanced by the synchronous RK4 (restriction,
    # preparation with coarse-fine interpolation, one box table over every
    # level, flux registers): base 6 x 64 x 48x48, 24 16^3 patches at level 1
    # (2x) over a 30-degree wedge, 576 8^3 patches at level 2 (4x) inside it;
    # Berta WENOZ + Roe, four-sector B_r and a magnetized blob
    "amr-3level-708boxes-berta-roe": dict(n_rad=64, n_ang=48, limiter=1, riemann=2, order=4, config=5, op=2,
                                          amr_rk4=True),
}
DEFAULT_WORKLOAD = "shell-6x256x128x128-berta-roe"

# Algorithmic HBM bytes per cell-update of one RK stage (8 FP64 components =
# 64 B): read U_stage + read U_n + read-modify-write the RK accumulator +
# write the next stage = 320 B; the RHS alone reads the state and writes K.
STAGE_BYTES = 320
RHS_BYTES = 128


def amr_table():
    """Levels and box table of the three-level AMR workload (652 boxes)."""
    from heliocubed_amd.workloads import amr_3level_table
    return amr_3level_table()


def pmc_summary(kernel, workload):
    """HBM traffic (bytes) and FP64 work (flop) per launch of `kernel` on
    `workload` from the committed rocprofv3 PMC summary, or Nones."""
    import re
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_pmc.txt")
    if not os.path.exists(path):
        return None, None, None
    txt = open(path).read()
    t = re.search(r"HBM traffic per stage: ([0-9.]+) GB", txt)
    f = re.search(r"FP64 work per stage: ([0-9.]+) GFLOP", txt)
    return (float(t.group(1)) * 1e9 if t else None), (float(f.group(1)) * 1e9 if f else None), \
        os.path.relpath(path, ROOT)


def algorithmic_flops(workload):
    """Algorithmic FP64 flops per cell-update of one stage RHS from the
    committed op count of the oracle (profiles/opcount_<workload>.txt), or None."""
    import re
    path = os.path.join(ROOT, "profiles", f"opcount_{workload}.txt")
    if not os.path.exists(path):
        return None, None
    m = re.search(r"flop per cell-stage: ([0-9.]+)", open(path).read())
    return (float(m.group(1)) if m else None), os.path.relpath(path, ROOT)


def host_threads():
    """Threads for the CPU leg: OMP_NUM_THREADS (the box's CPU share for one
    GPU) if set, else the cores this process may run on."""
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return (env or avail), avail


def rel_err(test, ref):
    """Per-variable relative L-infinity and L2 error of two [6][8][..] level
    arrays (torch tensors on the device)."""
    d = (test - ref).abs()
    dims = tuple(x for x in range(ref.dim()) if x != 1)
    linf = (d.amax(dim=dims) / ref.abs().amax(dim=dims).clamp_min(1e-300)).cpu().tolist()
    l2 = (d.square().sum(dim=dims).sqrt() / ref.square().sum(dim=dims).sqrt().clamp_min(1e-300)).cpu().tolist()
    return linf, l2


def timed_steps(L, U, G, dt, steps, warmup):
    """Warmup steps, then `steps` RK4 steps replayed from the level's captured
    hipGraph, bracketed by device synchronisation; returns seconds."""
    import torch
    for _ in range(warmup):
        L.rk4_advance(U, G, dt, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L.rk4_advance(U, G, dt, steps)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def profiled(L, U, G, dt, steps):
    """Per-kernel-group milliseconds per step from HIP events on the launch
    stream (a separate pass: the events bypass the graph)."""
    import torch
    L.profile(True)
    L.rk4_advance(U, G, dt, steps)
    torch.cuda.synchronize()
    prof = L.profile_read()
    L.profile(False)
    return prof


def interior_boxes(wl, count):
    """`count` 32^3 boxes of the level's box table that touch no panel edge
    (their RHS needs no seam reconciliation), spread over panels and radii."""
    nr, na = wl["n_rad"], wl["n_ang"]
    out = []
    for p in range(6):
        for k in range(32, na - 32, 32):
            for j in range(32, na - 32, 32):
                for i in range(0, nr, 32):
                    out.append((p, (i, j, k), (i + 31, j + 31, k + 31)))
    if not out:
        return []
    step = max(1, len(out) // count)
    return out[::step][:count]


def box_windows(W, boxes):
    """Ghost-5 windows of the stage level W[6][8][na+10][na+10][nr+10] for
    each box, concatenated in the hc_boxop USph layout (device tensor)."""
    import torch
    parts = []
    for p, lo, hi in boxes:
        parts.append(W[p, :, lo[2]:hi[2] + 11, lo[1]:hi[1] + 11, lo[0]:hi[0] + 11].reshape(-1))
    return torch.cat(parts)


def box_valid(K, boxes):
    import torch
    return torch.cat([K[p, :, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1].reshape(-1)
                      for p, lo, hi in boxes])


def per_var(flat, boxes):
    import numpy as np
    out, off = [[] for _ in range(8)], 0
    for _, lo, hi in boxes:
        n = int(np.prod([hi[d] - lo[d] + 1 for d in range(3)]))
        blk = flat[off:off + 8 * n].reshape(8, n)
        for c in range(8):
            out[c].append(blk[c])
        off += 8 * n
    return np.array([np.concatenate(o) for o in out])


def cpu_baseline(wl, boxes, Ub, reps=5, warm=3):
    """The FP64 oracle's box operator on `boxes` (one box per OpenMP thread
    at a time): `warm` untimed runs, then the median of `reps`. Returns the
    baseline record and the oracle RHS of the boxes."""
    import numpy as np
    threads, avail = host_threads()
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc  # the oracle is the timed CPU baseline and the checker only

    kw = dict(limiter=wl["limiter"], spatial_operator=wl.get("op", 1), riemann=wl["riemann"])
    ref = None
    times = []
    for r in range(warm + reps):
        t0 = time.perf_counter()
        ref = orc.o4_rhs_table(wl["n_rad"], wl["n_ang"], boxes, Ub, **kw)
        if r >= warm:
            times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    cells = sum((hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1) for _, lo, hi in boxes)
    rec = {"value": cells / t, "unit": "cell-updates/s per RK stage (stage RHS)", "cores": threads,
           "kind": "port",
           "sample": f"stage RHS of {len(boxes)} interior 32^3 boxes of the same shell and state on the CPU FP64 "
                     f"oracle (box operator, one box per OpenMP thread, {threads} threads of {avail} available, "
                     f"{os.cpu_count()} online); median of {reps} runs after {warm} warm-ups, {t:.2f} s per run; "
                     f"the stage preparation is not in the sample"}
    return rec, ref


def plain_per_box(hc, shell, opts, W, boxes, nbox_total):
    """The reference's calling pattern: one BoxOp evaluation per box
    (operator() per box in StageLevel), FP64, each box its own table; the
    per-box time of the sample times the table's box count."""
    import torch
    ops = [hc.BoxOp(shell, opts, [b]) for b in boxes]
    ins = [box_windows(W, [b]).contiguous() for b in boxes]
    outs = [torch.zeros(op.count(1), dtype=torch.float64, device=W.device) for op in ops]
    for op, u, r in zip(ops, ins, outs):
        op.apply(u, r)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for op, u, r in zip(ops, ins, outs):
        op.apply(u, r)
    e1.record()
    torch.cuda.synchronize()
    per_box_ms = e0.elapsed_time(e1) / len(boxes)
    del ops, ins, outs
    return per_box_ms, per_box_ms * nbox_total


def bench_level(args, wl):
    import dataclasses
    import numpy as np
    import torch
    import heliocubed_amd as hc

    shell = hc.Shell(wl["n_rad"], wl["n_ang"])
    base = hc.Options(spatial_order=wl["order"], limiter=wl["limiter"], riemann_solver=wl["riemann"],
                      spatial_operator=wl.get("op", 1))
    o64, omix = dataclasses.replace(base, precision=0), dataclasses.replace(base, precision=1)
    if wl["order"] == 4:
        # cell averages by a 2^3 Gauss rule (midpoint values on the 25M-cell
        # shell, where the host-side averaging would dominate the run)
        ju, inner = hc.init_level4(shell, base, wl["config"], 20261203, 1 if wl["n_rad"] > 128 else 2)
    else:
        ju, inner = hc.init_state(shell, base, wl["config"], 20261203)
    dt = 0.25 * (shell.inner_radius * np.pi / 2 / shell.n_ang) / 1.6
    dev = torch.device("cuda", 0)
    U0 = torch.from_numpy(ju).to(dev)
    G = torch.from_numpy(inner).to(dev)
    cells = 6 * wl["n_rad"] * wl["n_ang"] ** 2
    precisions = ["fp64", "mixed"] if args.precision == "both" else [args.precision]

    if args.drift_steps > 0:
        # state drift over many RK4 steps: mixed against FP64, and FP64 from
        # an input perturbed by a relative 1e-9 (the size of one mixed step's
        # state error) against FP64, from the same initial state
        import numpy as np
        n = args.drift_steps
        runs = {}
        rng = np.random.default_rng(20261203)
        pert = torch.from_numpy(1 + 1e-9 * (2 * rng.random(ju.shape) - 1)).to(dev)
        for name, o, U in (("fp64", o64, U0.clone()), ("mixed", omix, U0.clone()),
                           ("fp64_perturbed", o64, U0 * pert)):
            L = hc.Level(shell, o)
            L.rk4_advance(U, G, dt, n)
            torch.cuda.synchronize()
            if not torch.isfinite(U).all():
                raise RuntimeError(f"non-finite state after {n} steps ({name})")
            runs[name] = U
            del L
        out = {"workload": args.workload, "steps": n, "dt": dt, "cells": cells}
        for name in ("mixed", "fp64_perturbed"):
            linf, l2 = rel_err(runs[name], runs["fp64"])
            out[f"{name}_vs_fp64"] = {"rel_linf": linf, "rel_l2": l2, "max_rel_linf": max(linf), "max_rel_l2": max(l2)}
        print(json.dumps({"drift": out}), flush=True)
        return

    res = {}
    for prec in precisions:
        L = hc.Level(shell, o64 if prec == "fp64" else omix)
        if args.tile:
            t = [int(x) for x in args.tile.split(",")]
            L.set_tile(t[0] if len(t) == 1 else tuple(t))
        U = U0.clone()
        t = timed_steps(L, U, G, dt, args.steps, args.warmup)
        if not torch.isfinite(U).all():
            raise RuntimeError(f"non-finite state after the timed steps ({prec})")
        shape = L.tile_shape if wl["order"] == 4 else None
        r = {"L": L, "elapsed": t, "value": cells * 4 * args.steps / t, "ms_per_step": t / args.steps * 1e3,
             "tile": shape}
        if wl["order"] == 4 and not args.tile and not args.no_autotune:
            # the autotuner's tile shape (identical results for every shape),
            # timed over the same steps; the faster of the two is kept
            first = {"tile": list(shape), "ms_per_step": r["ms_per_step"], "value": r["value"]}
            tuned, tune_ms = L.autotune(U0, G)
            U = U0.clone()
            t2 = timed_steps(L, U, G, dt, args.steps, args.warmup)
            second = {"tile": list(tuned), "ms_per_step": t2 / args.steps * 1e3, "value": cells * 4 * args.steps / t2,
                      "autotuner_rhs_ms_per_stage": tune_ms}
            r["tiles_timed"] = [first, second]
            if t2 < t:
                r.update({"elapsed": t2, "value": second["value"], "ms_per_step": second["ms_per_step"],
                          "tile": tuned})
            else:
                L.set_tile(tuple(shape))
        res[prec] = r
        del U
    head = "mixed" if "mixed" in res else precisions[0]
    value, elapsed = res[head]["value"], res[head]["elapsed"]

    line = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else (
            "f64 state, f32 reconstruction weights and Riemann eigen-system" if wl.get("op", 1) == 2
            else "f64 state, f32 slopes / Riemann eigen-system"),
        "data": ("synthetic seeded radial-wind shell with four-sector B_r and a magnetized blob"
                 if wl["config"] == 5 else
                 "synthetic seeded radial-wind shell with a Gaussian pressure pulse"
                 + (", lognormal density noise and Alfvenic perturbations" if wl["config"] in (3, 4) else "")
                 + (", low-beta cone" if wl["config"] == 4 else "")),
        "config": {"workload": args.workload, "cells": cells, "panels": 6, "n_rad": wl["n_rad"],
                   "n_ang": wl["n_ang"], "spatial_order": wl["order"],
                   "operator": "berta-wenoz" if wl.get("op", 1) == 2 else ("standard" if wl["order"] == 4 else "plm"),
                   "limiter": wl["limiter"], "riemann_solver": wl["riemann"], "precision": head,
                   "boxes": ("x".join(str(x) for x in res[head]["tile"]) + " box table (32^3 configured; "
                             "the autotuned shape when faster over the timed steps)" if wl["order"] == 4
                             else "panel arrays")},
    }
    for prec in precisions:
        line[f"{prec}_batched"] = {"value": res[prec]["value"], "ms_per_step": res[prec]["ms_per_step"],
                                   "tile": list(res[prec]["tile"]) if res[prec]["tile"] else None}
        if "tiles_timed" in res[prec]:
            line[f"{prec}_batched"]["tiles_timed"] = res[prec]["tiles_timed"]
    if len(res) == 2:
        line["mixed_speedup_vs_fp64_batched"] = res["mixed"]["value"] / res["fp64"]["value"]

    if args.only_timed:
        print(json.dumps(line), flush=True)
        return

    # per-kernel timing of the headline precision and of FP64 (separate
    # passes: events on the launch stream around every launch group)
    nprof = min(args.steps, 5)
    kernels = {}
    for prec in precisions:
        U = U0.clone()
        prof = profiled(res[prec]["L"], U, G, dt, nprof)
        kernels[prec] = {k: v[0] / nprof for k, v in prof.items()}
        res[prec]["prof"] = prof
        del U
    line["kernels_ms_per_step"] = kernels

    dominant = "k4_rhs" if wl["order"] == 4 else "k_rhs_fused"
    rhs_ms = res[head]["prof"][dominant][0] / res[head]["prof"][dominant][1]
    algo_bytes = STAGE_BYTES * cells
    achieved = algo_bytes / (rhs_ms / 1e3) / 1e9
    traffic, pmc_flop, pmc_path = pmc_summary(dominant, args.workload)
    fpc, fpc_path = algorithmic_flops(args.workload)
    roof = {"bound": "hbm", "kernel": f"{dominant} (every launch of one stage RHS, {head})",
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic, "avg_launch_us": rhs_ms * 1e3,
            "algorithmic_bytes_per_launch": algo_bytes}
    if traffic is not None:
        roof["traffic_source"] = pmc_path
    if fpc is not None:
        # the fourth-order operators are FP64-VALU-bound: algorithmic flops of
        # the oracle's op count over the measured stage RHS time
        tf = fpc * cells / (rhs_ms / 1e3) / 1e12
        line["roofline"] = {"bound": "fp64-valu", "kernel": roof["kernel"], "achieved": tf, "peak": FP64_PEAK_TFS,
                            "unit": "TFLOP/s", "frac": tf / FP64_PEAK_TFS, "traffic": traffic,
                            "avg_launch_us": rhs_ms * 1e3, "flop_per_cell_stage": fpc, "flop_source": fpc_path,
                            "hbm": {k: roof[k] for k in ("achieved", "peak", "unit", "frac")}}
        if traffic is not None:
            # bytes the counters saw move per second (scratch round trips included)
            line["roofline"]["hbm"]["counted_GBs"] = traffic / (rhs_ms / 1e3) / 1e9
            line["roofline"]["traffic_source"] = pmc_path
    else:
        line["roofline"] = roof
    if pmc_flop is not None:
        # the counted FP64 instructions are those of the FP64 build: rate over its own stage RHS time
        pf = res["fp64"]["prof"][dominant] if "fp64" in res else res[head]["prof"][dominant]
        ms64 = pf[0] / pf[1]
        line["valu_fp64_counters"] = {"achieved": pmc_flop / (ms64 / 1e3) / 1e12, "peak": FP64_PEAK_TFS,
                                      "unit": "TFLOP/s", "flop_per_launch": pmc_flop, "precision": "fp64",
                                      "avg_launch_us": ms64 * 1e3, "source": pmc_path}

    # one stage RHS from the same prepared stage input in both precisions
    if wl["order"] == 4 and len(res) == 2:
        L64, Lm = res["fp64"]["L"], res["mixed"]["L"]
        W = torch.zeros(L64.stage_shape, dtype=torch.float64, device=dev)
        L64.prepare_stage(U0, G, W)
        K64 = torch.zeros(L64.valid_shape, dtype=torch.float64, device=dev)
        Km = torch.zeros_like(K64)
        L64.stage_rhs(W, K64)
        Lm.stage_rhs(W, Km)
        torch.cuda.synchronize()
        linf, l2 = rel_err(Km, K64)
        line["mixed_rhs_error_vs_fp64_gpu"] = {"rel_linf": linf, "rel_l2": l2, "max_rel_linf": max(linf),
                                               "cells": cells}
        sample = interior_boxes(wl, args.cpu_boxes or 2 * host_threads()[0])
        if sample and not args.no_cpu_baseline:
            Ub = box_windows(W, sample).cpu().numpy()
            rec, kref = cpu_baseline(wl, sample, Ub)
            line["cpu_baseline"] = rec
            kref = per_var(kref, sample)
            from_gpu = {p: per_var(box_valid(K, sample).cpu().numpy(), sample) for p, K in (("fp64", K64),
                                                                                           ("mixed", Km))}
            # relative to the largest |K| of each variable over the whole
            # shell (the FP64 level, which matches the oracle on the sample
            # to round-off), and to the largest |K| inside the sample
            shell_scale = K64.abs().amax(dim=(0, 2, 3, 4)).cpu().numpy()
            sample_scale = np.abs(kref).max(axis=1)
            shell_scale[shell_scale == 0] = 1.0
            sample_scale[sample_scale == 0] = 1.0
            line["rhs_error_vs_oracle"] = {
                "sample": f"{len(sample)} interior 32^3 boxes (the CPU baseline's sample)",
                **{f"{p}_rel_linf": (np.abs(from_gpu[p] - kref).max(axis=1) / shell_scale).tolist() for p in from_gpu},
                **{f"{p}_rel_linf_sample_scale": (np.abs(from_gpu[p] - kref).max(axis=1) / sample_scale).tolist()
                   for p in from_gpu}}
        if not args.no_plain:
            nbox = len(hc.shell_boxes(shell))
            psample = interior_boxes(wl, 8)
            if psample:
                per_box, total = plain_per_box(hc, shell, o64, W, psample, nbox)
                b64 = res["fp64"]["prof"]["k4_rhs"]
                batched64 = b64[0] / b64[1]
                bm = res["mixed"]["prof"]["k4_rhs"]
                batched_m = bm[0] / bm[1]
                line["plain_per_box_fp64"] = {
                    "rhs_ms_per_stage": total, "per_box_ms": per_box,
                    "sample": f"{len(psample)} boxes, one hc_boxop_apply per box (each box its own table), "
                              f"scaled to {nbox} boxes",
                    "batched_fp64_rhs_ms_per_stage": batched64, "batched_mixed_rhs_ms_per_stage": batched_m,
                    "batched_mixed_rhs_speedup_vs_plain_fp64": total / batched_m}
        del W, K64, Km
    print(json.dumps(line), flush=True)


def bench_amr(args, wl):
    """One step = one RHS evaluation of every box of the three-level table
    (hc_boxop_apply), inputs resident in HBM; timed with HIP events on the
    launch stream."""
    import numpy as np
    import torch
    import heliocubed_amd as hc

    levels, table = amr_table()
    shells = [hc.Shell(nr, na) for nr, na in levels]
    prec = "mixed" if args.precision == "mixed" else "fp64"
    opts = hc.Options(spatial_order=4, limiter=wl["limiter"], riemann_solver=wl["riemann"],
                      precision=1 if prec == "mixed" else 0)
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
        # replayed; one replay is checked against an eager evaluation bit for bit
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
    algo = RHS_BYTES * cells
    achieved = algo / (ev_ms / 1e3) / 1e9
    line = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == "fp64" else "f64 state, f32 Riemann eigen-system",
        "data": "synthetic seeded radial wind with density noise and Alfvenic perturbations on every box",
        "config": {"workload": args.workload, "cells": cells, "boxes": len(boxes),
                   "levels": [list(x) for x in levels], "box_sizes": "32x16x16, 16^3, 8^3",
                   "operator": "standard", "limiter": wl["limiter"], "riemann_solver": wl["riemann"],
                   "precision": prec},
        "roofline": {"bound": "hbm", "kernel": "box-table RHS (all launches of one evaluation)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "avg_launch_us": ev_ms * 1e3},
        "graph_ms_per_step": graph_ms,
    }
    print(json.dumps(line), flush=True)


def amr_rel_err(test, ref, shells, patches):
    """Per-variable max |test - ref| / max |ref| over a hierarchy's combined
    array (level 0, then every patch), numpy."""
    import numpy as np
    nr0, na0 = shells[0]
    n0 = 6 * 8 * na0 * na0 * nr0
    d0 = np.abs(test[:n0] - ref[:n0]).reshape(6, 8, -1).max(axis=(0, 2))
    r0 = np.abs(ref[:n0]).reshape(6, 8, -1).max(axis=(0, 2))
    off = n0
    for _, (_, lo, hi) in patches:
        m = 8 * (hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1)
        d0 = np.maximum(d0, np.abs(test[off:off + m] - ref[off:off + m]).reshape(8, -1).max(axis=1))
        r0 = np.maximum(r0, np.abs(ref[off:off + m]).reshape(8, -1).max(axis=1))
        off += m
    r0[r0 == 0] = 1.0
    return (d0 / r0).tolist()


def bench_amr_rk4(args, wl):
    """One step = one synchronous classical RK4 step of the whole hierarchy
    (hc_amr4_rk4_advance: per stage the restriction, the preparation of every
    level, one box table over every level and the flux registers), inputs
    resident in HBM; value = cell-updates per RK stage over all levels."""
    import numpy as np
    import torch
    import heliocubed_amd as hc
    from heliocubed_amd.workloads import amr_rk4_hierarchy

    levels, patches = amr_rk4_hierarchy()
    shells = [hc.Shell(nr, na) for nr, na in levels]
    base = hc.Options(spatial_order=4, limiter=wl["limiter"], riemann_solver=wl["riemann"],
                      spatial_operator=wl.get("op", 1))
    import dataclasses
    ju0, inner = hc.init_level4(shells[0], base, wl["config"], 20261203, 1)
    ju = np.concatenate([ju0.ravel()] + [hc.init_ju_boxes(shells[l], base, [b], wl["config"], 20261203, 1)
                                          for l, b in patches])
    dev = torch.device("cuda", 0)
    U0 = torch.from_numpy(ju).to(dev)
    G = torch.from_numpy(inner).to(dev)
    cells = 6 * levels[0][0] * levels[0][1] ** 2 + sum(
        (hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1) for _, (_, lo, hi) in patches)
    dt = 0.1 * (shells[0].inner_radius * np.pi / 2 / levels[-1][1]) / 1.6
    precisions = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    res = {}
    for prec in precisions:
        A = hc.Amr4(shells, dataclasses.replace(base, precision=1 if prec == "mixed" else 0), patches)
        U = U0.clone()
        for _ in range(args.warmup):
            A.rk4_advance(U, G, dt, 1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        A.rk4_advance(U, G, dt, args.steps)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        if not torch.isfinite(U).all():
            raise RuntimeError(f"non-finite state after the timed steps ({prec})")
        r = {"A": A, "elapsed": t, "value": cells * 4 * args.steps / t, "ms_per_step": t / args.steps * 1e3}
        if not args.only_timed:
            # one stage RHS (restriction, preparation, box table, registers)
            # timed with HIP events on the launch stream
            K = torch.zeros(A.state_count, dtype=torch.float64, device=dev)
            A.stage_rhs(U0, G, K)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(5):
                A.stage_rhs(U0, G, K)
            e1.record()
            torch.cuda.synchronize()
            r["stage_ms"] = e0.elapsed_time(e1) / 5
            r["K"] = K.cpu().numpy()
            # the RK4 step captured in a hipGraph and replayed: checked against
            # the eager step bit for bit, then timed
            Ue = U0.clone()
            A.rk4_advance(Ue, G, dt, 1)
            Ug = U0.clone()
            graph = torch.cuda.CUDAGraph()
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                with torch.cuda.graph(graph, stream=side):
                    A.rk4_advance(Ug, G, dt, 1)
            torch.cuda.current_stream().wait_stream(side)
            Ug.copy_(U0)
            graph.replay()
            torch.cuda.synchronize()
            if not torch.equal(Ug, Ue):
                raise RuntimeError("graph replay of the AMR RK4 step differs from the eager step")
            g0, g1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            g0.record()
            for _ in range(args.steps):
                graph.replay()
            g1.record()
            torch.cuda.synchronize()
            r["graph_ms_per_step"] = g0.elapsed_time(g1) / args.steps
            del graph, Ue, Ug, K
        res[prec] = r
        del U
    head = "mixed" if "mixed" in res else precisions[0]
    line = {
        "metric": METRIC, "value": res[head]["value"], "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": res[head]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f64 state, f32 reconstruction weights and Riemann eigen-system",
        "data": "synthetic seeded radial wind with four-sector B_r and a magnetized blob on every level",
        "config": {"workload": args.workload, "cells": cells, "levels": [list(x) for x in levels],
                   "patches": {"level1": sum(1 for l, _ in patches if l == 1),
                               "level2": sum(1 for l, _ in patches if l == 2)},
                   "boxes": "level 0: 32^3 tiles (108 boxes of 32x16x16 in the reference layout); 16^3; 8^3",
                   "time_stepping": "synchronous classical RK4 (advanceRungeKuttaSynchronous)",
                   "operator": "berta-wenoz" if wl.get("op", 1) == 2 else "standard",
                   "limiter": wl["limiter"], "riemann_solver": wl["riemann"], "precision": head},
    }
    for prec in precisions:
        line[f"{prec}_amr"] = {k: res[prec][k] for k in ("value", "ms_per_step", "stage_ms", "graph_ms_per_step")
                               if k in res[prec]}
    if len(res) == 2:
        line["mixed_speedup_vs_fp64"] = res["mixed"]["value"] / res["fp64"]["value"]
    if args.only_timed:
        print(json.dumps(line), flush=True)
        return
    stage_ms = res[head]["stage_ms"]
    fpc, fpc_path = algorithmic_flops(args.workload)
    algo_bytes = STAGE_BYTES * cells
    hbm = algo_bytes / (stage_ms / 1e3) / 1e9
    roof = {"bound": "hbm", "kernel": f"one AMR stage (every launch of hc_amr4_stage_rhs, {head})",
            "achieved": hbm, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": hbm / HBM_PEAK_GBS, "traffic": None,
            "avg_launch_us": stage_ms * 1e3}
    if fpc is not None:
        tf = fpc * cells / (stage_ms / 1e3) / 1e12
        roof = {"bound": "fp64-valu", "kernel": roof["kernel"], "achieved": tf, "peak": FP64_PEAK_TFS,
                "unit": "TFLOP/s", "frac": tf / FP64_PEAK_TFS, "traffic": None, "avg_launch_us": stage_ms * 1e3,
                "flop_per_cell_stage": fpc, "flop_source": fpc_path,
                "hbm": {"achieved": hbm, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": hbm / HBM_PEAK_GBS}}
    line["roofline"] = roof
    if len(res) == 2:
        line["mixed_rhs_error_vs_fp64_gpu"] = {
            "rel_linf": amr_rel_err(res["mixed"]["K"], res["fp64"]["K"], levels, patches), "cells": cells}
    if not args.no_cpu_baseline:
        threads, avail = host_threads()
        os.environ["OMP_NUM_THREADS"] = str(threads)
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc  # the oracle is the timed CPU baseline and the checker only

        O = orc.Amr4Oracle(levels, patches, limiter=wl["limiter"], riemann=wl["riemann"],
                           spatial_operator=wl.get("op", 1))
        times, kref = [], None
        for rep in range(3):
            t0 = time.perf_counter()
            kref = O.stage_rhs(ju, inner)
            if rep:
                times.append(time.perf_counter() - t0)
        t = float(np.median(times))
        line["cpu_baseline"] = {
            "value": cells / t, "unit": "cell-updates/s per RK stage (one AMR stage)", "cores": threads,
            "kind": "port",
            "sample": f"one stage of the whole hierarchy ({cells} cells: restriction, preparation of every level, "
                      f"every patch's box operator, flux registers) on the CPU FP64 oracle, {threads} OpenMP "
                      f"threads of {avail} available; median of 2 runs after 1 warm-up, {t:.2f} s per run"}
        line["rhs_error_vs_oracle"] = {
            "sample": "one stage RHS of the whole hierarchy",
            **{f"{p}_rel_linf": amr_rel_err(res[p]["K"], kref, levels, patches) for p in res}}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD, choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="both", choices=["both", "fp64", "mixed"],
                    help="both: mixed is `value`, FP64 is reported beside it")
    ap.add_argument("--tile", default="", help="box-table tile edge or radial,alpha,beta shape of the level "
                                                "(default: 32^3, then the autotuner's shape if faster)")
    ap.add_argument("--no-autotune", action="store_true", help="time the configured 32^3 tiles only")
    ap.add_argument("--cpu-boxes", type=int, default=0, help="boxes in the CPU sample (0: two per thread)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the one-call-per-box FP64 measurement")
    ap.add_argument("--drift-steps", type=int, default=0,
                    help="instead of timing: the state after this many RK4 steps, mixed and FP64 from an input "
                         "perturbed by 1e-9, each against FP64 (per-variable relative L-inf and L2)")
    ap.add_argument("--only-timed", action="store_true",
                    help="run the warmup and timed steps only: the command profiled under rocprofv3")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the operator shards by boxes across GPUs, which this build does not run)", file=sys.stderr)
        sys.exit(2)
    wl = WORKLOADS[args.workload]
    if wl.get("amr_rk4"):
        bench_amr_rk4(args, wl)
    elif wl.get("amr"):
        bench_amr(args, wl)
    else:
        bench_level(args, wl)


if __name__ == "__main__":
    main()
