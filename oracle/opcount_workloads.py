#!/usr/bin/env python3
"""Algorithmic operation counts of the fourth-order workloads (test
infrastructure): the oracle's box operator, built with the counting scalar
(hc_opcount.cpp), on sample 32^3 boxes of each workload's grid and seeded
state. Writes profiles/opcount_<workload>.txt with the counts per cell-stage;
bench.py reads the flop figure for the FP64 VALU roofline.

Usage (from the repository root): python3 oracle/opcount_workloads.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import heliocubed_amd as hc  # noqa: E402  (host-side initial data only)
import oracle as orc  # noqa: E402

WORKLOADS = {
    "shell-6x256x128x128-berta-roe": dict(n_rad=256, n_ang=128, op=2, config=5),
    "shell-6x128x64x64-berta-roe": dict(n_rad=128, n_ang=64, op=2, config=4),
    "shell-6x128x64x64-o4-roe": dict(n_rad=128, n_ang=64, op=1, config=3),
}


def sample_boxes(nr, na):
    """Four 32^3 boxes of the workload's table: panel edges and interior,
    inner and outer radii, different panels."""
    return [(1, (0, 0, 0), (31, 31, 31)), (4, (nr // 2, na // 2, na // 2 - 32), (nr // 2 + 31, na // 2 + 31, na // 2 - 1)),
            (2, (nr - 32, na - 32, 0), (nr - 1, na - 1, 31)), (5, (32, 32, na - 32), (63, 63, na - 1))]


def amr_count():
    """The three-level AMR workload: each box of the reference's layout
    recomputes its own ghost cells, so the count per box is taken on one
    sample box of every size (level 0: 32x16x16, level 1: 16^3, level 2: 8^3,
    inside the refined wedge) and weighted by the box counts of the table."""
    from heliocubed_amd.workloads import amr_rk4_hierarchy
    levels, patches = amr_rk4_hierarchy()
    name = "amr-3level-708boxes-berta-roe"
    opts = hc.Options(spatial_order=4, limiter=1, riemann_solver=2, spatial_operator=2)
    samples = [(0, (1, (32, 16, 16), (63, 31, 31)), 108),
               (1, (1, (48, 32, 32), (63, 47, 47)), sum(1 for l, _ in patches if l == 1)),
               (2, (1, (128, 96, 96), (135, 103, 103)), sum(1 for l, _ in patches if l == 2))]
    tot = {k: 0.0 for k in ("add", "mul", "div", "sqrt", "cmp", "other")}
    cells = 0
    for lev, box, count in samples:
        nr, na = levels[lev]
        shell = hc.Shell(nr, na)
        U = hc.init_native_boxes(shell, opts, [box], config=5, seed=20261203, quad=1)
        rhs, cnt = orc.opcount_box(nr, na, box, U, spatial_operator=2)
        ref = orc.o4_rhs_boxes(nr, na, [box], U, spatial_operator=2)
        assert np.array_equal(rhs, ref), "the counting build must reproduce the oracle"
        n = int(np.prod([box[2][d] - box[1][d] + 1 for d in range(3)]))
        for k in tot:
            tot[k] += cnt[k] * count
        cells += n * count
    per = {k: v / cells for k, v in tot.items()}
    flop = per["add"] + per["mul"] + per["div"] + per["sqrt"]
    lines = [
        f"# Algorithmic operation count of one stage RHS: {name}",
        "# the oracle's box operator compiled with a counting scalar (oracle/hc_opcount.cpp) on one box of",
        "# every size of the hierarchy's table (level 0: 32x16x16, 108 boxes; level 1: 16^3; level 2: 8^3,",
        "# the refined patches), weighted by the box counts: each box recomputes its ghost cells as the",
        "# reference's per-box operator does. The preparation, the coarse-fine interpolation and the",
        "# flux registers are not counted. Seeded state of bench.py (config 5).",
        "# command (repository root): python3 oracle/opcount_workloads.py",
        "#",
        "# per cell-stage: " + ", ".join(f"{k} {v:.1f}" for k, v in per.items()),
        f"flop per cell-stage: {flop:.1f}",
        "# flop = add/sub + mul + div + sqrt (each counted as one operation).",
    ]
    path = os.path.join(ROOT, "profiles", f"opcount_{name}.txt")
    open(path, "w").write("\n".join(lines) + "\n")
    print(path, f"{flop:.1f} flop per cell-stage")


def main():
    amr_count()
    for name, wl in WORKLOADS.items():
        nr, na = wl["n_rad"], wl["n_ang"]
        shell = hc.Shell(nr, na)
        opts = hc.Options(spatial_order=4, limiter=1, riemann_solver=2, spatial_operator=wl["op"])
        tot = {k: 0 for k in ("add", "mul", "div", "sqrt", "cmp", "other")}
        cells = 0
        for box in sample_boxes(nr, na):
            U = hc.init_native_boxes(shell, opts, [box], config=wl["config"], seed=20261203, quad=1)
            rhs, cnt = orc.opcount_box(nr, na, box, U, spatial_operator=wl["op"])
            ref = orc.o4_rhs_boxes(nr, na, [box], U, spatial_operator=wl["op"])
            assert np.array_equal(rhs, ref), "the counting build must reproduce the oracle"
            for k in tot:
                tot[k] += cnt[k]
            cells += 32 ** 3
        per = {k: v / cells for k, v in tot.items()}
        flop = per["add"] + per["mul"] + per["div"] + per["sqrt"]
        lines = [
            f"# Algorithmic operation count of one stage RHS: {name}",
            "# the oracle's box operator compiled with a counting scalar (oracle/hc_opcount.cpp), on four",
            "# 32^3 boxes of the workload (the box size of its table; each box recomputes its ghost cells as",
            "# the reference's per-box operator does), seeded state of bench.py; metric and geometry helpers",
            "# not counted (the device path reads precomputed tables). The counting build reproduces the",
            "# oracle's RHS bit for bit on every sample box.",
            "# command (repository root): python3 oracle/opcount_workloads.py",
            "#",
            "# per cell-stage: " + ", ".join(f"{k} {v:.1f}" for k, v in per.items()),
            f"flop per cell-stage: {flop:.1f}",
            "# flop = add/sub + mul + div + sqrt (each counted as one operation); comparisons, min/max and",
            "# fabs are in cmp/other. Divisions are {:.1f} % of the flops.".format(100 * per["div"] / flop),
        ]
        path = os.path.join(ROOT, "profiles", f"opcount_{name}.txt")
        open(path, "w").write("\n".join(lines) + "\n")
        print(path, f"{flop:.1f} flop per cell-stage")


if __name__ == "__main__":
    main()
