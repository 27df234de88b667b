"""The batched and large workloads on the GPU.

* The 652-box three-level AMR table (base 6 x 64 x 48x48 in 32x16x16 boxes,
  16^3 boxes of a refined wedge, 8^3 boxes at its front) evaluated in one set
  of launches: every box against the FP64 oracle at its own level within
  relative L-infinity 1e-12, and the mixed-precision table within 1e-6 of the
  oracle (per variable, over the table).
* One stage of the 25M-cell shell (6 x 256 x 128x128, Berta WENOZ + Roe,
  four-sector B_r with a magnetized blob): a finite stage RHS; mixed within
  1e-6 of FP64 per variable over the whole shell; FP64 against the oracle on
  a sample of interior boxes within 1e-12; and the flux form of the divergence
  per box (the sum of a box's RHS equals minus the net flux through its faces
  divided by the cell widths, to round-off) through the box operator's face
  fluxes.
"""
import time

import numpy as np
import pytest
import torch

import heliocubed_amd as hc
import oracle as orc
from heliocubed_amd.workloads import amr_3level_table
from helpers import per_var, rel_linf_per_var

TOL = 1e-12


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _oracle_table(levels, table, U, opts):
    """Oracle RHS of every box of a multi-level table, in table order."""
    offs, off = [], 0
    for lev, (blk, lo, hi) in table:
        offs.append(off)
        off += 8 * np.prod([hi[d] - lo[d] + 11 for d in range(3)])
    out = [None] * len(table)
    for lev, (nr, na) in enumerate(levels):
        idx = [n for n, (l, _) in enumerate(table) if l == lev]
        boxes = [table[n][1] for n in idx]
        Ul = np.concatenate([U[offs[n]:offs[n] + 8 * np.prod([table[n][1][2][d] - table[n][1][1][d] + 11
                                                               for d in range(3)])] for n in idx])
        r = orc.o4_rhs_table(nr, na, boxes, Ul, limiter=opts.limiter, spatial_operator=opts.spatial_operator,
                             riemann=opts.riemann_solver)
        o = 0
        for n, (_, lo, hi) in zip(idx, boxes):
            m = 8 * np.prod([hi[d] - lo[d] + 1 for d in range(3)])
            out[n] = r[o:o + m]
            o += m
    return np.concatenate(out)


@pytest.mark.gpu
@pytest.mark.parametrize("operator", [1, 2])
def test_amr_652_box_table_matches_oracle_box_by_box(device, operator):
    t0 = time.time()
    levels, table = amr_3level_table()
    assert len(table) == 652
    shells = [hc.Shell(nr, na) for nr, na in levels]
    opts = hc.Options(spatial_order=4, limiter=1, riemann_solver=2, spatial_operator=operator)
    print(f"652-box table, operator {operator}: host init", flush=True)
    U = np.concatenate([hc.init_native_boxes(shells[l], opts, [b], config=3, seed=20261203, quad=1)
                        for l, b in table])
    boxes, lev = [b for _, b in table], [l for l, _ in table]
    op = hc.BoxOp(shells, opts, boxes, levels=lev)
    R = torch.zeros(op.count(1), dtype=torch.float64, device=device)
    op.apply(_dev(U), R)
    mixed = hc.Options(spatial_order=4, limiter=1, riemann_solver=2, spatial_operator=operator, precision=1)
    opm = hc.BoxOp(shells, mixed, boxes, levels=lev)
    Rm = torch.zeros_like(R)
    opm.apply(_dev(U), Rm)
    torch.cuda.synchronize()
    print(f"GPU done ({time.time() - t0:.1f} s), oracle", flush=True)
    ref = _oracle_table(levels, table, U, opts)
    got, gotm = R.cpu().numpy(), Rm.cpu().numpy()
    assert np.isfinite(got).all()
    # box by box at its own level: every box's error, relative to the largest
    # |RHS| of each variable over the table (the bar's normalization), within
    # 1e-12; relative to the box's own largest |RHS| (stricter: a near-steady
    # box amplifies round-off differences of its face fluxes) within 1e-10
    scale = np.abs(per_var(ref, boxes)).max(axis=1)
    off, worst_table, worst_box = 0, 0.0, 0.0
    for _, (blk, lo, hi) in table:
        m = 8 * np.prod([hi[d] - lo[d] + 1 for d in range(3)])
        d = np.abs(got[off:off + m] - ref[off:off + m]).reshape(8, -1).max(axis=1)
        worst_table = max(worst_table, (d / scale).max())
        own = np.abs(ref[off:off + m]).reshape(8, -1).max(axis=1)
        worst_box = max(worst_box, (d / np.where(own == 0, 1.0, own)).max())
        off += m
    errm = rel_linf_per_var(per_var(gotm, boxes), per_var(ref, boxes), axis_var=0)
    print(f"worst per-box FP64 error {worst_table:.2e} (table scale), {worst_box:.2e} (box scale); "
          f"mixed over the table {errm.max():.2e} ({time.time() - t0:.1f} s)", flush=True)
    assert worst_table <= TOL
    assert worst_box <= 1e-10
    assert errm.max() <= 1e-6, errm


@pytest.mark.gpu
def test_25m_cell_shell_stage(device):
    t0 = time.time()
    nr, na = 256, 128
    shell = hc.Shell(nr, na)
    o64 = hc.Options(spatial_order=4, limiter=1, riemann_solver=2, spatial_operator=2)
    omx = hc.Options(spatial_order=4, limiter=1, riemann_solver=2, spatial_operator=2, precision=1)
    print("25M-cell shell: host init", flush=True)
    ju, inner = hc.init_level4(shell, o64, 5, 20261203, 1)
    L64, Lm = hc.Level(shell, o64), hc.Level(shell, omx)
    U, G = _dev(ju), _dev(inner)
    W = torch.zeros(L64.stage_shape, dtype=torch.float64, device=device)
    L64.prepare_stage(U, G, W)
    K64 = torch.zeros(L64.valid_shape, dtype=torch.float64, device=device)
    Km = torch.zeros_like(K64)
    L64.stage_rhs(W, K64)
    Lm.stage_rhs(W, Km)
    torch.cuda.synchronize()
    print(f"GPU stages done ({time.time() - t0:.1f} s)", flush=True)
    assert torch.isfinite(K64).all() and torch.isfinite(Km).all()
    d = (Km - K64).abs().amax(dim=(0, 2, 3, 4)) / K64.abs().amax(dim=(0, 2, 3, 4))
    print("mixed vs FP64 stage RHS over the shell", d.cpu().numpy(), flush=True)
    assert d.max().item() <= 1e-6
    # interior boxes (no seam faces) against the oracle's box operator
    sample = [(p, (i, j, k), (i + 31, j + 31, k + 31)) for p in (0, 3, 5) for (i, j, k) in
              ((0, 32, 32), (96, 64, 32), (224, 32, 64), (128, 64, 64))]
    Ub = torch.cat([W[p, :, lo[2]:hi[2] + 11, lo[1]:hi[1] + 11, lo[0]:hi[0] + 11].reshape(-1)
                    for p, lo, hi in sample])
    kref = orc.o4_rhs_table(nr, na, sample, Ub.cpu().numpy(), limiter=1, spatial_operator=2)
    kgot = torch.cat([K64[p, :, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1].reshape(-1)
                      for p, lo, hi in sample]).cpu().numpy()
    err = rel_linf_per_var(per_var(kgot, sample), per_var(kref, sample), axis_var=0)
    print("FP64 vs oracle on 12 interior boxes", err, flush=True)
    assert err.max() <= TOL, err
    # flux form: the box operator's RHS of each sample box against its face fluxes
    op = hc.BoxOp(shell, o64, sample)
    R = torch.zeros(op.count(1), dtype=torch.float64, device=device)
    F = [torch.zeros(op.count(2 + d), dtype=torch.float64, device=device) for d in range(3)]
    op.apply(Ub.contiguous(), R, F)
    torch.cuda.synchronize()
    Rh, Fh = R.cpu().numpy(), [f.cpu().numpy() for f in F]
    ro, fo = 0, [0, 0, 0]
    dxi = (1.0 / nr, 1.0 / na, 1.0 / na)
    for _, lo, hi in sample:
        n = [hi[d] - lo[d] + 1 for d in range(3)]
        r = Rh[ro:ro + 8 * np.prod(n)].reshape(8, n[2], n[1], n[0])
        ro += 8 * np.prod(n)
        net = np.zeros(8)
        mag = np.zeros(8)
        for d in range(3):
            m = [n[a] + (a == d) for a in range(3)]
            f = Fh[d][fo[d]:fo[d] + 8 * np.prod(m)].reshape(8, m[2], m[1], m[0])
            fo[d] += 8 * np.prod(m)
            ax = 3 - d
            hi_f = np.take(f, [m[d] - 1], axis=ax).sum(axis=(1, 2, 3))
            lo_f = np.take(f, [0], axis=ax).sum(axis=(1, 2, 3))
            net -= (hi_f - lo_f) / dxi[d]
            mag += np.abs(f).sum(axis=(1, 2, 3)) / dxi[d]
        total = r.sum(axis=(1, 2, 3))
        assert (np.abs(total - net) <= 1e-12 * mag).all(), (total, net)
    print(f"flux form holds on every sample box ({time.time() - t0:.1f} s)", flush=True)
