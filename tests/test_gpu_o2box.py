"""GPU parity of the second-order per-box operator (hc_boxop2: secondOrderOperator
with Cartesian primitives, src/BoxOp_SecondOrder.H:155-370, as StageLevel calls
it per box at order 2, src/BoxOp_EulerCubedSphereImplem.H:1003-1004).

The boxes' Cartesian primitives (ghost 2) are cut out of a prepared stage
level; the box table is compared with the oracle's per-box restatement within
relative L-infinity 1e-12 per variable (RHS and face fluxes, several
limiters, both Riemann solvers, the Powell source, the divergence scale), and
on boxes whose faces avoid the panel seams with the level operator's RHS.
"""
import ctypes

import numpy as np
import pytest
import torch

import heliocubed_amd as hc
import oracle as orc
from helpers import per_var, rel_linf_per_var

TOL = 1e-12
NR, NA = 12, 10
BOXES = [(0, (0, 1, 2), (5, 6, 8)), (4, (6, 3, 1), (11, 8, 7)), (1, (2, 0, 3), (7, 4, 6)),
         (5, (0, 0, 0), (11, 9, 9)), (3, (4, 7, 0), (4, 9, 2))]


def _stage(limiter, riemann, powell, seed=7):
    opts = hc.Options(limiter=limiter, riemann_solver=riemann, divergence_cleaning=powell)
    ju, inner = hc.init_state(hc.Shell(NR, NA), opts, 0, seed)
    O = orc.Oracle(NR, NA, limiter=limiter, riemann=riemann, divergence_cleaning=powell)
    return opts, O, O.prepare(ju, inner)


def _cut(W, boxes, g=2):
    return np.concatenate([W[p, :, lo[2]:hi[2] + 1 + 2 * g, lo[1]:hi[1] + 1 + 2 * g, lo[0]:hi[0] + 1 + 2 * g].ravel()
                           for p, lo, hi in boxes])


def test_boxop2_rejects_bad_arguments():
    lib, h = hc.lib(), ctypes.c_void_p()
    sh = hc.Shell(NR, NA).c()
    tab = hc.box_table([(0, (0, 0, 0), (3, 3, 3))])
    for bad in (dict(spatial_order=4), dict(riemann_solver=3), dict(limiter=4), dict(precision=1)):
        op = hc.Options(**bad).c()
        assert lib.hc_boxop2_create(ctypes.byref(sh), ctypes.byref(op), tab, 1, ctypes.byref(h)) != 0, bad
    op = hc.Options().c()
    out = hc.box_table([(0, (0, 0, 0), (NR, 3, 3))])
    assert lib.hc_boxop2_create(ctypes.byref(sh), ctypes.byref(op), out, 1, ctypes.byref(h)) != 0
    assert b"outside" in lib.hc_last_error()


@pytest.mark.gpu
@pytest.mark.parametrize("limiter,riemann,powell,scale", [(1, 1, 0, 1.0), (3, 2, 0, 1.0), (0, 2, 0, 0.5),
                                                          (2, 1, 1, 1.0), (1, 2, 1, 2.0)])
def test_boxop2_matches_oracle(device, limiter, riemann, powell, scale):
    opts, O, W = _stage(limiter, riemann, powell)
    Wb = _cut(W, BOXES)
    ref, fref = orc.o2_rhs_boxes(NR, NA, BOXES, Wb, limiter=limiter, riemann=riemann, scale=scale,
                                 with_fluxes=True, divergence_cleaning=powell)
    op = hc.BoxOp2(hc.Shell(NR, NA), opts, BOXES)
    R = torch.zeros(op.count(1), dtype=torch.float64, device=device)
    F = [torch.zeros(op.count(2 + d), dtype=torch.float64, device=device) for d in range(3)]
    print("o2box case:", limiter, riemann, powell, scale, flush=True)
    op.apply(torch.from_numpy(Wb).to(device), R, F, scale=scale)
    torch.cuda.synchronize()
    err = rel_linf_per_var(per_var(R.cpu().numpy(), BOXES), per_var(ref, BOXES), axis_var=0)
    assert err.max() <= TOL, err
    for d in range(3):
        fb = [(p, lo, tuple(hi[a] + (a == d) for a in range(3))) for p, lo, hi in BOXES]
        ef = rel_linf_per_var(per_var(F[d].cpu().numpy(), fb), per_var(fref[d], fb), axis_var=0)
        assert ef.max() <= TOL, (d, ef)


@pytest.mark.gpu
def test_boxop2_equals_level_rhs_away_from_seams(device):
    opts, O, W = _stage(1, 2, 0)
    boxes = [BOXES[0], BOXES[1]]
    L = hc.Level(hc.Shell(NR, NA), opts)
    K = torch.zeros(L.valid_shape, dtype=torch.float64, device=device)
    L.stage_rhs(torch.from_numpy(np.ascontiguousarray(W)).to(device), K)
    op = hc.BoxOp2(hc.Shell(NR, NA), opts, boxes)
    R = torch.zeros(op.count(1), dtype=torch.float64, device=device)
    op.apply(torch.from_numpy(_cut(W, boxes)).to(device), R)
    torch.cuda.synchronize()
    Kc, Rc, off = K.cpu().numpy(), R.cpu().numpy(), 0
    for p, lo, hi in boxes:
        n = [hi[d] - lo[d] + 1 for d in range(3)]
        got = Rc[off:off + 8 * int(np.prod(n))].reshape(8, n[2], n[1], n[0])
        want = Kc[p, :, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
        assert rel_linf_per_var(got, want, axis_var=0).max() <= TOL
        off += got.size
