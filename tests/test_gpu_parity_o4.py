"""GPU parity of the fourth-order box operator (O4 standard + Roe 8-wave).

The oracle restates src/BoxOp_EulerCubedSphereImplem.H:97-402 with
src/BoxOp_Limiters.H, src/BoxOp_RoeCubedSphere.H and src/MHD_Roe8Wave.H.
Both evaluate the same boxes from the same native conserved averages; the
GPU must match within relative L-infinity 1e-12 per variable.
"""
import numpy as np
import pytest
import torch

import heliocubed_amd as hc
import oracle as orc
from helpers import per_var as _per_var, rel_linf_per_var

TOL = 1e-12


CASES = [
    # (n_rad, n_ang, boxes, limiter, config)
    (32, 16, [(3, (8, 0, 0), (15, 7, 7)), (0, (16, 4, 8), (23, 11, 15))], 1, 3),
    (32, 16, [(1, (12, 0, 4), (19, 7, 11))], 0, 1),
    (32, 16, [(p, (8, 0, 8), (15, 7, 15)) for p in range(6)], 1, 0),
    # limiter 2 (Limiters_minmod with the metric cell widths), boxes at panel edges
    (32, 16, [(3, (8, 0, 0), (15, 7, 7)), (0, (4, 4, 8), (11, 11, 15)), (4, (20, 8, 2), (27, 15, 9))], 2, 3),
    (32, 16, [(2, (4, 4, 4), (11, 11, 11))], 2, 1),
]


@pytest.mark.gpu
@pytest.mark.parametrize("nr,na,boxes,limiter,config", CASES)
def test_o4_boxop_matches_oracle(device, nr, na, boxes, limiter, config):
    shell = hc.Shell(nr, na)
    opts = hc.Options(spatial_order=4, limiter=limiter, riemann_solver=2)
    U = hc.init_native_boxes(shell, opts, boxes, config=config, seed=5, quad=3)
    ref, Fref = orc.o4_rhs_boxes(nr, na, boxes, U, limiter=limiter, with_fluxes=True)
    op = hc.BoxOp(shell, opts, boxes)
    Ud = torch.from_numpy(U).to(device)
    R = torch.zeros(op.count(1), dtype=torch.float64, device=device)
    F = [torch.zeros(op.count(2 + d), dtype=torch.float64, device=device) for d in range(3)]
    op.apply(Ud, R, F)
    torch.cuda.synchronize()
    got = R.cpu().numpy()
    assert np.isfinite(got).all()
    err = rel_linf_per_var(_per_var(got, boxes), _per_var(ref, boxes), axis_var=0)
    print("o4 rhs rel err per var", err)
    assert err.max() <= TOL, err
    for d in range(3):
        fg = F[d].cpu().numpy()
        scale = np.abs(Fref[d]).max()
        assert np.abs(fg - Fref[d]).max() <= TOL * scale, d


@pytest.mark.gpu
def test_o4_rhs_accumulates(device):
    nr, na = 32, 16
    boxes = [(2, (8, 4, 4), (15, 11, 11))]
    shell = hc.Shell(nr, na)
    opts = hc.Options(spatial_order=4, limiter=1, riemann_solver=2)
    U = torch.from_numpy(hc.init_native_boxes(shell, opts, boxes, config=3, seed=1, quad=2)).to(device)
    op = hc.BoxOp(shell, opts, boxes)
    R1 = torch.zeros(op.count(1), dtype=torch.float64, device=device)
    op.apply(U, R1)
    R2 = torch.ones(op.count(1), dtype=torch.float64, device=device)
    op.apply(U, R2)
    torch.cuda.synchronize()
    assert torch.allclose(R2 - 1.0, R1, rtol=0, atol=1e-12 * R1.abs().max().item())


RUSANOV_CASES = [
    # (spatial_operator, limiter, config, eta_d)
    (1, 1, 3, 0.02),
    (2, 1, 3, 0.02),
    (2, 1, 1, 10.0),
    (1, 2, 3, 0.02),
]


@pytest.mark.gpu
@pytest.mark.parametrize("operator,limiter,config,eta_d", RUSANOV_CASES)
def test_o4_rusanov_boxop_matches_oracle(device, operator, limiter, config, eta_d):
    """Fourth-order operators with Rusanov dissipation (riemann_solver 1):
    primToMappedRusanovFlux (src/BoxOp_EulerCubedSphereImplem.H:536-733) and
    BertaPrimPointToMappedRusanovFlux (src/BoxOp_BertaCubedSphereImplem.H:1541-1730)."""
    nr, na = 32, 16
    boxes = [(3, (8, 0, 0), (15, 7, 7)), (0, (16, 4, 8), (23, 11, 15)), (5, (4, 8, 3), (11, 15, 10))]
    shell = hc.Shell(nr, na)
    opts = hc.Options(spatial_order=4, limiter=limiter, riemann_solver=1, spatial_operator=operator,
                      berta_eta_d=eta_d)
    U = hc.init_native_boxes(shell, opts, boxes, config=config, seed=9, quad=3)
    ref = orc.o4_rhs_boxes(nr, na, boxes, U, limiter=limiter, riemann=1, spatial_operator=operator, eta_d=eta_d)
    op = hc.BoxOp(shell, opts, boxes)
    R = torch.zeros(op.count(1), dtype=torch.float64, device=device)
    op.apply(torch.from_numpy(U).to(device), R)
    torch.cuda.synchronize()
    got = R.cpu().numpy()
    err = rel_linf_per_var(_per_var(got, boxes), _per_var(ref, boxes), axis_var=0)
    print("o4 rusanov rhs rel err per var", operator, err)
    assert err.max() <= TOL, err


# the smallest box the operator takes is 4 cells in each direction
RAGGED = [(0, (0, 0, 0), (3, 3, 3)), (1, (28, 12, 12), (31, 15, 15)), (2, (0, 5, 0), (3, 9, 4)),
          (3, (12, 0, 12), (18, 3, 15)), (4, (7, 7, 7), (10, 15, 10)), (5, (26, 0, 3), (31, 6, 11))]


@pytest.mark.gpu
@pytest.mark.parametrize("operator,riemann", [(1, 2), (2, 2), (1, 1)])
def test_ragged_boxes_match_oracle(device, operator, riemann):
    """Boxes of 4 x 4 x 4 up to 6 x 7 x 9 cells (odd and uneven sizes), at panel
    corners and at both radial boundaries, in one table: each matches the
    oracle's box RHS."""
    nr, na, r1 = 32, 16, 60.0
    shell = hc.Shell(nr, na, 21.5, r1)
    opts = hc.Options(spatial_order=4, limiter=1, riemann_solver=riemann, spatial_operator=operator)
    U = hc.init_native_boxes(shell, opts, RAGGED, config=3, seed=21, quad=2)
    ref = orc.o4_rhs_boxes(nr, na, RAGGED, U, riemann=riemann, spatial_operator=operator, r1=r1)
    op = hc.BoxOp(shell, opts, RAGGED)
    R = torch.zeros(op.count(1), dtype=torch.float64, device=device)
    op.apply(torch.from_numpy(U).to(device), R)
    torch.cuda.synchronize()
    got = R.cpu().numpy()
    assert np.isfinite(got).all()
    err = rel_linf_per_var(_per_var(got, RAGGED), _per_var(ref, RAGGED), axis_var=0)
    print("ragged boxes rel err per var", operator, riemann, err)
    assert err.max() <= TOL, err
