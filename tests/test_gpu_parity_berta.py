"""GPU parity of the Berta WENOZ fourth-order operator (spatial_operator 2):
Jameson detector, WENOZ with WENO3 / PLM fallback, theta-aware face averages,
separable-metric mapped products and theta-aware Roe dissipation
(src/BoxOp_BertaCubedSphereImplem.H:1737-2076), as a box table and over the
level with seam reconciliation and RK4, against the CPU oracle within
relative L-infinity 1e-12 per variable; mixed precision (FP32 reconstruction
weights and Roe eigen-system) within 1e-6 of the FP64 state."""
import numpy as np
import pytest
import torch

import heliocubed_amd as hc
import oracle as orc
from helpers import per_var as _per_var, rel_linf_per_var

TOL = 1e-12
R1 = 60.0


CASES = [
    # (n_rad, n_ang, boxes, config, eta_d, fallback)
    (32, 16, [(3, (8, 0, 0), (15, 7, 7)), (0, (16, 4, 8), (23, 11, 15))], 3, 0.02, 0),
    (32, 16, [(1, (12, 0, 4), (19, 7, 11))], 1, 0.02, 1),
    (32, 16, [(p, (8, 0, 8), (15, 7, 15)) for p in range(6)], 0, 10.0, 0),
    (32, 16, [(4, (6, 3, 2), (13, 10, 9))], 3, 1e-9, 0),
]


@pytest.mark.gpu
@pytest.mark.parametrize("nr,na,boxes,config,eta_d,fallback", CASES)
def test_berta_boxop_matches_oracle(device, nr, na, boxes, config, eta_d, fallback):
    shell = hc.Shell(nr, na)
    opts = hc.Options(spatial_order=4, limiter=1, riemann_solver=2, spatial_operator=2, berta_eta_d=eta_d,
                      berta_fallback=fallback)
    U = hc.init_native_boxes(shell, opts, boxes, config=config, seed=5, quad=3)
    ref, Fref = orc.o4_rhs_boxes(nr, na, boxes, U, with_fluxes=True, spatial_operator=2, eta_d=eta_d,
                                 fallback=fallback)
    op = hc.BoxOp(shell, opts, boxes)
    Ud = torch.from_numpy(U).to(device)
    R = torch.zeros(op.count(1), dtype=torch.float64, device=device)
    F = [torch.zeros(op.count(2 + d), dtype=torch.float64, device=device) for d in range(3)]
    op.apply(Ud, R, F)
    torch.cuda.synchronize()
    got = R.cpu().numpy()
    assert np.isfinite(got).all()
    err = rel_linf_per_var(_per_var(got, boxes), _per_var(ref, boxes), axis_var=0)
    print("berta rhs rel err per var", err)
    assert err.max() <= TOL, err
    for d in range(3):
        fg = F[d].cpu().numpy()
        scale = np.abs(Fref[d]).max()
        assert np.abs(fg - Fref[d]).max() <= TOL * scale, d


def _setup(nr, na, config, seed=3, precision=0, eta_d=0.02):
    shell = hc.Shell(nr, na, 21.5, R1)
    opts = hc.Options(spatial_order=4, limiter=1, riemann_solver=2, precision=precision, spatial_operator=2,
                      berta_eta_d=eta_d)
    ju, inner = hc.init_level4(shell, opts, config, seed, 3)
    return shell, opts, ju, inner, orc.Oracle4(nr, na, r1=R1, spatial_operator=2, eta_d=eta_d)


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.gpu
@pytest.mark.parametrize("nr,na,config,eta_d", [(16, 16, 3, 0.02), (12, 20, 1, 0.02), (16, 16, 0, 10.0),
                                               (24, 16, 5, 0.02)])
def test_berta_level_rhs_matches_oracle(device, nr, na, config, eta_d):
    shell, opts, ju, inner, O = _setup(nr, na, config, eta_d=eta_d)
    W = O.prepare(ju, inner)
    ref = O.rhs(W)
    L = hc.Level(shell, opts)
    K = torch.zeros(L.valid_shape, dtype=torch.float64, device=device)
    L.stage_rhs(_dev(W), K)
    torch.cuda.synchronize()
    err = rel_linf_per_var(K.cpu().numpy(), ref)
    print("berta level rhs err", err)
    assert err.max() <= TOL, err


@pytest.mark.gpu
def test_berta_rk4_step_matches_oracle(device):
    nr, na = 16, 16
    shell, opts, ju, inner, O = _setup(nr, na, 3)
    dt = 0.2 * (21.5 * np.pi / 2 / na) / 1.6
    ref = O.rk4(ju, inner, dt, 1)
    L = hc.Level(shell, opts)
    U = _dev(ju)
    L.rk4_advance(U, _dev(inner), dt, 1)
    torch.cuda.synchronize()
    got = U.cpu().numpy()
    err = rel_linf_per_var(got, ref)
    errd = rel_linf_per_var(got - ju, ref - ju)
    print("berta rk4 err", err, "increment", errd)
    assert err.max() <= TOL, err
    assert errd.max() <= 1e-10, errd


@pytest.mark.gpu
def test_berta_mixed_precision_error(device):
    """FP32 WENO weights, minmod and Roe eigen-system vs the FP64 oracle, one RK4 step."""
    nr, na = 16, 16
    shell, opts, ju, inner, O = _setup(nr, na, 3, precision=1)
    dt = 0.2 * (21.5 * np.pi / 2 / na) / 1.6
    ref = O.rk4(ju, inner, dt, 1)
    L = hc.Level(shell, opts)
    U = _dev(ju)
    L.rk4_advance(U, _dev(inner), dt, 1)
    torch.cuda.synchronize()
    got = U.cpu().numpy()
    err = rel_linf_per_var(got, ref)
    errd = rel_linf_per_var(got - ju, ref - ju)
    print("berta mixed state err", err, "increment err", errd)
    assert errd.max() <= 1e-6, errd
    assert err.max() <= 1e-6, err


@pytest.mark.gpu
@pytest.mark.parametrize("operator", [1, 2])
def test_mixed_precision_low_beta_cone(device, operator):
    """Low-beta cone (p = 1e-4 |B|^2/2): mixed precision evaluates such faces'
    Roe eigen-system in FP64, so the state stays within 1e-6 of the FP64 oracle."""
    nr, na = 16, 16
    shell = hc.Shell(nr, na, 21.5, R1)
    opts = hc.Options(spatial_order=4, limiter=1, riemann_solver=2, precision=1, spatial_operator=operator)
    ju, inner = hc.init_level4(shell, opts, 4, 5, 3)
    O = orc.Oracle4(nr, na, r1=R1, spatial_operator=operator)
    dt = 0.2 * (21.5 * np.pi / 2 / na) / 1.6
    ref = O.rk4(ju, inner, dt, 1)
    L = hc.Level(shell, opts)
    U = _dev(ju)
    L.rk4_advance(U, _dev(inner), dt, 1)
    torch.cuda.synchronize()
    got = U.cpu().numpy()
    err = rel_linf_per_var(got, ref)
    errd = rel_linf_per_var(got - ju, ref - ju)
    # the FP64 oracle's own increment change from a 1e-9 relative input
    # perturbation: the band of branch flips inside the cone
    rng = np.random.default_rng(7)
    up = O.rk4(ju * (1 + 1e-9 * (2 * rng.random(ju.shape) - 1)), inner, dt, 1)
    sens = rel_linf_per_var(up - ju, ref - ju)
    print("mixed low-beta state err", operator, err, "increment err", errd, "FP64 sensitivity", sens)
    assert errd.max() <= max(1e-6, 2 * sens.max()), (errd, sens)
    assert err.max() <= 1e-6, err
