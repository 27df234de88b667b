"""GPU parity of the O2 operator (PLM + Rusanov / Roe) against the CPU oracle.

The oracle restates src/BoxOp_SecondOrder.H, src/MHD_SecondOrderFlux.H,
src/MHD_Roe8Wave.H and src/BoxOp_SecondOrderPreStage.H of the reference; the
GPU path must match it within relative L-infinity 1e-12 per variable.
"""
import numpy as np
import pytest
import torch

import heliocubed_amd as hc
import oracle as orc
from helpers import cfl_dt, rel_linf_per_var

TOL = 1e-12

CASES = [
    # (n_rad, n_ang, limiter, riemann)
    (16, 8, 1, 1),
    (32, 32, 1, 1),     # config 0: 6 x 32^3, PLM minmod + Rusanov
    (12, 12, 3, 2),     # MC limiter + Roe 8-wave
    (10, 6, 0, 2),      # unlimited + Roe
]


def _setup(nr, na, limiter, riemann, seed=20261203):
    shell = hc.Shell(nr, na)
    opts = hc.Options(limiter=limiter, riemann_solver=riemann)
    ju, inner = hc.init_state(shell, opts, 0, seed)
    O = orc.Oracle(nr, na, limiter=limiter, riemann=riemann)
    return shell, opts, ju, inner, O


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.gpu
@pytest.mark.parametrize("nr,na,limiter,riemann", CASES)
def test_prepare_stage_matches_oracle(device, nr, na, limiter, riemann):
    shell, opts, ju, inner, O = _setup(nr, na, limiter, riemann)
    L = hc.Level(shell, opts)
    W = torch.zeros(L.stage_shape, dtype=torch.float64, device=device)
    L.prepare_stage(_dev(ju), _dev(inner), W)
    torch.cuda.synchronize()
    Wref = O.prepare(ju, inner)
    # Every stage cell the operator reads: valid, angular ghosts and radial ghosts
    # (corner ghost columns included).
    err = rel_linf_per_var(W.cpu().numpy(), Wref)
    assert err.max() <= TOL, err


@pytest.mark.gpu
@pytest.mark.parametrize("nr,na,limiter,riemann", CASES)
def test_stage_rhs_matches_oracle(device, nr, na, limiter, riemann):
    shell, opts, ju, inner, O = _setup(nr, na, limiter, riemann)
    Wref = O.prepare(ju, inner)
    kref, Fref = O.rhs(Wref, with_fluxes=True)
    L = hc.Level(shell, opts)
    W = _dev(Wref)
    K = torch.zeros(L.valid_shape, dtype=torch.float64, device=device)
    F = [torch.zeros(L.face_shape(d), dtype=torch.float64, device=device) for d in range(3)]
    L.stage_rhs(W, K, fluxes=F)
    torch.cuda.synchronize()
    err = rel_linf_per_var(K.cpu().numpy(), kref)
    assert err.max() <= TOL, err
    # Face fluxes per variable, scaled by the same variable's radial-face flux
    # magnitude: through angular faces of a radial wind the mass and energy
    # fluxes are nearly zero, so their own maximum would amplify the 1e-15
    # difference between the two independent face-normal computations.
    for d in range(3):
        diff = np.abs(F[d].cpu().numpy() - Fref[d]).max(axis=(0, 2, 3, 4))
        scale = np.maximum(np.abs(Fref[d]).max(axis=(0, 2, 3, 4)), np.abs(Fref[0]).max(axis=(0, 2, 3, 4)))
        assert (diff <= TOL * scale).all(), (d, diff / scale)


@pytest.mark.gpu
@pytest.mark.parametrize("nr,na,limiter,riemann", CASES[:3])
def test_rk4_step_matches_oracle(device, nr, na, limiter, riemann):
    shell, opts, ju, inner, O = _setup(nr, na, limiter, riemann)
    dt = cfl_dt(na)
    uref = O.rk4(ju, inner, dt, 2)
    L = hc.Level(shell, opts)
    U = _dev(ju)
    L.rk4_advance(U, _dev(inner), dt, 2)
    torch.cuda.synchronize()
    err = rel_linf_per_var(U.cpu().numpy(), uref)
    assert err.max() <= TOL, err
    # the increment itself, not only the state, must match
    errd = rel_linf_per_var(U.cpu().numpy() - ju, uref - ju)
    assert errd.max() <= 1e-10, errd


@pytest.mark.gpu
def test_rhs_conserves_to_roundoff(device):
    """Sum of the RHS over the shell equals minus the net radial boundary flux."""
    nr, na = 24, 16
    shell, opts, ju, inner, O = _setup(nr, na, 1, 2)
    L = hc.Level(shell, opts)
    W = torch.zeros(L.stage_shape, dtype=torch.float64, device=device)
    L.prepare_stage(_dev(ju), _dev(inner), W)
    K = torch.zeros(L.valid_shape, dtype=torch.float64, device=device)
    F = [torch.zeros(L.face_shape(d), dtype=torch.float64, device=device) for d in range(3)]
    L.stage_rhs(W, K, fluxes=F)
    k = K.cpu().numpy()
    F0 = F[0].cpu().numpy()
    vol = (1.0 / nr) * (1.0 / na) ** 2
    total = k.sum(axis=(0, 2, 3, 4)) * vol
    boundary = (F0[..., nr].sum(axis=(0, 2, 3)) - F0[..., 0].sum(axis=(0, 2, 3))) * (1.0 / na) ** 2
    scale = np.abs(F0).sum(axis=(0, 2, 3, 4)) * (1.0 / na) ** 2
    assert np.all(np.abs(total + boundary) <= 1e-13 * scale), (total + boundary) / scale


@pytest.mark.gpu
def test_second_call_smaller_level(device):
    """Two levels of different size in one process give independent results."""
    for nr, na in [(16, 12), (8, 6)]:
        shell, opts, ju, inner, O = _setup(nr, na, 1, 1, seed=7)
        L = hc.Level(shell, opts)
        U = _dev(ju)
        dt = cfl_dt(na)
        L.rk4_advance(U, _dev(inner), dt, 1)
        torch.cuda.synchronize()
        err = rel_linf_per_var(U.cpu().numpy(), O.rk4(ju, inner, dt, 1))
        assert err.max() <= TOL, (nr, na, err)


@pytest.mark.gpu
@pytest.mark.parametrize("riemann", [1, 2])
def test_mixed_precision_error_vs_fp64_oracle(device, riemann):
    """Mixed precision (FP32 slopes and Riemann dissipation) stays within
    1e-6 relative L-infinity per variable of the FP64 oracle after one step."""
    nr, na = 24, 16
    shell = hc.Shell(nr, na)
    opts = hc.Options(limiter=1, riemann_solver=riemann, precision=1)
    ju, inner = hc.init_state(shell, opts, 0, 11)
    O = orc.Oracle(nr, na, limiter=1, riemann=riemann)
    dt = cfl_dt(na)
    ref = O.rk4(ju, inner, dt, 1)
    L = hc.Level(shell, opts)
    U = _dev(ju)
    L.rk4_advance(U, _dev(inner), dt, 1)
    torch.cuda.synchronize()
    got = U.cpu().numpy()
    err = rel_linf_per_var(got, ref)
    errd = rel_linf_per_var(got - ju, ref - ju)
    print("mixed riemann", riemann, "state err", err, "increment err", errd)
    assert errd.max() <= 1e-6, errd
    assert err.max() <= 1e-6, err
