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
    dt = 0.2 * (21.5 * np.pi / 2 / na) / 1.6
    uref = O.rk4(ju, inner, dt, 1)
    U = _dev(ju)
    L.rk4_advance(U, _dev(inner), dt, 1)
    torch.cuda.synchronize()
    got = U.cpu().numpy()
    errd = rel_linf_per_var(got - ju, uref - ju)
    print("mixed stage RHS err", errk, "increment err", errd, flush=True)
    assert errk.max() <= BAR, errk
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
    print(f"full-size Berta cone mixed stage RHS err ({time.time() - t0:.1f} s)", errk, flush=True)
    assert errk.max() <= BAR, errk
