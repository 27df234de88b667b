"""GPU parity of the inner radial ghosts from boundary-file data
(hc_level_inner_ghosts_from_bc) with the oracle's restatement of the
PreStagePatch file branch (src/BoxOp_PreStage.H:161-317) and, at second
order, the ghost conversion to Cartesian primitives
(src/BoxOp_SecondOrderPreStage.H:194-214), within relative 1e-12 per
component."""
import numpy as np
import pytest
import torch

import heliocubed_amd as hc
import oracle as orc
from helpers import rel_linf_per_var
from test_oracle_bc_pins import _strip

R1 = 60.0


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.gpu
@pytest.mark.parametrize("order,nr,na", [(2, 12, 8), (2, 16, 16), (4, 16, 16), (4, 12, 20)])
def test_inner_ghosts_from_bc_match_oracle(device, order, nr, na):
    shell = hc.Shell(nr, na, 21.5, R1)
    opts = hc.Options(spatial_order=order, limiter=1, riemann_solver=2)
    L = hc.Level(shell, opts)
    g = L.ghost()
    bc = _strip(na, g, 31 + order)
    inner = torch.zeros((6, 8, na + 2 * g, na + 2 * g, g), dtype=torch.float64, device=device)
    L.inner_ghosts_from_bc(_dev(bc), inner)
    torch.cuda.synchronize()
    got = inner.cpu().numpy()
    ref = orc.inner_ghosts_from_bc(nr, na, bc, order=order, r1=R1)
    err = rel_linf_per_var(got, ref)
    print("bc ghosts err", order, err)
    assert np.isfinite(got).all() and err.max() <= 1e-12, err


@pytest.mark.gpu
@pytest.mark.parametrize("order,strength,time", [(2, 0.0, 0.0), (2, 20.0, 3.0e5), (4, 20.0, 1.0e6)])
def test_outflow_ghosts_match_oracle(device, order, strength, time):
    nr, na = 12, 16
    shell = hc.Shell(nr, na, 21.5, R1)
    L = hc.Level(shell, hc.Options(spatial_order=order, limiter=1, riemann_solver=2))
    g = L.ghost()
    inner = torch.zeros((6, 8, na + 2 * g, na + 2 * g, g), dtype=torch.float64, device=device)
    L.inner_ghosts_outflow(inner, strength, time)
    torch.cuda.synchronize()
    got = inner.cpu().numpy()
    ref = orc.inner_ghosts_outflow(nr, na, strength, time, order=order, r1=R1)
    err = rel_linf_per_var(got, ref)
    print("outflow ghosts err", order, strength, err)
    assert np.isfinite(got).all() and err.max() <= 1e-12, err
