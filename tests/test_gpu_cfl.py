"""GPU parity of the level CFL time step (hc_level_dt_cfl) against the CPU
oracle: dtCFLSecondOrderHost (src/BoxOp_SecondOrderCFL.H:99-210) at second
order, dtCFL + dtInv + f_speedmax (src/BoxOp_EulerCubedSphereImplem.H:878-954,
1013-1131; src/BoxOp_EulerCubedSphere.H:127-194) at fourth order. The result
is a maximum over cells of the same per-cell arithmetic, so it must agree to
relative 1e-13; the full-size case covers the 6 x 128 x 64 x 64 shell."""
import numpy as np
import pytest
import torch

import heliocubed_amd as hc
import oracle as orc

TOL = 1e-13


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.gpu
@pytest.mark.parametrize("nr,na,config,cone", [(12, 8, 0, 0.0), (16, 16, 1, 0.0), (12, 8, 2, 0.0), (16, 16, 1, 0.02)])
def test_o2_dt_cfl_matches_oracle(device, nr, na, config, cone):
    shell = hc.Shell(nr, na)
    opts = hc.Options(limiter=1, riemann_solver=2)
    ju, _ = hc.init_state(shell, opts, config, 7)
    L = hc.Level(shell, opts)
    got = L.dt_cfl(_dev(ju), cone)
    ref = orc.dt_cfl(nr, na, ju, order=2, cone_density=cone)
    print("o2 dt_cfl", config, cone, got, ref)
    assert np.isfinite(got) and abs(got - ref) <= TOL * ref


@pytest.mark.gpu
@pytest.mark.parametrize("nr,na,config,operator", [(16, 16, 3, 1), (12, 20, 0, 1), (16, 16, 4, 2), (20, 16, 1, 2)])
def test_o4_dt_cfl_matches_oracle(device, nr, na, config, operator):
    r1 = 60.0
    shell = hc.Shell(nr, na, 21.5, r1)
    opts = hc.Options(spatial_order=4, limiter=1, riemann_solver=2, spatial_operator=operator)
    ju, _ = hc.init_level4(shell, opts, config, 5, 3)
    L = hc.Level(shell, opts)
    U = _dev(ju)
    got = L.dt_cfl(U)
    ref = orc.dt_cfl(nr, na, ju, order=4, r1=r1)
    # the fourth-order time step takes no cone bound
    assert L.dt_cfl(U, 1e-3) == got
    print("o4 dt_cfl", config, operator, got, ref)
    assert np.isfinite(got) and abs(got - ref) <= TOL * ref


@pytest.mark.gpu
def test_dt_cfl_full_shell_and_scaling(device):
    """6 x 128 x 64 x 64 shell (the production shape): parity with the oracle,
    and the exact scaling dt(rho, lambda v, lambda B, lambda^2 p) = dt / lambda
    for lambda a power of two (every product scales exactly)."""
    nr, na = 128, 64
    shell = hc.Shell(nr, na)
    opts = hc.Options(spatial_order=4, limiter=1, riemann_solver=2)
    ju, _ = hc.init_level4(shell, opts, 3, 20261204, 1)
    L = hc.Level(shell, opts)
    got = L.dt_cfl(_dev(ju))
    ref = orc.dt_cfl(nr, na, ju, order=4)
    print("full shell dt_cfl", got, ref)
    assert abs(got - ref) <= TOL * ref
    lam = 4.0
    scaled = ju.copy()
    scaled[:, 1:4] *= lam
    scaled[:, 5:8] *= lam
    scaled[:, 4] *= lam * lam
    assert L.dt_cfl(_dev(scaled)) == got / lam


def test_dt_cfl_rejects_bad_arguments_without_device():
    import ctypes
    out = ctypes.c_double()
    assert hc.lib().hc_level_dt_cfl(None, None, 0.0, ctypes.byref(out), None) != 0
    assert b"null" in hc.lib().hc_last_error()
