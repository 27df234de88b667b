"""GPU parity of the conserved-state check (hc_level_valid_state) with the
oracle's validConservedState restatement (src/BoxOp_EulerCubedSphereImplem.H:813-876):
the same verdict and the same first offending index, on small levels with
planted bad values and on the 6 x 128 x 64 x 64 shell."""
import numpy as np
import pytest
import torch

import heliocubed_amd as hc
import oracle as orc


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.gpu
@pytest.mark.parametrize("plant", [[], [(0, 0, 0, 0, 0, 0.0)], [(5, 7, 7, 7, 11, np.nan)],
                                   [(3, 4, 2, 5, 6, np.inf), (2, 0, 7, 1, 0, -0.0)],
                                   [(1, 1, 0, 0, 3, -np.inf), (1, 1, 0, 0, 2, np.nan), (4, 0, 3, 3, 3, -2.0)]])
def test_state_check_matches_oracle(device, plant):
    nr, na = 12, 8
    shell = hc.Shell(nr, na)
    opts = hc.Options(limiter=1, riemann_solver=2)
    ju, _ = hc.init_state(shell, opts, 1, 3)
    ju[:, 1:4] *= -1.0                       # signs of momentum do not matter
    for p, c, k, j, i, val in plant:
        ju[p, c, k, j, i] = val
    L = hc.Level(shell, opts)
    ok, first = L.valid_state(_dev(ju))
    ref = orc.first_invalid(ju)
    print("state check", plant, ok, first, ref)
    assert first == ref and ok == (ref < 0)


@pytest.mark.gpu
def test_state_check_full_shell(device):
    nr, na = 128, 64
    shell = hc.Shell(nr, na)
    opts = hc.Options(spatial_order=4, limiter=1, riemann_solver=2)
    ju, _ = hc.init_level4(shell, opts, 3, 9, 1)
    L = hc.Level(shell, opts)
    U = _dev(ju)
    assert L.valid_state(U) == (True, -1)
    last = ju.size - 1
    U.view(-1)[last] = float("nan")
    assert L.valid_state(U) == (False, last)
    x = int(np.ravel_multi_index((2, 0, 40, 17, 99), ju.shape))
    U.view(-1)[x] = 0.0
    assert L.valid_state(U) == (False, x)
