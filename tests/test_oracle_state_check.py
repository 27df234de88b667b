"""CPU pins of the conserved-state check (validConservedState,
src/BoxOp_EulerCubedSphereImplem.H:813-876): a state is invalid when any
value is not finite or any density is not > 0; other signs are allowed.
The expectations are constructed cell by cell from that predicate."""
import numpy as np
import pytest

import oracle as orc

SHAPE = (6, 8, 4, 4, 5)


def _valid():
    rng = np.random.default_rng(2)
    u = rng.standard_normal(SHAPE)
    u[:, 0] = 1.0 + rng.random(SHAPE[2:])
    return u


def _flat(p, c, k, j, i):
    return int(np.ravel_multi_index((p, c, k, j, i), SHAPE))


def test_valid_state_passes_with_negative_momentum_and_energy():
    u = _valid()
    u[:, 1:] = -np.abs(u[:, 1:])
    u[3, 0, 1, 2, 3] = 5e-324            # smallest positive density (denormal) is valid
    assert orc.first_invalid(u) == -1


@pytest.mark.parametrize("comp,value", [(0, 0.0), (0, -0.0), (0, -1e-300), (0, np.nan), (4, np.inf),
                                        (7, -np.inf), (2, np.nan)])
def test_single_bad_value_is_found(comp, value):
    u = _valid()
    u[2, comp, 3, 1, 4] = value
    assert orc.first_invalid(u) == _flat(2, comp, 3, 1, 4)


def test_first_in_scan_order():
    u = _valid()
    u[4, 6, 0, 0, 0] = np.nan
    u[1, 3, 3, 3, 4] = np.inf
    u[1, 3, 2, 0, 1] = np.nan            # same panel and component, earlier (k, j, i)
    assert orc.first_invalid(u) == _flat(1, 3, 2, 0, 1)
