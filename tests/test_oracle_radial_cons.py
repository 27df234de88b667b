"""Pin of the oracle's conservation sums (radialCons as StageLevel calls it,
src/BoxOp_EulerCubedSphereImplem.H:1010, with the stage weight times dt;
Proto's sum restated as the net radial boundary inflow): over RK4 steps of a
fourth-order level the sums equal the change of sum(<JU>) dxi^3 per component
(the seam fluxes cancel after reconciliation, the radial boundaries are the
only open faces), for both operators."""
import numpy as np
import pytest

import heliocubed_amd as hc
import oracle as orc


@pytest.mark.parametrize("operator,config", [(1, 3), (2, 4)])
def test_radial_cons_is_the_change_of_the_totals(operator, config):
    nr, na = 16, 16
    shell = hc.Shell(nr, na, 21.5, 60.0)
    opts = hc.Options(spatial_order=4, limiter=1, riemann_solver=2, spatial_operator=operator)
    ju, inner = hc.init_level4(shell, opts, config, 5, 2)
    O = orc.Oracle4(nr, na, r1=60.0, spatial_operator=operator)
    dt = 0.2 * (21.5 * np.pi / 2 / na) / 1.6
    u, cons = O.rk4_cons(ju, inner, dt, 2)
    change = (u - ju).sum(axis=(0, 2, 3, 4)) / (nr * na * na)
    scale = np.abs(u - ju).sum(axis=(0, 2, 3, 4)) / (nr * na * na)
    print("cons", cons, "change", change)
    assert np.abs(cons - change).max() <= 1e-12 * scale.max()
    assert (np.abs(cons - change) <= 1e-11 * scale).all()
    assert np.abs(cons[0]) > 1e-3 * scale[0]      # the boundaries do carry mass
