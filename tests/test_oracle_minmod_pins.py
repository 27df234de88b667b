"""Pins of the oracle's fourth-order minmod limiter (limiter 2,
Limiters_minmod / minmod_calc, src/BoxOp_Limiters.H:360-446, with the cell
widths of src/BoxOp_EulerCubedSphereImplem.H:220-240):

* minmod_calc known answers: linear data are reproduced at both faces, a local
  extremum or a flat side gives a zero slope, the smaller of two one-sided
  slopes is taken, and on cells of unequal widths data linear in the physical
  coordinate are reproduced exactly at the cell's edges (the width-weighted
  differences are the point slope);
* the widths: the radial width is r_hi - r_lo of the cell (closed form); an
  angular width divided by the cell's radial factor dVolr / adjDr_t (here from
  the closed-form radii) summed across a panel along the centre rows tends to
  the great-circle length pi/2 at second order;
* the operator: with limiter 2 the box RHS of a smooth state converges to the
  exact RHS (numpy quadrature of the closed-form map) at about second order in
  L1 (first order at the minmod's clipped extrema in L-inf).
"""
import numpy as np
import pytest

import exact_shell as ex
import oracle as orc


def test_minmod_known_answers():
    assert orc.minmod_cell(1.0, 2.0, 3.0, 1.0, 1.0, 1.0) == (2.5, 1.5)          # linear: exact faces
    assert orc.minmod_cell(1.0, 3.0, 2.0, 1.0, 1.0, 1.0) == (3.0, 3.0)          # extremum: slope 0
    assert orc.minmod_cell(1.0, 1.0, 2.0, 1.0, 1.0, 1.0) == (1.0, 1.0)          # flat side: slope 0
    assert orc.minmod_cell(0.0, 1.0, 3.0, 1.0, 1.0, 1.0) == (1.5, 0.5)          # smaller slope (behind)
    assert orc.minmod_cell(3.0, 2.0, 1.5, 1.0, 1.0, 1.0) == (1.75, 2.25)        # smaller slope (ahead)
    # equal magnitudes take the slope behind (B): same value either way
    assert orc.minmod_cell(0.0, 1.0, 2.0, 1.0, 1.0, 1.0) == (1.5, 0.5)


@pytest.mark.parametrize("widths", [(1.0, 2.0, 0.5), (0.3, 0.7, 1.9), (2.0, 2.0, 2.0)])
def test_minmod_exact_for_linear_data_on_unequal_cells(widths):
    dxb, dx, dxa = widths
    s, x0 = -1.7, 0.4
    xc = x0
    xb = xc - 0.5 * (dxb + dx)
    xa = xc + 0.5 * (dx + dxa)
    hi, lo = orc.minmod_cell(s * xb, s * xc, s * xa, dxb, dx, dxa)
    assert abs(hi - s * (xc + 0.5 * dx)) < 1e-14
    assert abs(lo - s * (xc - 0.5 * dx)) < 1e-14


def test_radial_width_is_the_cell_width():
    r0, r1 = 21.5, 215.0
    for n in (16, 32):
        for i in (0, 5, n - 1):
            w = orc.minmod_width(n, 16, 2, 0, i, 3, 9)
            assert abs(w - (r1 - r0) / n) < 1e-12 * r1


def test_angular_width_row_sums_converge_to_the_arc_length():
    r0, r1 = 21.5, 215.0
    errs = []
    for na in (16, 32, 64):
        nr = 8
        i = 3
        dxi = 1.0 / nr
        rl, rh = r0 + (r1 - r0) * i * dxi, r0 + (r1 - r0) * (i + 1) * dxi
        radial = ((rh ** 3 - rl ** 3) / 3.0 / dxi) / ((rh ** 2 - rl ** 2) / dxi / 2.0)
        for d in (1, 2):
            # the two rows of cells beside the panel's centre line
            tot = 0.0
            for k in (na // 2 - 1, na // 2):
                for j in range(na):
                    jj, kk = (j, k) if d == 1 else (k, j)
                    tot += orc.minmod_width(nr, na, 1, d, i, jj, kk) / radial
            errs.append(abs(tot / 2 - np.pi / 2))
    errs = np.array(errs).reshape(3, 2)
    rates = np.log2(errs[:-1] / errs[1:])
    print("errors", errs, "rates", rates)
    assert (rates > 1.8).all(), rates
    assert (errs[-1] < 2e-3).all(), errs


def _box(n):
    lo = (3 * n // 8,) * 3
    hi = (5 * n // 8 - 1,) * 3
    return lo, hi


def test_limiter2_box_rhs_converges_to_the_exact_rhs():
    l1, linf = [], []
    p = 4
    for n in (16, 32, 64):
        sh = ex.Shell(n, n)
        lo, hi = _box(n)
        rng_ = [np.arange(lo[a] - 5, hi[a] + 6) for a in range(3)]
        U = sh.native_averages(p, *rng_)
        got = orc.o4_rhs_boxes(n, n, [(p, lo, hi)], U.ravel(), limiter=2).reshape(
            8, *(hi[a] - lo[a] + 1 for a in (2, 1, 0)))
        want = sh.exact_rhs(p, lo, hi)
        scale = np.abs(want).max(axis=(1, 2, 3))
        linf.append(np.abs(got - want).max(axis=(1, 2, 3)) / scale)
        l1.append(np.abs(got - want).mean(axis=(1, 2, 3)) / scale)
    l1, linf = np.array(l1), np.array(linf)
    l1_rates = np.log2(l1[:-1] / l1[1:])
    linf_rates = np.log2(linf[:-1] / linf[1:])
    print("L1", l1, "L1 rates", l1_rates, "Linf rates", linf_rates)
    assert (l1_rates[-1] > 1.6).all(), l1_rates
    assert (linf_rates[-1] > 0.8).all(), linf_rates
