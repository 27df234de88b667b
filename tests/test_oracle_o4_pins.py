"""CPU pins of the fourth-order oracle operator.

* Face averages of det A, det(A) A and the cofactor rows: 8-point
  Gauss-Legendre (oracle) against the product's closed forms / 12-point rule.
* The radial wind with a radial field has no induction: the B rows of the RHS
  vanish to round-off (v x B = 0 exactly for radial v and B).
* Fourth-order convergence of the RHS on a smooth state (the reference's
  convergence methodology, exec/main.cpp:539-603: coarsen the fine solution
  and compare consecutive resolutions).
"""
import ctypes

import numpy as np
import pytest

import heliocubed_amd as hc
import oracle as orc


def test_face_averages_match_product_geometry():
    na = 16
    cfg = orc.OrcCfg(8, na, 5, 21.5, 215.0, 0.0, 5 / 3, 1, 2, 0.3)
    rng = np.random.default_rng(1)
    for _ in range(40):
        p, d = int(rng.integers(6)), int(rng.integers(3))
        j, k = int(rng.integers(-5, na + 5)), int(rng.integers(-5, na + 5))
        det = np.zeros(1); dA = np.zeros(9); n = np.zeros(3)
        orc.lib().orc_face_average(ctypes.byref(cfg), p, d, j, k, 0, det.ctypes.data)
        orc.lib().orc_face_average(ctypes.byref(cfg), p, d, j, k, 1, dA.ctypes.data)
        orc.lib().orc_face_average(ctypes.byref(cfg), p, d, j, k, 2, n.ctypes.data)
        e = np.zeros(3)
        hc.lib().hc_geom_face_normal(na, p, d, j, k, e.ctypes.data)
        assert np.abs(n - e).max() <= 1e-13 * np.abs(e).max()
        if d == 0:
            assert abs(det[0] - hc.lib().hc_geom_cell_detA(na, j, k)) <= 1e-13 * det[0]
        # det(A) A against a brute-force midpoint rule (independent, low accuracy)
        h = 1.0 / na
        M = 200
        acc = np.zeros(9)
        for s in range(M):
            t = (s + 0.5) / M
            if d == 0:
                continue
            a = -np.pi / 4 + (j + (t if d == 2 else 0.0)) * h * np.pi / 2
            b = -np.pi / 4 + (k + (t if d == 1 else 0.0)) * h * np.pi / 2
            A = orc.basis(p, a, b)
            acc += np.linalg.det(A) * A.ravel() / M
        if d != 0:
            assert np.abs(acc - dA).max() <= 1e-5 * np.abs(dA).max()


def _o4_boxes_rhs(nr, na, boxes, config, quad=4, limiter=0):
    shell = hc.Shell(nr, na)
    opts = hc.Options(spatial_order=4, limiter=limiter, riemann_solver=2)
    U = hc.init_native_boxes(shell, opts, boxes, config=config, seed=3, quad=quad)
    return orc.o4_rhs_boxes(nr, na, boxes, U, limiter=limiter)


def test_radial_field_radial_flow_has_no_induction():
    nr, na = 32, 16
    boxes = [(3, (6, 2, 2), (13, 9, 9)), (5, (10, 6, 3), (17, 13, 10))]
    r = _o4_boxes_rhs(nr, na, boxes, config=2, limiter=1)
    off = 0
    for _, lo, hi in boxes:
        n = int(np.prod([hi[d] - lo[d] + 1 for d in range(3)]))
        blk = r[off:off + 8 * n].reshape(8, n)
        off += 8 * n
        assert np.abs(blk[5:8]).max() <= 1e-9 * np.abs(blk[0:5]).max()


def test_o4_rhs_converges_at_fourth_order():
    """Smooth monopole wind without pulse noise (config 1), unlimited:
    RHS at three resolutions on the same physical box, finer ones averaged
    onto the coarse cells; the error ratio between levels approaches 16."""
    errs = []
    fields = []
    for lev, n in enumerate((8, 16, 32)):
        box = [(3, (n // 2, n // 4, n // 4), (n // 2 + n // 4 - 1, n // 2 - 1, n // 2 - 1))]
        r = _o4_boxes_rhs(4 * n, n, box, config=1, quad=5)
        m = n // 4
        fields.append(r.reshape(8, m, m, m))
    def coarsen(a):
        c, m = a.shape[0], a.shape[1] // 2
        return a.reshape(c, m, 2, m, 2, m, 2).mean(axis=(2, 4, 6))
    e0 = np.abs(fields[0] - coarsen(fields[1])).max(axis=(1, 2, 3))
    e1 = np.abs(fields[1] - coarsen(fields[2])).max(axis=(1, 2, 3))
    rate = np.log2(e0[[0, 1, 4]] / e1[[0, 1, 4]])
    print("O4 convergence rates (rho, mom_r, E):", rate)
    assert (rate > 3.3).all(), rate


def test_o4_prepare_converges_at_fourth_order():
    """JU (J-weighted averages) -> native averages: the error against the exact
    native averages (quadrature of the analytic state) falls by ~16 per
    refinement on valid cells and on the seam ghost cells."""
    errs = []
    for na in (16, 32):
        nr = na
        shell = hc.Shell(nr, na, 21.5, 60.0)
        opts = hc.Options(spatial_order=4, limiter=1, riemann_solver=2)
        ju, inner = hc.init_level4(shell, opts, 1, 3, 5)
        W = orc.Oracle4(nr, na, r1=60.0).prepare(ju, inner)
        boxes = [(p, (0, 0, 0), (nr - 1, na - 1, na - 1)) for p in range(6)]
        ex = hc.init_native_boxes(shell, opts, boxes, config=1, seed=3, quad=5).reshape(W.shape)
        g = 5
        sl = np.s_[:, :, g - 2:g + na + 2, g - 2:g + na + 2, g:g + nr]
        d = np.abs(W[sl] - ex[sl])
        errs.append(np.array([d[:, c].max() for c in (0, 1, 4)]))
    rate = np.log2(errs[0] / errs[1])
    print("prepare convergence rates (rho, mom_r, E):", rate)
    assert (rate > 3.4).all(), rate


def test_o4_rusanov_matches_roe_when_face_states_agree():
    """Unlimited fourth-order face states are equal on both sides, so the
    mapped Rusanov jump (primToMappedRusanovFlux) and the Roe jump both vanish:
    the two solvers give the same RHS bit for bit."""
    nr, na = 32, 16
    boxes = [(3, (8, 2, 2), (15, 9, 9)), (0, (16, 4, 8), (23, 11, 15))]
    shell = hc.Shell(nr, na)
    opts = hc.Options(spatial_order=4, limiter=0, riemann_solver=2)
    U = hc.init_native_boxes(shell, opts, boxes, config=3, seed=4, quad=3)
    roe = orc.o4_rhs_boxes(nr, na, boxes, U, limiter=0, riemann=2)
    rus = orc.o4_rhs_boxes(nr, na, boxes, U, limiter=0, riemann=1)
    assert np.array_equal(roe, rus)


@pytest.mark.parametrize("op", [1, 2])
def test_o4_rusanov_and_roe_converge_together(op):
    """Both dissipations vanish at high order on smooth data (limited standard
    operator / Berta WENOZ with the detector off): the RHS difference between
    the Rusanov and Roe solvers is round-off for the standard operator and
    falls faster than h^3 for Berta."""
    diffs = []
    for n in (8, 16, 32):
        box = [(1, (n // 2, n // 4, n // 4), (n // 2 + n // 4 - 1, n // 2 - 1, n // 2 - 1))]
        shell = hc.Shell(4 * n, n)
        opts = hc.Options(spatial_order=4, limiter=1, riemann_solver=2)
        U = hc.init_native_boxes(shell, opts, box, config=1, seed=3, quad=5)
        kw = dict(limiter=1, spatial_operator=op, eta_d=10.0)
        a = orc.o4_rhs_boxes(4 * n, n, box, U, riemann=2, **kw).reshape(8, -1)
        b = orc.o4_rhs_boxes(4 * n, n, box, U, riemann=1, **kw).reshape(8, -1)
        diffs.append(np.abs(a - b).max(axis=1)[[0, 1, 4]] / np.abs(a).max())
    print("Rusanov - Roe differences relative to the largest RHS entry", diffs)
    if op == 1:
        # the limiter leaves this smooth state alone, so both jumps are round-off
        assert all((d < 1e-11).all() for d in diffs), diffs
    else:
        r = np.log2(diffs[1] / diffs[2])
        assert (r > 3.0).all(), r
