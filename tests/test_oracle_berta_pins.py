"""CPU pins of the Berta WENOZ operator of the oracle (oracle/hc_oracle.cpp).

Reconstruction (src/BoxOp_BertaCubedSphereImplem.H:747-991):
* WENOZ reproduces quadratics exactly (equal smoothness indicators give the
  ideal weights) and quartics that are even about the upwind cell (beta0 =
  beta2, tau5 = 0), i.e. the ideal weights are the 5-point Lagrange
  interpolant (3, -20, 90, 60, -5)/128; it converges at fifth order on smooth
  data and takes the smooth side at a step to 1e-30.
* WENO3 reproduces linear data and converges at third order; PLM (minmod3,
  theta = 1.5) reproduces linear data and is flat at an extremum.
Detector (:346-507): a radial density step flags exactly the two cells whose
3-point stencil straddles it, with the closed-form indicator 1/5 and 1/7; the
magnetic indicator is the 3-point ratio of |B|^2/2; the three directions combine
in the 2-norm (a single-cell spike is flagged only through it); dilation by
the 6 neighbours; zero in the two outer layers.
Operator: fourth-order convergence with the detector off (theta = 1), the
difference to the standard operator shrinks at fourth order, and the
induction residual of a radial wind with a radial field (only the dissipation
of the WENOZ face jumps) vanishes faster than h^4.
"""
import numpy as np
import pytest

import heliocubed_amd as hc
import oracle as orc


def test_wenoz_exact_for_quadratics():
    rng = np.random.default_rng(2)
    for _ in range(20):
        a, b, c = rng.normal(size=3)
        f = lambda x: a + b * x + c * x * x
        v = np.array([f(i) for i in range(-3, 3)])       # cells i-3 .. i+2, face at -0.5
        left, right = orc.berta_face(v, "wenoz")
        exact = f(-0.5)
        assert abs(left - exact) <= 1e-14 * max(1.0, abs(exact))
        assert abs(right - exact) <= 1e-14 * max(1.0, abs(exact))


def test_wenoz_ideal_weights_are_the_five_point_interpolant():
    rng = np.random.default_rng(3)
    for _ in range(20):
        a, b, c = rng.normal(size=3)
        # even quartic about the left cell (index -1): the left state is exact
        f = lambda x: a + b * (x + 1) ** 2 + c * (x + 1) ** 4
        v = np.array([f(i) for i in range(-3, 3)])
        left, _ = orc.berta_face(v, "wenoz")
        assert abs(left - f(-0.5)) <= 1e-13 * max(1.0, np.abs(v).max())
        # even quartic about the right cell (index 0): the right state is exact
        g = lambda x: a + b * x ** 2 + c * x ** 4
        v = np.array([g(i) for i in range(-3, 3)])
        _, right = orc.berta_face(v, "wenoz")
        assert abs(right - g(-0.5)) <= 1e-13 * max(1.0, np.abs(v).max())
        # and equals the Lagrange weights on the five upwind points
        w = np.array([3, -20, 90, 60, -5]) / 128.0
        assert abs(right - np.dot(w, v[::-1][:5])) <= 1e-13 * max(1.0, np.abs(v).max())


@pytest.mark.parametrize("method,order", [("wenoz", 4.6), ("weno3", 2.6)])
def test_weno_convergence_on_smooth_data(method, order):
    f = lambda x: np.sin(2.0 * x + 0.3)
    errs = []
    for h in (0.05, 0.025, 0.0125):
        e = 0.0
        for x0 in np.linspace(0.1, 0.6, 6):     # f' != 0 there
            if method == "wenoz":
                v = np.array([f(x0 + o * h) for o in range(-3, 3)])
            else:
                v = np.array([f(x0 + o * h) for o in range(-2, 2)])
            l, r = orc.berta_face(v, method, dx2=h * h)
            exact = f(x0 - 0.5 * h)
            e = max(e, abs(l - exact), abs(r - exact))
        errs.append(e)
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    print(method, "rates", rates)
    assert (rates > order).all(), rates


def test_wenoz_takes_the_smooth_side_at_a_step():
    v = np.array([2.0, 2.0, 2.0, 3.0, 3.0, 3.0])
    left, right = orc.berta_face(v, "wenoz")
    assert abs(left - 2.0) < 1e-30 * 3 + 1e-15 and abs(right - 3.0) < 1e-15
    # a zero stencil gives zero states
    assert orc.berta_face(np.zeros(6), "wenoz") == (0.0, 0.0)


def test_weno3_and_plm_linear_and_extremum():
    v = np.array([1.0, 2.0, 3.0, 4.0])
    for m in ("weno3", "plm"):
        l, r = orc.berta_face(v, m, dx2=1e-4)
        assert abs(l - 2.5) < 1e-14 and abs(r - 2.5) < 1e-14
    # extremum in the left cell: minmod3 gives a flat left state
    l, r = orc.berta_face(np.array([1.0, 3.0, 2.0, 1.0]), "plm")
    assert l == 3.0
    # r: slopes (2-3, 1-2, (1-3)/2) all negative -> 1.5*(-1) vs -1 vs -1.5: -1
    assert r == 2.0 - 0.5 * -1.0


def _uniform_box(n, rho, p=1.0, b1=0.0, gamma=5.0 / 3.0):
    """Native conserved averages [8][n+10]^3 with v = 0; b1 is the native
    radial field component; E = p/(gamma-1) + |A B|^2/(8 pi) at radial B."""
    e = n + 10
    U = np.zeros((8, e, e, e))
    U[0] = rho
    U[4] = p / (gamma - 1.0) + (b1 ** 2) / (8 * np.pi)   # A e_r is a unit vector (column 0 of A)
    U[5] = b1
    return U


def test_detector_radial_density_step():
    n, i0 = 8, 4                       # step at radial cell index lo+i0 (box-local)
    lo, hi = (8, 4, 4), (15, 11, 11)
    rho = np.ones(n + 10)
    rho[5 + i0:] = 2.0                 # box-local index 5+i0 is cell lo+i0
    U = _uniform_box(n, 1.0)
    U[0] = rho[None, None, :]
    for eta_d, flagged in ((0.15, [i0 - 1]), (0.18, [i0 - 1]), (0.25, []), (0.1, [i0 - 1, i0])):
        th = orc.berta_theta(64, 16, 3, lo, hi, U.ravel(), eta_d=eta_d)
        exp = np.zeros_like(th)
        exp[2:-2, 2:-2, 2:-2] = 1.0
        for c in flagged:               # raw flag at box-local radial index 5+c; dilation +-1
            for o in (-1, 0, 1):
                if 2 <= 5 + c + o < n + 8:
                    exp[2:-2, 2:-2, 5 + c + o] = 0.0
        assert np.array_equal(th, exp), (eta_d, np.argwhere(th != exp)[:5])


def test_detector_magnetic_indicator():
    n, i0 = 8, 4
    lo, hi = (8, 4, 4), (15, 11, 11)
    b = np.ones(n + 10)
    b[5 + i0:] = 2.0
    U = _uniform_box(n, 1.0, p=1e3)
    U[5] = b[None, None, :]
    U[4] = 1e3 / (2.0 / 3.0) + (U[5] ** 2) / (8 * np.pi)
    # |B|^2/2: 0.5 | 2 -> indicator 1.5/3.5 = 0.43 at i0-1 and 1.5/6.5 = 0.23 at i0
    th = orc.berta_theta(64, 16, 3, lo, hi, U.ravel(), eta_d=0.3)
    exp = np.zeros_like(th)
    exp[2:-2, 2:-2, 2:-2] = 1.0
    exp[2:-2, 2:-2, 5 + i0 - 2:5 + i0 + 1] = 0.0
    assert np.array_equal(th, exp)


@pytest.mark.parametrize("delta,flag", [(0.02, False), (0.025, True)])
def test_detector_two_norm_over_directions(delta, flag):
    """A density spike of relative size delta has per-direction indicator
    2 delta/(4 + 2 delta): 0.0099 / 0.0123 (each below 0.02), 2-norm over the
    three directions 0.0171 / 0.0214."""
    n = 8
    lo, hi = (8, 4, 4), (15, 11, 11)
    U = _uniform_box(n, 1.0)
    s = (9, 9, 9)
    U[0][s] = 1.0 + delta
    th = orc.berta_theta(64, 16, 3, lo, hi, U.ravel(), eta_d=0.02)
    exp = np.zeros_like(th)
    exp[2:-2, 2:-2, 2:-2] = 1.0
    if flag:
        for ax in range(3):
            for o in (-1, 0, 1):
                q = list(s)
                q[ax] += o
                exp[tuple(q)] = 0.0
    assert np.array_equal(th, exp)


def _berta_rhs(nr, na, boxes, config, quad=4, eta_d=0.02, op=2, r1=215.0, limiter=0):
    shell = hc.Shell(nr, na, 21.5, r1)
    opts = hc.Options(spatial_order=4, limiter=1, riemann_solver=2)
    U = hc.init_native_boxes(shell, opts, boxes, config=config, seed=3, quad=quad)
    return orc.o4_rhs_boxes(nr, na, boxes, U, limiter=limiter, spatial_operator=op, eta_d=eta_d, r1=r1)


def test_berta_rhs_converges_at_fourth_order():
    """Detector off (eta_d above any indicator): WENOZ everywhere. RHS at three
    resolutions of the same physical box, finer ones averaged onto the coarse
    cells; the error ratio approaches 16."""
    fields = []
    for n in (8, 16, 32):
        box = [(3, (n // 2, n // 4, n // 4), (n // 2 + n // 4 - 1, n // 2 - 1, n // 2 - 1))]
        r = _berta_rhs(4 * n, n, box, config=1, quad=5, eta_d=10.0)
        m = n // 4
        fields.append(r.reshape(8, m, m, m))

    def coarsen(a):
        c, m = a.shape[0], a.shape[1] // 2
        return a.reshape(c, m, 2, m, 2, m, 2).mean(axis=(2, 4, 6))
    e0 = np.abs(fields[0] - coarsen(fields[1])).max(axis=(1, 2, 3))
    e1 = np.abs(fields[1] - coarsen(fields[2])).max(axis=(1, 2, 3))
    rate = np.log2(e0[[0, 1, 4]] / e1[[0, 1, 4]])
    print("Berta convergence rates (rho, mom_r, E):", rate)
    assert (rate > 3.3).all(), rate


def test_berta_and_standard_operators_agree_to_fourth_order():
    """Both operators are fourth-order approximations of the same flux
    divergence: their difference falls by ~16 per refinement in every
    component (the steady-wind density RHS is itself truncation error, so the
    difference is measured absolutely)."""
    diffs = []
    for n in (8, 16, 32):
        box = [(1, (n // 2, n // 4, n // 4), (n // 2 + n // 4 - 1, n // 2 - 1, n // 2 - 1))]
        rb = _berta_rhs(4 * n, n, box, config=1, quad=5, eta_d=10.0, op=2).reshape(8, -1)
        rs = _berta_rhs(4 * n, n, box, config=1, quad=5, op=1).reshape(8, -1)
        diffs.append(np.abs(rb - rs).max(axis=1))
    r0, r1 = np.log2(diffs[0] / diffs[1]), np.log2(diffs[1] / diffs[2])
    print("Berta - standard |diff| rates", r0, r1)
    assert (r0 > 3.5).all() and (r1 > 3.5).all(), (r0, r1)


def test_berta_induction_residual_of_a_radial_wind_is_high_order():
    """Radial wind with a radial field (config 2): the point induction fluxes
    v_n B - v B_n cancel exactly, so the B rows of the RHS are the Roe
    dissipation of the WENOZ face jumps alone, which vanish faster than h^4."""
    res = []
    for n in (8, 16, 32):
        box = [(3, (n // 2, n // 4, n // 4), (n // 2 + n // 4 - 1, n // 2 - 1, n // 2 - 1))]
        r = _berta_rhs(4 * n, n, box, config=2, quad=5, eta_d=10.0).reshape(8, -1)
        res.append(np.abs(r[5:8]).max(axis=1))
    r0, r1 = np.log2(res[0] / res[1]), np.log2(res[1] / res[2])
    print("Berta induction residual rates", r0, r1)
    assert (r0 > 4.0).all() and (r1 > 4.0).all(), (r0, r1)
