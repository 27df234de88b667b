"""Known answers for the point pieces of the fourth-order oracle.

* Limiters_4O limiter (src/BoxOp_Limiters.H:28-148): the limited face values
  of one cell from its 7-cell line, against values worked out by hand from
  the published rules (McCorquodale & Colella 2011, Colella & Sekora 2008):
  constant and linear data unchanged, a smooth monotone profile unchanged, a
  steep monotone profile clipped to twice the opposite deviation, a step
  clipped on its steep side, a smooth extremum bounded by 1.25 times the
  smallest neighbouring second difference, and a cubic extremum (constant
  third differences) left alone.
* Flattening eta_tilde (src/BoxOp_Limiters.H:151-188): 1 without compression
  or with a weak pressure jump, 1 - (z - 0.75)/0.10 in the ramp, 0 above
  z = 0.85, 1 below 0.75 (Colella & Woodward 1984 with z0 = 0.75, z1 = 0.85,
  delta = 0.33).
* Point flux (energies + f_primflux, src/BoxOp_EulerCubedSphere.H:195-225,
  src/BoxOp_Utilities.H:508-597): with the pressure term through the cofactor
  row it is the textbook ideal-MHD flux through that row, in Gaussian units.
* factoredRoe8WaveDissipation (src/MHD_Roe8Wave.H:634-740): zero for equal
  states; for a small jump with continuous normal field, |J_n| (U_R - U_L)
  of the textbook flux Jacobian along the face normal (eigen-decomposed by
  numpy), times the face area, returned in the native factored form (scalars
  / det A, vectors A^-1 (.) / det A).
"""
import numpy as np
import pytest

import oracle as orc

GAMMA = 5.0 / 3.0
PI = np.pi


# ---------------------------------------------------------------- limiter
def test_limiter_constant_and_linear_data_unchanged():
    La, Rr = orc.limit_line([2.5] * 7)
    assert La == 2.5 and Rr == 2.5
    # w_i = a + b i: the fourth-order face values a + b/2 and a - b/2 are exact
    # and the profile is monotone with equal deviations: nothing is clipped
    a, b = 1.25, -0.375
    w = [a + b * i for i in range(-3, 4)]
    La, Rr = orc.limit_line(w)
    assert La == pytest.approx(a + b / 2, abs=1e-15) and Rr == pytest.approx(a - b / 2, abs=1e-15)


def test_limiter_smooth_monotone_profile_unchanged():
    # w_i = (i + 10)^2: face values u^2 + u + 1/6 (u = 10) and (u-1)^2 + (u-1) + 1/6;
    # deviations 10 + 1/6 and 10 - 1/6 differ by far less than a factor 2
    w = [(i + 10.0) ** 2 for i in range(-3, 4)]
    La, Rr = orc.limit_line(w)
    assert La == pytest.approx(100 + 10 + 1 / 6, rel=1e-15)
    assert Rr == pytest.approx(81 + 9 + 1 / 6, rel=1e-15)


def test_limiter_steep_monotone_side_is_clipped_to_twice_the_other():
    # w = (0, .1, .2, .3, 1.5, 2.7, 3.9), cell value .3: unlimited face values
    # 7/12 (.3 + 1.5) - 1/12 (.2 + 2.7) = .808333.. and 7/12 (.2 + .3) - 1/12 (.1 + 1.5) = .158333..;
    # high deviation .508333 >= 2 x low deviation .141667: the high face value
    # becomes .3 + 2 x .141667 = .583333..; the low face value is kept
    La, Rr = orc.limit_line([0.0, 0.1, 0.2, 0.3, 1.5, 2.7, 3.9])
    assert La == pytest.approx(0.3 + 2 * (0.3 - 0.95 / 6), rel=1e-14)
    assert Rr == pytest.approx(0.95 / 6, rel=1e-14)
    assert abs(La - 0.3) <= 2 * abs(0.3 - Rr) * (1 + 1e-14)


def test_limiter_step_is_clipped_on_its_steep_side():
    # cell value 0 with the step to 1 above it: unlimited 1/2 (high) and -1/12
    # (low); the second differences change sign, so the curvature ratio is 0
    # and the high face is clipped to 0 + 2 x (1/12) = 1/6
    La, Rr = orc.limit_line([0, 0, 0, 0, 1, 1, 1])
    assert La == pytest.approx(1 / 6, rel=1e-15)
    assert Rr == pytest.approx(-1 / 12, rel=1e-15)


def test_limiter_smooth_extremum_is_bounded_by_the_neighbouring_curvature():
    # w = (0, 3, 5, 6, 5, 3, 0): both face values 69/12 = 5.75; second
    # differences -1, -1, -2, -1, -1 and the face curvature 6 (5.75 - 12 + 5.75) = -3
    # all negative: the limited curvature is -min(3, 1.25 x 1, 1.25 x 2, 1.25 x 1) = -1.25,
    # ratio 1.25/3; third differences vary (0, -1, 1, 0), so the faces are
    # pulled to 6 - (1.25/3) x 0.25
    La, Rr = orc.limit_line([0, 3, 5, 6, 5, 3, 0])
    want = 6 - (1.25 / 3) * 0.25
    assert La == pytest.approx(want, rel=1e-15) and Rr == pytest.approx(want, rel=1e-15)


def test_limiter_leaves_an_extremum_with_constant_third_differences():
    # w_i = -i^3 + 3 i^2: a local extremum of a cubic, third differences all
    # -6 (their spread 0 < 0.1 x 6): no limiting whatever the curvature ratio
    w = np.array([-(i ** 3) + 3.0 * i * i + 1.0 for i in range(-3, 4)])
    # unlimited face values of the middle cell
    La0 = 7 / 12 * (w[3] + w[4]) - 1 / 12 * (w[2] + w[5])
    Rr0 = 7 / 12 * (w[2] + w[3]) - 1 / 12 * (w[1] + w[4])
    La, Rr = orc.limit_line(w)
    assert La == La0 and Rr == Rr0


# ------------------------------------------------------------- flattening
@pytest.mark.parametrize("p,v,eta", [
    # compression, |dp|/min p = 0.8 > 0.33, z = 0.8 / 1.0 in the ramp: 1 - 0.05/0.10
    ([1.0, 1.0, 1.4, 1.8, 2.0], [0, 1, 0, -1, 0], 0.5),
    # expansion: no flattening
    ([1.0, 1.0, 1.4, 1.8, 2.0], [0, -1, 0, 1, 0], 1.0),
    # weak jump |dp|/min p = 0.2 <= 0.33
    ([1.0, 1.0, 1.1, 1.2, 2.0], [0, 1, 0, -1, 0], 1.0),
    # z = 0.9 / 1.0 above 0.85: fully flattened
    ([1.0, 1.0, 1.5, 1.9, 2.0], [0, 1, 0, -1, 0], 0.0),
    # z = 0.4 / 1.0 below 0.75 (jump 0.4/1.2 > 0.33): not flattened
    ([1.0, 1.2, 1.4, 1.6, 2.0], [0, 1, 0, -1, 0], 1.0),
    # z = 0.78: 1 - 0.03/0.10
    ([1.0, 1.0, 1.39, 1.78, 2.0], [0, 2, 0, 1, 0], 0.7),
])
def test_flattening_known_values(p, v, eta):
    assert orc.eta_tilde(p, v) == pytest.approx(eta, abs=1e-13)


# ------------------------------------------------------------- point flux
def _cofactor_row(A, d):
    return np.linalg.det(A) * np.linalg.inv(A)[d]


def _textbook_flux_through(n, rho, v, p, b):
    """N . F of the ideal-MHD flux tensor (Gaussian units) for a vector N."""
    vn, bn = v @ n, b @ n
    pt = p + b @ b / (8 * PI)
    E = p / (GAMMA - 1) + 0.5 * rho * v @ v + b @ b / (8 * PI)
    f = np.zeros(8)
    f[0] = rho * vn
    f[1:4] = rho * vn * v - bn * b / (4 * PI) + pt * n
    f[4] = (E + pt) * vn - bn * (v @ b) / (4 * PI)
    f[5:8] = vn * b - bn * v
    return f


@pytest.mark.parametrize("d", [0, 1, 2])
def test_point_flux_is_the_textbook_flux_through_the_cofactor_row(d):
    rng = np.random.default_rng(20 + d)
    for trial in range(6):
        p_ = int(rng.integers(6))
        A = orc.basis(p_, rng.uniform(-0.7, 0.7), rng.uniform(-0.7, 0.7)) if trial % 2 == 0 \
            else np.eye(3) + 0.3 * rng.normal(size=(3, 3))
        w = np.array([rng.uniform(0.5, 2), *rng.normal(size=3), rng.uniform(0.2, 2), *rng.normal(size=3)])
        f, (ke, me, vb) = orc.prim_flux(A, w, d, GAMMA)
        vc, bc = A @ w[1:4], A @ w[5:8]
        assert ke == pytest.approx(0.5 * w[0] * vc @ vc, rel=1e-13)
        assert me == pytest.approx(bc @ bc / (8 * PI), rel=1e-13)
        assert vb == pytest.approx(vc @ bc, rel=1e-13)
        N = _cofactor_row(A, d)
        det = np.linalg.det(A)
        got = np.zeros(8)
        got[0], got[4] = det * f[0], det * f[4]
        got[1:4] = det * (A @ f[1:4]) + (w[4] + me) * N
        got[5:8] = det * (A @ f[5:8])
        ref = _textbook_flux_through(N, w[0], vc, w[4], bc)
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max()), (got, ref)


# ----------------------------------------------------------- factored Roe
def _cons(rho, v, p, b):
    return np.array([rho, *(rho * v), p / (GAMMA - 1) + 0.5 * rho * v @ v + b @ b / (8 * PI), *b])


def _prim(u):
    rho, v, b = u[0], u[1:4] / u[0], u[5:8]
    return rho, v, (GAMMA - 1) * (u[4] - 0.5 * rho * v @ v - b @ b / (8 * PI)), b


def test_factored_roe_zero_for_equal_states():
    A = orc.basis(2, 0.3, -0.2)
    w = np.array([1.3, 0.4, -0.2, 0.1, 0.9, 0.7, -0.3, 0.5])
    assert np.abs(orc.factored_roe(w, w, A, 2.7, 1, GAMMA, 0.3)).max() == 0.0


@pytest.mark.parametrize("d", [0, 1, 2])
def test_factored_roe_small_jump_is_abs_jacobian_times_area(d):
    rng = np.random.default_rng(5 + d)
    A = orc.basis(int(rng.integers(6)), rng.uniform(-0.6, 0.6), rng.uniform(-0.6, 0.6))
    area = 3.1
    N = _cofactor_row(A, d)
    n = N / np.linalg.norm(N)
    det = np.linalg.det(A)
    rho, vc, p, bc = 1.2, np.array([0.5, -0.3, 0.2]), 0.9, np.array([1.1, 1.6, -0.8])
    u0 = _cons(rho, vc, p, bc)
    # Jacobian of N.F / |N| along n, with the normal field held fixed (7 waves)
    Q = np.zeros((8, 8))
    Q[0, 0] = Q[4, 4] = 1.0
    tang = np.linalg.svd(np.outer(n, n) - np.eye(3))[0][:, :2]
    frame = np.vstack([n, tang.T])
    Q[1:4, 1:4] = frame
    Q[5:8, 5:8] = frame

    def flux_frame(uf):
        u = Q.T @ uf
        r_, v_, p_, b_ = _prim(u)
        return Q @ _textbook_flux_through(n, r_, v_, p_, b_)

    uf0 = Q @ u0
    idx = [0, 1, 2, 3, 4, 6, 7]
    J = np.zeros((7, 7))
    for c in range(7):
        e = np.zeros(8)
        e[idx[c]] = 1e-6 * max(1.0, abs(uf0[idx[c]]))
        J[:, c] = ((flux_frame(uf0 + e) - flux_frame(uf0 - e)) / (2 * e[idx[c]]))[idx]
    lam, R = np.linalg.eig(J)
    absJ = (R @ np.diag(np.abs(lam)) @ np.linalg.inv(R)).real
    Ainv = np.linalg.inv(A)
    wl = np.array([rho, *(Ainv @ vc), p, *(Ainv @ bc)])
    for _ in range(3):
        duf = np.zeros(8)
        duf[idx] = rng.normal(size=7) * 1e-7
        u1 = Q.T @ (uf0 + duf)
        r1, v1, p1, b1 = _prim(u1)
        wr = np.array([r1, *(Ainv @ v1), p1, *(Ainv @ b1)])
        out = orc.factored_roe(wl, wr, A, area, d, GAMMA, 1e-12)
        dcart = np.zeros(8)
        dcart[idx] = absJ @ duf[idx]
        dcart = Q.T @ dcart * area * np.linalg.norm(N)
        want = np.zeros(8)
        want[0], want[4] = dcart[0] / det, dcart[4] / det
        want[1:4] = Ainv @ dcart[1:4] / det
        want[5:8] = Ainv @ dcart[5:8] / det
        assert np.allclose(out, want, rtol=0, atol=2e-4 * np.abs(want).max()), (out, want)
