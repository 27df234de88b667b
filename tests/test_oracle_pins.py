"""Pins of the CPU oracle (no GPU): closed forms, invariants and known answers.

Each test checks an oracle function against something the mathematics fixes,
chosen so that a plausible mistake (dropped term, wrong sign, transposed
index) fails:
* metric averages: 8x8 Gauss-Legendre quadrature (oracle) against the exact
  solid angle / great-circle closed forms (product host code), the total
  panel solid angle 4 pi / 6, and the discrete metric identity;
* the pointwise basis of src/MHD_BoundaryAverages.H:62-70 against finite
  differences of the panel map;
* the seam interpolation rule: exact for cubic polynomials in the donor
  panel's angles, donor lists identical to the product's tables;
* PLM (src/BoxOp_SecondOrder.H:7-37): constant and linear data, extrema;
* Rusanov (src/MHD_Roe8Wave.H:604-629): closed form;
* Roe 8-wave (src/MHD_Roe8Wave.H:16-600): D(W,W)=0 and, for small jumps,
  D = |A| dU with A the numerically differentiated conservative MHD flux
  Jacobian eigen-decomposed by numpy;
* the rotated face flux (src/MHD_SecondOrderFlux.H:34-74): equal states give
  the textbook MHD flux through the face;
* whole operator: conservation to round-off, six-panel symmetry of a
  spherically symmetric state, fourth-order temporal convergence of RK4.
"""
import ctypes

import numpy as np
import pytest

import heliocubed_amd as hc
import oracle as orc
from helpers import cfl_dt

GAMMA = 5.0 / 3.0
PI = np.pi


# ---------------------------------------------------------------- geometry
@pytest.mark.parametrize("na", [8, 32, 64])
def test_cell_detA_quadrature_matches_exact_solid_angle(na):
    O = orc.Oracle(4, na)
    h = 1.0 / na
    total = 0.0
    for j in range(na):
        for k in range(0, na, max(1, na // 8)):
            q = O.cell_detA(j, k)
            e = hc.lib().hc_geom_cell_detA(na, j, k)
            assert abs(q - e) <= 1e-13 * e
    for j in range(na):
        for k in range(na):
            total += hc.lib().hc_geom_cell_detA(na, j, k) * h * h
    assert abs(total - 4 * PI / 6) <= 1e-13


@pytest.mark.parametrize("na", [8, 32])
def test_face_normals_quadrature_matches_closed_form_and_gcl(na):
    O = orc.Oracle(4, na)
    rng = np.random.default_rng(3)
    for _ in range(60):
        p, d = rng.integers(6), rng.integers(3)
        j = rng.integers(na + (d == 1))
        k = rng.integers(na + (d == 2))
        q = O.face_adjA(p, d, j, k)
        e = np.zeros(3)
        hc.lib().hc_geom_face_normal(na, int(p), int(d), int(j), int(k), e.ctypes.data)
        assert np.abs(q - e).max() <= 1e-13 * np.abs(e).max()
    # discrete metric identity with the oracle's own normals
    h = 1.0 / na
    for p in range(6):
        for (j, k) in [(0, 0), (na // 2, 3), (na - 1, na - 1)]:
            s = O.face_adjA(p, 0, j, k) + (O.face_adjA(p, 1, j + 1, k) - O.face_adjA(p, 1, j, k)
                                            + O.face_adjA(p, 2, j, k + 1) - O.face_adjA(p, 2, j, k)) / (2 * h)
            assert np.abs(s).max() <= 1e-12


def _panel_point(p, a, b):
    perm = [[1, 2, 0], [1, 2, 0], [1, 0, 2], [0, 1, 2], [1, 0, 2], [0, 1, 2]]
    sign = [[1, -1, -1], [1, 1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1], [-1, -1, 1]]
    loc = np.array([1.0, np.tan(a), np.tan(b)])
    loc /= np.linalg.norm(loc)
    return np.array([sign[p][c] * loc[perm[p][c]] for c in range(3)])


def test_basis_is_derivative_of_panel_map():
    eps = 1e-6
    for p in range(6):
        for (a, b) in [(0.0, 0.0), (0.3, -0.5), (-0.7, 0.7)]:
            A = orc.basis(p, a, b)
            x = _panel_point(p, a, b)
            assert np.allclose(A[:, 0], x, atol=1e-14)
            da = (_panel_point(p, a + eps, b) - _panel_point(p, a - eps, b)) / (2 * eps) * (PI / 2)
            db = (_panel_point(p, a, b + eps) - _panel_point(p, a, b - eps)) / (2 * eps) * (PI / 2)
            assert np.allclose(A[:, 1], da, atol=1e-8)
            assert np.allclose(A[:, 2], db, atol=1e-8)
            assert np.linalg.det(A) > 0


def test_panels_tile_the_sphere():
    """Six panel centres are the six axis directions, each exactly once."""
    centres = np.array([_panel_point(p, 0.0, 0.0) for p in range(6)])
    assert np.allclose(np.abs(centres).sum(axis=1), 1.0)
    assert np.allclose(sorted(map(tuple, np.round(centres))),
                       sorted([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]))


# ------------------------------------------------------------ seam stencils
def _local_angles(q, x):
    perm = [[1, 2, 0], [1, 2, 0], [1, 0, 2], [0, 1, 2], [1, 0, 2], [0, 1, 2]]
    sign = [[1, -1, -1], [1, 1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1], [-1, -1, 1]]
    L = np.zeros(3)
    for a in range(3):
        L[perm[q][a]] = sign[q][a] * x[a]
    return np.arctan2(L[1], L[0]), np.arctan2(L[2], L[0])


@pytest.mark.parametrize("na,g", [(8, 2), (16, 5)])
def test_seam_donors_exact_for_cubics_and_match_product(na, g):
    O = orc.Oracle(4, na, ghost=g)
    h = 1.0 / na
    coef = np.random.default_rng(5).normal(size=(4, 4))

    def f(a, b):  # cubic polynomial in the donor panel's angles
        return sum(coef[m, n] * a ** m * b ** n for m in range(4) for n in range(4))

    qa = np.zeros(16, np.int32); ja = np.zeros(16, np.int32); ka = np.zeros(16, np.int32); wa = np.zeros(16)
    checked = 0
    for p in range(6):
        for k in range(-g, na + g):
            for j in range(-g, na + g):
                if 0 <= j < na and 0 <= k < na:
                    continue
                donors = O.ghost_donors(p, j, k)
                n = hc.lib().hc_geom_ghost_donors(na, p, j, k, qa.ctypes.data, ja.ctypes.data,
                                                  ka.ctypes.data, wa.ctypes.data)
                assert n == len(donors)
                for x, (q, dj, dk, w) in enumerate(donors):
                    assert (q, dj, dk) == (qa[x], ja[x], ka[x])
                    assert abs(w - wa[x]) <= 1e-15 * max(1.0, abs(w))
                    assert 0 <= dj < na and 0 <= dk < na and q != p
                assert abs(sum(w for *_, w in donors) - 1.0) <= 1e-13
                x = _panel_point(p, -PI / 4 + (j + 0.5) * h * PI / 2, -PI / 4 + (k + 0.5) * h * PI / 2)
                q = donors[0][0]
                a, b = _local_angles(q, x)
                interp = sum(w * f(-PI / 4 + (dj + 0.5) * h * PI / 2, -PI / 4 + (dk + 0.5) * h * PI / 2)
                             for (_, dj, dk, w) in donors)
                assert abs(interp - f(a, b)) <= 1e-11, (p, j, k)
                checked += 1
    assert checked == 6 * ((na + 2 * g) ** 2 - na * na)


@pytest.mark.parametrize("na", [8, 32])
def test_seam_edges_are_symmetric_and_geometric(na):
    O = orc.Oracle(4, na)
    h = 1.0 / na
    for p in range(6):
        for d in (1, 2):
            for s in (0, 1):
                q, dq, sq, rev = O.seam_edge(p, d, s)
                assert O.seam_edge(q, dq, sq) == (p, d, s, rev)
                for t in (0, na // 2, na - 1):
                    tq = na - 1 - t if rev else t
                    fa = (na if s else 0) * h if d == 1 else (t + 0.5) * h
                    fb = (t + 0.5) * h if d == 1 else (na if s else 0) * h
                    ga = (na if sq else 0) * h if dq == 1 else (tq + 0.5) * h
                    gb = (tq + 0.5) * h if dq == 1 else (na if sq else 0) * h
                    xp = _panel_point(p, -PI / 4 + fa * PI / 2, -PI / 4 + fb * PI / 2)
                    xq = _panel_point(q, -PI / 4 + ga * PI / 2, -PI / 4 + gb * PI / 2)
                    assert np.allclose(xp, xq, atol=1e-13)


# -------------------------------------------------------------- pointwise
def _state(rho, v, p, b):
    return np.array([rho, *v, p, *b], dtype=np.float64)


def test_plm_known_answers():
    c = _state(1.0, (0.1, 0.2, 0.3), 0.5, (1.0, 2.0, 3.0))
    for lim in (0, 1, 3):
        lo, hi = orc.plm(c, c, c, c, lim)
        assert np.array_equal(lo, c) and np.array_equal(hi, c)
    slope = _state(0.1, (0.01, -0.02, 0.03), 0.05, (0.1, 0.0, -0.1))
    cells = [c + m * slope for m in range(-2, 2)]
    for lim in (0, 1, 3):
        lo, hi = orc.plm(*cells, lim)
        face = c - 0.5 * slope
        assert np.allclose(lo, face, rtol=0, atol=1e-15) and np.allclose(hi, face, rtol=0, atol=1e-15)
    # extremum at m: minmod/MC slopes vanish, the unlimited slope does not
    cells = [c, c + slope, c - 0.5 * slope, c - slope]
    lo, hi = orc.plm(*cells, 1)
    assert np.array_equal(lo, cells[1])
    lo0, _ = orc.plm(*cells, 0)
    assert not np.array_equal(lo0, cells[1])
    # negative extrapolated density falls back to the complete cell state
    a = _state(1.0, (0, 0, 0), 1.0, (0, 0, 0)); b = _state(0.05, (1, 0, 0), 1.0, (0, 0, 0))
    lo, hi = orc.plm(b, b, b, a, 0)   # unlimited hi density 0.05 - 0.2375 < 0
    assert np.array_equal(hi, b)
    lo, hi = orc.plm(b, b, b, a, 1)   # minmod keeps the slope zero: no fallback needed
    assert np.array_equal(hi, b)


def test_rusanov_closed_form():
    l = _state(1.0, (0.0, 0.0, 0.0), 1.0, (0.0, 0.0, 0.0))
    r = _state(2.0, (0.0, 0.0, 0.0), 1.0, (0.0, 0.0, 0.0))
    d = orc.rusanov(l, r, GAMMA)
    speed = np.sqrt(GAMMA * 1.0 / 1.0)
    assert np.allclose(d, speed * np.array([1.0, 0, 0, 0, 0, 0, 0, 0]), rtol=1e-15)
    w = _state(1.3, (0.4, -0.2, 0.1), 0.7, (0.5, -1.0, 2.0))
    assert np.array_equal(orc.rusanov(w, w, GAMMA), np.zeros(8))


def _conserved(w):
    rho, v, p, b = w[0], w[1:4], w[4], w[5:8]
    return np.array([rho, *(rho * v), p / (GAMMA - 1) + 0.5 * rho * v @ v + b @ b / (8 * PI), *b])


def _primitive(u):
    rho = u[0]; v = u[1:4] / rho; b = u[5:8]
    p = (GAMMA - 1) * (u[4] - 0.5 * rho * v @ v - b @ b / (8 * PI))
    return np.array([rho, *v, p, *b])


def _flux_x(w):
    """Textbook ideal-MHD flux in x, Gaussian units (magnetic pressure B^2/8pi)."""
    rho, v, p, b = w[0], w[1:4], w[4], w[5:8]
    pt = p + b @ b / (8 * PI)
    E = p / (GAMMA - 1) + 0.5 * rho * v @ v + b @ b / (8 * PI)
    f = np.zeros(8)
    f[0] = rho * v[0]
    f[1:4] = rho * v[0] * v - b[0] * b / (4 * PI)
    f[1] += pt
    f[4] = (E + pt) * v[0] - b[0] * (v @ b) / (4 * PI)
    f[5:8] = v[0] * b - v * b[0]
    return f


def test_roe_zero_jump_and_small_jump_is_abs_jacobian():
    w = _state(1.2, (0.7, -0.3, 0.2), 0.9, (1.5, 2.0, -1.1))
    ok, d = orc.roe8(w, w, GAMMA, 0.3)
    assert ok and np.abs(d).max() == 0.0
    # |A| dU for a small jump with B_x fixed (7x7 conservative system)
    idx = [0, 1, 2, 3, 4, 6, 7]
    u0 = _conserved(w)

    def flux_of(u7):
        u = u0.copy(); u[idx] = u7
        return _flux_x(_primitive(u))[idx]

    J = np.zeros((7, 7))
    for c in range(7):
        e = np.zeros(7); e[c] = 1e-6 * max(1.0, abs(u0[idx][c]))
        J[:, c] = (flux_of(u0[idx] + e) - flux_of(u0[idx] - e)) / (2 * e[c])
    lam, R = np.linalg.eig(J)
    absA = (R @ np.diag(np.abs(lam)) @ np.linalg.inv(R)).real
    rng = np.random.default_rng(11)
    for _ in range(5):
        du = rng.normal(size=7) * 1e-7
        u1 = u0.copy(); u1[idx] += du
        wr = _primitive(u1)
        ok, d = orc.roe8(w, wr, GAMMA, 1e-12)
        assert ok
        ref = absA @ du
        assert abs(d[5]) <= 1e-12 * np.abs(ref).max()
        assert np.allclose(d[idx], ref, rtol=0, atol=1e-4 * np.abs(ref).max())


def test_roe_rejects_inadmissible_states():
    good = _state(1.0, (0.0, 0.0, 0.0), 1.0, (0.5, 0.5, 0.0))
    bad = _state(1.0, (0.0, 0.0, 0.0), -1.0, (0.5, 0.5, 0.0))
    assert not orc.roe8(good, bad, GAMMA, 0.3)[0]
    # strong two-rarefaction pair: the linearized fan has a negative intermediate state
    l = _state(1.0, (-20.0, 0.0, 0.0), 0.01, (0.0, 0.0, 0.0))
    r = _state(1.0, (20.0, 0.0, 0.0), 0.01, (0.0, 0.0, 0.0))
    assert not orc.roe8(l, r, GAMMA, 0.3)[0]
    assert orc.roe8(l, r, GAMMA, 0.3, check=False)[0]


@pytest.mark.parametrize("roe", [False, True])
def test_face_flux_of_equal_states_is_textbook_flux(roe):
    w = _state(1.1, (0.6, -0.4, 0.25), 0.8, (1.2, -0.7, 0.9))
    rng = np.random.default_rng(2)
    for _ in range(5):
        n = rng.normal(size=3)
        area = 3.7
        out = orc.face_flux(w, w, n, area, GAMMA, 0.3, roe)
        nh = n / np.linalg.norm(n)
        rho, v, p, b = w[0], w[1:4], w[4], w[5:8]
        vn, bn = v @ nh, b @ nh
        pt = p + b @ b / (8 * PI)
        E = p / (GAMMA - 1) + 0.5 * rho * v @ v + b @ b / (8 * PI)
        ref = np.zeros(8)
        ref[0] = rho * vn
        ref[1:4] = rho * vn * v + pt * nh - bn * b / (4 * PI)
        ref[4] = (E + pt) * vn - bn * (v @ b) / (4 * PI)
        ref[5:8] = vn * b - bn * v
        ref *= area * np.linalg.norm(n)
        assert np.allclose(out, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())


# ----------------------------------------------------------- whole operator
@pytest.mark.parametrize("limiter,riemann", [(1, 1), (3, 2)])
def test_oracle_rhs_conserves(limiter, riemann):
    nr, na = 12, 8
    shell = hc.Shell(nr, na)
    ju, inner = hc.init_state(shell, hc.Options(limiter=limiter, riemann_solver=riemann), 0, 9)
    O = orc.Oracle(nr, na, limiter=limiter, riemann=riemann)
    k, F = O.rhs(O.prepare(ju, inner), with_fluxes=True)
    total = k.sum(axis=(0, 2, 3, 4)) / nr / na ** 2
    boundary = (F[0][..., nr].sum(axis=(0, 2, 3)) - F[0][..., 0].sum(axis=(0, 2, 3))) / na ** 2
    scale = np.abs(F[0]).sum(axis=(0, 2, 3, 4)) / na ** 2
    assert np.all(np.abs(total + boundary) <= 1e-13 * scale)


def test_oracle_spherically_symmetric_state_gives_identical_panels():
    nr, na = 8, 8
    shell = hc.Shell(nr, na)
    ju, inner = hc.init_state(shell, hc.Options(), 2, 1)
    O = orc.Oracle(nr, na, limiter=1, riemann=2)
    k = O.rhs(O.prepare(ju, inner))
    for c in (0, 4):  # scalars: density and energy
        ref = np.sort(k[0, c].ravel())
        for p in range(1, 6):
            assert np.allclose(np.sort(k[p, c].ravel()), ref, rtol=0, atol=1e-11 * np.abs(ref).max())
    # radial momentum magnitude is also panel-independent
    mag = np.sqrt((k[:, 1:4] ** 2).sum(axis=1))
    for p in range(1, 6):
        assert np.allclose(np.sort(mag[p].ravel()), np.sort(mag[0].ravel()), rtol=0,
                           atol=1e-11 * mag.max())


def test_oracle_rk4_driver_follows_reference_tableau():
    """One RK4 step equals the stage composition with the classical tableau
    a = {0; 1/2; 0,1/2; 0,0,1}, b = {1/6,1/3,1/3,1/6} held in
    src/MHD_AMRRungeKutta.H:8-12 (stage sums and update order of :23-39).
    The composite operator is only piecewise smooth (|v_n| and max() in the
    Riemann speeds), so a global temporal-order test is not well posed."""
    nr, na = 8, 8
    shell = hc.Shell(nr, na)
    ju, inner = hc.init_state(shell, hc.Options(), 0, 4)
    O = orc.Oracle(nr, na, limiter=1, riemann=2)
    dt = cfl_dt(na)
    a = [[0, 0, 0], [0.5, 0, 0], [0, 0.5, 0], [0, 0, 1]]
    b = [1 / 6, 1 / 3, 1 / 3, 1 / 6]
    ks = []
    for s in range(4):
        stage = ju.copy()
        for prev in range(s):
            if a[s][prev]:
                stage = stage + (dt * a[s][prev]) * ks[prev]
        ks.append(O.rhs(O.prepare(stage, inner)))
    u = ju.copy()
    for s in range(4):
        u = u + (dt * b[s]) * ks[s]
    got = O.rk4(ju, inner, dt, 1)
    assert np.array_equal(got, u)
