"""Pins of the oracle's boundary-file path: the frame step of
MHDReader::BoxData_to_BC (src/MHDReader.H:221-331) and the maps of
MHD_BoundaryAverages (src/MHD_BoundaryAverages.H:17-266).

* frame step: the identity at t = 0; a rotation by whole degrees is a roll of
  the longitude columns by that many cells per integer 360 / nphi degrees
  (so nphi = 256 rolls one cell per degree, as the reference computes it);
  between two frames the result is the time blend of the single-frame
  rotations of each (fixed frames: both at t; rotating frames: each from its
  own start); past the last frame the last frame alone;
* fillPoints reproduces a field bilinear in (theta, phi) at every strip cell
  centre whose angles lie inside the node range, with the cell centre's polar
  angle and azimuth computed here from the closed-form panel map, and clamps
  in theta outside it;
* fill of a uniform radial state gives the closed forms (density, radial
  momentum, energy, radial field times the cell averages of (r0/r)^k,
  exact integrals here); a uniform azimuthal velocity or polar field gives the
  native contravariant components of e_phi / e_theta at the cell centre
  (second order) and their Gauss average (fourth order).
"""
import numpy as np
import pytest

import exact_shell as X
import oracle as orc

CARR = 25.38 * 24 * 60 * 60


def _frames(nf, nphi, ntheta, seed=1):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((nf, 8, nphi, ntheta))


def test_frame_identity_and_whole_degree_rolls():
    for nphi in (360, 180, 256):
        F = _frames(1, nphi, 7)
        assert np.array_equal(orc.bc_frame_at(F, 1.0, 0, 0.0), F[0])
        per_cell = 360 // nphi
        for m in (3, 10, 37):
            t = m * per_cell * CARR / 360
            out = orc.bc_frame_at(F, 1.0, 0, t)
            ref = np.roll(F[0], m, axis=1)
            assert np.abs(out - ref).max() < 1e-9, (nphi, m)


def test_frame_time_blend_and_end():
    F = _frames(3, 90, 5, seed=2)
    cad = 2.0
    for t in (0.3 * cad * 3600, 1.7 * cad * 3600):
        low = int(t / (cad * 3600))
        ft = (t - low * cad * 3600) / (cad * 3600)
        out = orc.bc_frame_at(F, cad, 0, t)
        ref = (1 - ft) * orc.bc_frame_at(F[low:low + 1], cad, 0, t) + ft * orc.bc_frame_at(F[low + 1:low + 2], cad, 0, t)
        assert np.abs(out - ref).max() < 1e-13
        out = orc.bc_frame_at(F, cad, 1, t)
        t1, t2 = t + low * cad * 3600, t + (low + 1) * cad * 3600
        ref = (1 - ft) * orc.bc_frame_at(F[low:low + 1], cad, 0, t1) + ft * orc.bc_frame_at(F[low + 1:low + 2], cad, 0, t2)
        assert np.abs(out - ref).max() < 1e-13
    t = 2.4 * cad * 3600
    assert np.array_equal(orc.bc_frame_at(F, cad, 0, t), orc.bc_frame_at(F[2:3], cad, 0, t))


def _centre_angles(n_ang, g, phi0):
    """Polar angle and azimuth of every strip cell centre [6][NJ][NJ]."""
    NJ = n_ang + 2 * g
    th = np.zeros((6, NJ, NJ))
    ph = np.zeros((6, NJ, NJ))
    c = (np.arange(-g, n_ang + g) + 0.5) / n_ang
    for p in range(6):
        M = X.panel_matrix(p)
        xi1, xi2 = np.meshgrid(c, c, indexing="xy")       # [k][j]
        n = X.unit_dir(M, xi1, xi2)
        th[p] = np.arccos(np.clip(n[2], -1, 1))
        ph[p] = phi0 + np.mod(np.arctan2(n[1], n[0]) - phi0 + 4 * np.pi, 2 * np.pi)
    return th, ph


def test_points_reproduce_bilinear_fields():
    nt, npf = 37, 72
    theta = np.linspace(0.1, np.pi - 0.1, nt)
    phi = np.linspace(0.2, 0.2 + 2 * np.pi, npf, endpoint=False)
    coef = np.array([[1.0, 0.3, -0.2, 0.05], [2.0, -0.1, 0.4, 0.02]])
    data0 = np.zeros((8, npf, nt))
    T, P = np.meshgrid(theta, phi, indexing="xy")            # [phi][theta]
    for c in range(8):
        a, b, cc, d = coef[c % 2] * (1 + 0.1 * c)
        data0[c] = a + b * T + cc * P + d * T * P
    n_ang, g = 16, 5
    bc = orc.bc_points(24, n_ang, g, data0, theta, phi)
    th, ph = _centre_angles(n_ang, g, phi[0])
    inside = (th >= theta[0]) & (th <= theta[-1]) & (ph <= phi[-1])
    assert inside.sum() > 0.8 * inside.size
    for c in range(8):
        a, b, cc, d = coef[c % 2] * (1 + 0.1 * c)
        exact = a + b * th + cc * ph + d * th * ph
        assert np.abs(bc[:, c][inside] - exact[inside]).max() < 1e-12 * np.abs(exact).max()
    # beyond the last latitude node: the value at the clamped node
    th_clamped = np.clip(th, theta[0], theta[-1])
    out = (th < theta[0]) | (th > theta[-1])
    a, b, cc, d = coef[0]
    near = out & (ph <= phi[-1])
    assert near.sum() > 0
    assert np.abs(bc[:, 0][near] - (a + b * th_clamped + cc * ph + d * th_clamped * ph)[near]).max() < 1e-12


def _s_average(ir, nr, k, r0, r1):
    """Exact cell average in xi of (r0/r)^k over radial ghost cell ir (linear r)."""
    x, w = np.polynomial.legendre.leggauss(64)
    xi = (ir + 0.5 * (x + 1)) / nr
    r = r0 + (r1 - r0) * xi
    return float(0.5 * (w * (r0 / r) ** k).sum())


@pytest.mark.parametrize("order", [2, 4])
def test_averages_of_a_uniform_radial_state(order):
    nt, npf = 19, 36
    theta = np.linspace(0, np.pi, nt)
    phi = np.linspace(0, 2 * np.pi, npf, endpoint=False)
    n0, vr, p0, br = 300.0, 4e7, 10.0, 150.0
    data0 = np.zeros((8, npf, nt))
    data0[0], data0[1], data0[4], data0[5] = n0, vr, p0, br
    nr, na, g, r0, r1, gamma = 24, 16, 5 if order == 4 else 2, 21.5, 60.0, 5.0 / 3.0
    inner = orc.bc_averages(nr, na, g, order, data0, theta, phi, r0=r0, r1=r1, gamma=gamma)
    rho = n0 * 1.67262192e-24
    for ig in range(g):
        ir = ig - g
        if order == 4:
            s1, s2, s4, sg = (_s_average(ir, nr, k, r0, r1) for k in (1, 2, 4, 2 * gamma))
        else:
            xi = (ir + 0.5) / nr
            ratio = r0 / (r0 + (r1 - r0) * xi)
            s1, s2, s4, sg = ratio, ratio ** 2, ratio ** 4, ratio ** (2 * gamma)
        expect = [rho * s2, rho * vr * s2, 0, 0,
                  p0 * 1e-12 / (gamma - 1) * sg + 0.5 * rho * vr * vr * s2 + 1e-12 * br * br / (8 * np.pi) * s4,
                  1e-6 * br * s2, 0, 0]
        for c in range(8):
            got = inner[:, c, :, :, ig]
            scale = max(abs(expect[c]), abs(expect[0]) if c in (2, 3) else 1e-6 * br * s2 if c in (6, 7) else 0)
            assert np.abs(got - expect[c]).max() <= 1e-11 * scale, (order, ig, c)


def _native_tangent(p, j, k, n_ang, e, nq):
    """Native (xi1, xi2) components of a unit tangential field e in {'phi', 'theta'}
    averaged over angular cell (j, k) by an nq x nq Gauss rule (nq = 1: centre)."""
    M = X.panel_matrix(p)
    if nq == 1:
        x, w = np.array([0.5]), np.array([1.0])
    else:
        x, w = np.polynomial.legendre.leggauss(nq)
        x, w = 0.5 * (x + 1), 0.5 * w
    tot = np.zeros(2)
    for a, wa in zip(x, w):
        for b, wb in zip(x, w):
            n, d1, d2 = X.basis(M, np.array((j + a) / n_ang), np.array((k + b) / n_ang))
            t, ph = np.arccos(np.clip(n[2], -1, 1)), np.arctan2(n[1], n[0])
            vec = np.array([-np.sin(ph), np.cos(ph), 0.0]) if e == "phi" else \
                np.array([np.cos(t) * np.cos(ph), np.cos(t) * np.sin(ph), -np.sin(t)])
            A = np.column_stack([n, d1, d2])
            tot += wa * wb * np.linalg.solve(A, vec)[1:]
    return tot


@pytest.mark.parametrize("order", [2, 4])
def test_averages_of_tangential_fields(order):
    nt, npf = 19, 36
    theta = np.linspace(0, np.pi, nt)
    phi = np.linspace(0, 2 * np.pi, npf, endpoint=False)
    n0 = 100.0
    data0 = np.zeros((8, npf, nt))
    data0[0], data0[2], data0[7] = n0, 1.0, 1.0          # v_phi = 1, B_theta = 1
    nr, na, r0, r1 = 24, 16, 21.5, 60.0
    g = 5 if order == 4 else 2
    inner = orc.bc_averages(nr, na, g, order, data0, theta, phi, r0=r0, r1=r1)
    rho = n0 * 1.67262192e-24
    ratio_mid = r0 / (r0 + (r1 - r0) * (0.5 - g) / nr) if order == 2 else None
    for (p, j, k) in [(0, 3, 4), (1, 8, 8), (3, 0, 15), (5, 12, 2)]:
        nq = 1 if order == 2 else 8
        vp = _native_tangent(p, j, k, na, "phi", nq)
        bt = _native_tangent(p, j, k, na, "theta", nq)
        s2 = ratio_mid ** 2 if order == 2 else _s_average(-g, nr, 2, r0, r1)
        s1 = ratio_mid if order == 2 else _s_average(-g, nr, 1, r0, r1)
        got = inner[p, :, k + g, j + g, 0]
        assert np.allclose(got[2:4], rho * vp * s2, rtol=1e-9, atol=1e-12 * rho * s2)
        assert np.allclose(got[6:8], 1e-6 * bt * s1, rtol=1e-9, atol=1e-15)
