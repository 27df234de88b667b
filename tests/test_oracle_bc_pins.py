"""CPU pins of the oracle's inner radial ghosts from boundary-file data
(PreStagePatch file branch, src/BoxOp_PreStage.H:161-317; second-order ghost
conversion, src/BoxOp_SecondOrderPreStage.H:194-214).

The expectations are the physics of the boundary data, built without the
oracle's conversion code:
* second order: the ghost cell's Cartesian primitives are the file's
  spherical vectors rotated to Cartesian at the cell centre
  (v = u_r r + u_theta theta + u_phi phi), density in g/cm^3 (x m_p) scaled by
  (r_BC/r)^2, pressure in dyne/cm^2 scaled by (r_BC/r)^(2 gamma), and a
  field in Gauss with the radial part scaled by (r_BC/r)^2 and the
  tangential parts by r_BC/r (Parker-spiral scaling);
* fourth order: native components through the inverse of the cell-averaged
  <det(A) A>/<det A>, with the averages taken here by 8-point Gauss-Legendre
  (independent of the oracle's closed forms), so the point basis would fail.
"""
import ctypes

import numpy as np
import pytest

import oracle as orc

GAMMA = 5.0 / 3.0
R0, R1 = 21.5, 60.0
MP = 1.67262192e-24


def _strip(na, g, seed):
    rng = np.random.default_rng(seed)
    n = na + 2 * g
    bc = np.zeros((6, 8, n, n))
    bc[:, 0] = 300 + 400 * rng.random((6, n, n))               # amu / cm^3
    bc[:, 1] = 3e7 + 4e7 * rng.random((6, n, n))               # v_r cm/s
    bc[:, 2] = 2e6 * rng.standard_normal((6, n, n))            # v_phi
    bc[:, 3] = 2e6 * rng.standard_normal((6, n, n))            # v_theta
    bc[:, 4] = 3e3 + 4e3 * rng.random((6, n, n))               # pdyne / cm^2
    bc[:, 5] = 1e3 * rng.standard_normal((6, n, n))            # B_r uG
    bc[:, 6] = 5e2 * rng.standard_normal((6, n, n))            # B_phi
    bc[:, 7] = 5e2 * rng.standard_normal((6, n, n))            # B_theta
    return bc


def _frame(p, j, k, na):
    h = 1.0 / na
    ang = lambda x: -np.pi / 4 + np.pi / 2 * x
    A = orc.basis(p, ang((j + .5) * h), ang((k + .5) * h))
    rhat = A[:, 0]
    th, ph = np.arccos(rhat[2]), np.arctan2(rhat[1], rhat[0])
    that = np.array([np.cos(th) * np.cos(ph), np.cos(th) * np.sin(ph), -np.sin(th)])
    phat = np.array([-np.sin(ph), np.cos(ph), 0.0])
    return A, rhat, that, phat


def _radii(nr, g):
    dr = (R1 - R0) / nr
    return np.array([R0 + (ig - g + 0.5) * dr for ig in range(g)])


def test_second_order_ghosts_are_the_rotated_scaled_file_state():
    nr, na, g = 12, 8, 2
    bc = _strip(na, g, 1)
    got = orc.inner_ghosts_from_bc(nr, na, bc, order=2, r1=R1)
    rad = _radii(nr, g)
    rng = np.random.default_rng(9)
    for _ in range(60):
        p, k, j = int(rng.integers(6)), int(rng.integers(-g, na + g)), int(rng.integers(-g, na + g))
        A, rhat, that, phat = _frame(p, j, k, na)
        w = bc[p, :, k + g, j + g]
        for ig in range(g):
            s = R0 / rad[ig]
            v = w[1] * rhat + w[3] * that + w[2] * phat
            B = 1e-6 * (s ** 2 * w[5] * rhat + s * (w[6] * phat + w[7] * that))
            exp = np.concatenate([[w[0] * MP * s ** 2], v, [w[4] * 1e-12 * s ** (2 * GAMMA)], B])
            o = got[p, :, k + g, j + g, ig]
            assert abs(o[0] - exp[0]) <= 1e-13 * exp[0]
            assert np.abs(o[1:4] - v).max() <= 1e-12 * np.abs(v).max(), (p, k, j, ig, o, exp)
            assert np.abs(o[5:8] - B).max() <= 1e-12 * np.abs(B).max(), (p, k, j, ig, o, exp)
            assert abs(o[4] - exp[4]) <= 1e-10 * exp[4], (o[4], exp[4])


def test_fourth_order_ghosts_use_the_averaged_matrix():
    nr, na, g = 12, 16, 5
    bc = _strip(na, g, 2)
    got = orc.inner_ghosts_from_bc(nr, na, bc, order=4, r1=R1)
    rad = _radii(nr, g)
    cfg = orc.OrcCfg(nr, na, g, R0, R1, 0.0, GAMMA, 1, 2, 0.3)
    rng = np.random.default_rng(4)
    worst_point = 0.0
    for _ in range(40):
        p, k, j = int(rng.integers(6)), int(rng.integers(-g, na + g)), int(rng.integers(-g, na + g))
        det, dA = np.zeros(1), np.zeros(9)
        orc.lib().orc_face_average(ctypes.byref(cfg), p, 0, j, k, 0, det.ctypes.data)
        orc.lib().orc_face_average(ctypes.byref(cfg), p, 0, j, k, 1, dA.ctypes.data)
        M = dA.reshape(3, 3) / det[0]
        A, rhat, that, phat = _frame(p, j, k, na)
        w = bc[p, :, k + g, j + g]
        vc = w[1] * rhat + w[3] * that + w[2] * phat
        bcart = w[5] * rhat + w[7] * that + w[6] * phat
        for ig in range(g):
            s = R0 / rad[ig]
            rho = w[0] * MP * s ** 2
            mom = rho * np.linalg.solve(M, vc)
            bn = np.linalg.solve(M, bcart) * 1e-6 * np.array([s ** 2, s, s])
            o = got[p, :, k + g, j + g, ig]
            assert abs(o[0] - rho) <= 1e-13 * rho
            assert np.abs(o[1:4] - mom).max() <= 1e-11 * np.abs(mom).max(), (o[1:4], mom)
            assert np.abs(o[5:8] - bn).max() <= 1e-11 * np.abs(bn).max(), (o[5:8], bn)
            # the point basis differs at second order in h
            mom_point = rho * np.linalg.solve(A, vc)
            worst_point = max(worst_point, np.abs(mom_point - mom).max() / np.abs(mom).max())
    assert worst_point > 1e-6


def test_density_only_strip_gives_scaled_proton_density():
    nr, na, g = 12, 8, 2
    n = na + 2 * g
    bc = np.zeros((6, 8, n, n))
    bc[:, 0] = 1.0
    got = orc.inner_ghosts_from_bc(nr, na, bc, order=2, r1=R1)
    rad = _radii(nr, g)
    for ig in range(g):
        assert np.allclose(got[:, 0, :, :, ig], MP * (R0 / rad[ig]) ** 2, rtol=1e-14)
    assert np.abs(got[:, 1:]).max() == 0.0


def test_outflow_boundary_closed_form_and_rotation():
    """Radial-outflow test boundary (f_radialoutflowIBCs, src/BoxOp_PreStage.H:5-48):
    at strength 0 the ghosts are 700 m_p (r_BC/r)^2, 5e7 cm/s radially and
    1e-7 (r_BC/r)^(2 gamma); the perturbation co-rotates, so advancing time by
    a full Carrington period reproduces the state and a seventh of it does not."""
    nr, na, g = 12, 8, 2
    rad = _radii(nr, g)
    base = orc.inner_ghosts_outflow(nr, na, 0.0, 0.0, order=2, r1=R1)
    for p in range(6):
        for k in range(-g, na + g):
            for j in range(-g, na + g):
                A, rhat, that, phat = _frame(p, j, k, na)
                for ig in range(g):
                    s = R0 / rad[ig]
                    o = base[p, :, k + g, j + g, ig]
                    assert abs(o[0] - 700 * MP * s ** 2) <= 1e-13 * o[0]
                    assert np.abs(o[1:4] - 5e7 * rhat).max() <= 1e-12 * 5e7
                    assert abs(o[4] - 1e-7 * s ** (2 * GAMMA)) <= 1e-11 * o[4]
                    assert np.abs(o[5:8]).max() == 0.0
    period = 25.38 * 24 * 60 * 60
    a = orc.inner_ghosts_outflow(nr, na, 20.0, 1000.0, order=4, r1=R1)
    b = orc.inner_ghosts_outflow(nr, na, 20.0, 1000.0 + period, order=4, r1=R1)
    c = orc.inner_ghosts_outflow(nr, na, 20.0, 1000.0 + period / 7, order=4, r1=R1)
    assert np.abs(a - b).max() <= 1e-9 * np.abs(a).max()
    assert np.abs(a[:, 1] - c[:, 1]).max() > 1e-3 * np.abs(a[:, 1]).max()
    # the speed perturbation is bounded by strength/300 * 2 of u0
    speed = a[:, 1] / a[:, 0]
    assert np.abs(speed / 5e7 - 1).max() <= 20.0 / 300 * 2
