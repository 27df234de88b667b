"""CPU pins of the oracle's CFL time step (dtCFLSecondOrderHost,
src/BoxOp_SecondOrderCFL.H:99-210; dtCFL/dtInv/f_speedmax,
src/BoxOp_EulerCubedSphereImplem.H:878-954,1013-1131, src/BoxOp_EulerCubedSphere.H:127-194).

The expected values are built from the physics, not from the oracle's code:
* States whose native components are recovered exactly (JU = dVolr <det(A) A> w
  for a native vector w) give closed-form wave speeds: a radial wind with a
  radial field has |A_d v| = |v_r| and |A_d b| = |B_r| at any point.
* The metric widths converge at second order to widths from spherical
  trigonometry (radial spacing; radius x solid angle of the strip centred on
  the face / arc length of the face), which the test evaluates exactly.
* A tangential native velocity pins where each direction evaluates its basis
  (cell centre for d = 0, centre of the low face for d = 1, 2).
* Floor and cone: rho = RHO00 for an empty cell; the cone density bounds
  only the Alfven term, at second order only.
The geometry in the expectations comes from the 8-point Gauss-Legendre face
averages (orc_face_average), independent of the closed forms the oracle's CFL
uses.
"""
import ctypes

import numpy as np
import pytest

import oracle as orc

GAMMA = 5.0 / 3.0
R0, R1 = 21.5, 215.0
RHO00 = 1.67262192e-26


def _geometry(nr, na):
    cfg = orc.OrcCfg(nr, na, 5, R0, R1, 0.0, GAMMA, 1, 2, 0.3)
    det0 = np.zeros((6, na, na))
    dA0 = np.zeros((6, na, na, 3, 3))
    detd = np.zeros((3, 6, na, na))
    nrm = np.zeros((3, 6, na, na))
    x1, x9, x3 = np.zeros(1), np.zeros(9), np.zeros(3)
    L = orc.lib()
    for p in range(6):
        for k in range(na):
            for j in range(na):
                L.orc_face_average(ctypes.byref(cfg), p, 0, j, k, 0, x1.ctypes.data)
                det0[p, k, j] = x1[0]
                L.orc_face_average(ctypes.byref(cfg), p, 0, j, k, 1, x9.ctypes.data)
                dA0[p, k, j] = x9.reshape(3, 3)
                for d in (1, 2):
                    L.orc_face_average(ctypes.byref(cfg), p, d, j, k, 0, x1.ctypes.data)
                    detd[d, p, k, j] = x1[0]
                    L.orc_face_average(ctypes.byref(cfg), p, d, j, k, 2, x3.ctypes.data)
                    nrm[d, p, k, j] = np.linalg.norm(x3)
    dxi = 1.0 / nr
    r = R0 + (R1 - R0) * np.arange(nr + 1) / nr
    rlo, rhi = r[:-1], r[1:]
    dvol = (rhi ** 3 - rlo ** 3) / 3.0 / dxi
    areaT = (rhi ** 2 - rlo ** 2) / dxi / 2.0
    h = 1.0 / na
    # widths [d][p][k][j][i]
    w = np.zeros((3, 6, na, na, nr))
    w[0] = (rhi - rlo)[None, None, None, :]
    for d in (1, 2):
        w[d] = dvol[None, None, None, :] * detd[d][..., None] * h / areaT[None, None, None, :] / nrm[d][..., None]
    return dict(det0=det0, dA0=dA0, dvol=dvol, w=w, rlo=rlo, rhi=rhi)


def _state(G, rho, vnat, bnat, E):
    """JU of native primitives: rho, E scalars and native vectors [..,3] per cell [p][k][j][i]."""
    det0, dA0, dvol = G["det0"], G["dA0"], G["dvol"]
    J = det0[..., None] * dvol[None, None, None, :]
    ju = np.zeros((6, 8) + rho.shape[1:])
    ju[:, 0] = rho * J
    ju[:, 4] = E * J
    mom = np.einsum("pkjab,pkjib->pkjia", dA0, rho[..., None] * vnat) * dvol[None, None, None, :, None]
    bb = np.einsum("pkjab,pkjib->pkjia", dA0, bnat) * dvol[None, None, None, :, None]
    for a in range(3):
        ju[:, 1 + a] = mom[..., a]
        ju[:, 5 + a] = bb[..., a]
    return ju


def _radial_case(nr, na, seed=0):
    G = _geometry(nr, na)
    rng = np.random.default_rng(seed)
    shape = (6, na, na, nr)
    rmid = 0.5 * (G["rlo"] + G["rhi"])
    base = (R0 / rmid) ** 2
    rho = base * (1 + 0.3 * rng.random(shape))
    vr = 1 + rng.standard_normal(shape) * 0.5
    br = (0.7 + 0.3 * rng.random(shape)) * base * np.sign(rng.standard_normal(shape))
    p = 0.1 * rho * (1 + rng.random(shape))
    E = p / (GAMMA - 1) + 0.5 * rho * vr ** 2 + br ** 2 / (8 * np.pi)
    vnat = np.zeros(shape + (3,)); vnat[..., 0] = vr
    bnat = np.zeros(shape + (3,)); bnat[..., 0] = br
    ju = _state(G, rho, vnat, bnat, E)
    return G, ju, rho, vr, br, p


@pytest.mark.parametrize("order", [2, 4])
def test_radial_wind_closed_form(order):
    nr, na = 10, 8
    G, ju, rho, vr, br, p = _radial_case(nr, na)
    cf = np.sqrt(GAMMA * p / rho + br ** 2 / (4 * np.pi * rho))
    S = (cf + np.abs(vr)) * (1 / G["w"][0] + 1 / G["w"][1] + 1 / G["w"][2])
    expect = 1.0 / S.max()
    got = orc.dt_cfl(nr, na, ju, order=order)
    assert abs(got - expect) <= 1e-12 * expect, (got, expect)


def test_cone_density_bounds_alfven_term_at_second_order_only():
    nr, na = 10, 8
    G, ju, rho, vr, br, p = _radial_case(nr, na, seed=3)
    cone = 0.5 * np.median(rho)
    rm = np.minimum(rho, cone)
    cf = np.sqrt(GAMMA * p / rho + br ** 2 / (4 * np.pi * rm))
    S = (cf + np.abs(vr)) * (1 / G["w"][0] + 1 / G["w"][1] + 1 / G["w"][2])
    got = orc.dt_cfl(nr, na, ju, order=2, cone_density=cone)
    assert abs(got - 1.0 / S.max()) <= 1e-12 / S.max()
    assert orc.dt_cfl(nr, na, ju, order=4, cone_density=cone) == orc.dt_cfl(nr, na, ju, order=4)


def test_tangential_velocity_uses_low_face_basis():
    nr, na = 6, 8
    G = _geometry(nr, na)
    shape = (6, na, na, nr)
    rng = np.random.default_rng(5)
    rho = np.ones(shape)
    vnat = np.zeros(shape + (3,))
    vnat[..., 1] = 1 + rng.random(shape)
    vnat[..., 2] = rng.standard_normal(shape)
    bnat = np.zeros(shape + (3,))
    h = 1.0 / na
    ang = lambda x: -np.pi / 4 + np.pi / 2 * x
    p0 = 1e-3
    S = np.zeros(shape)
    E = np.zeros(shape)
    for pp in range(6):
        for k in range(na):
            for j in range(na):
                Ac = orc.basis(pp, ang((j + .5) * h), ang((k + .5) * h))
                A1 = orc.basis(pp, ang(j * h), ang((k + .5) * h))
                A2 = orc.basis(pp, ang((j + .5) * h), ang(k * h))
                v = vnat[pp, k, j]                       # [nr, 3]
                E[pp, k, j] = p0 / (GAMMA - 1) + 0.5 * np.sum((v @ Ac.T) ** 2, axis=1)
                c = np.sqrt(GAMMA * p0)
                S[pp, k, j] = sum((c + np.linalg.norm(v @ A.T, axis=1)) / G["w"][d][pp, k, j]
                                  for d, A in enumerate((Ac, A1, A2)))
    ju = _state(G, rho, vnat, bnat, E)
    expect = 1.0 / S.max()
    for order in (2, 4):
        got = orc.dt_cfl(nr, na, ju, order=order)
        assert abs(got - expect) <= 1e-11 * expect, (order, got, expect)
    # the cell-centred basis in every direction would differ at O(h)
    S_cell = np.zeros(shape)
    for pp in range(6):
        for k in range(na):
            for j in range(na):
                Ac = orc.basis(pp, ang((j + .5) * h), ang((k + .5) * h))
                v = vnat[pp, k, j]
                S_cell[pp, k, j] = sum((np.sqrt(GAMMA * p0) + np.linalg.norm(v @ Ac.T, axis=1)) / G["w"][d][pp, k, j]
                                       for d in range(3))
    assert abs(1.0 / S_cell.max() - expect) > 1e-4 * expect


def test_empty_cell_takes_density_floor():
    nr, na = 6, 8
    G, ju, rho, vr, br, p = _radial_case(nr, na, seed=7)
    pp, k, j, i = 2, 3, 5, 4
    E = 1e-20
    J = G["det0"][pp, k, j] * G["dvol"][i]
    ju[pp, :, k, j, i] = 0.0
    ju[pp, 4, k, j, i] = E * J
    c = np.sqrt(GAMMA * (GAMMA - 1) * E / RHO00)
    expect = 1.0 / (c * sum(1.0 / G["w"][d][pp, k, j, i] for d in range(3)))
    for order in (2, 4):
        got = orc.dt_cfl(nr, na, ju, order=order)
        assert abs(got - expect) <= 1e-12 * expect, (order, got, expect)


def test_negative_pressure_is_clipped():
    nr, na = 6, 8
    G, ju, rho, vr, br, p = _radial_case(nr, na, seed=11)
    # remove the thermal energy and more: p < 0 -> 0 in every cell
    ju[:, 4] -= 2 * (p / (GAMMA - 1)) * G["det0"][..., None] * G["dvol"][None, None, None, :]
    S = (np.abs(br) / np.sqrt(4 * np.pi * rho) + np.abs(vr)) * (1 / G["w"][0] + 1 / G["w"][1] + 1 / G["w"][2])
    got = orc.dt_cfl(nr, na, ju, order=4)
    assert abs(got - 1.0 / S.max()) <= 1e-12 / S.max()


def _spherical_dt(nr, na, r1):
    """sqrt(gamma) / max sum_d 1/width_d from spherical trigonometry: width_0 =
    dr, width_d = r_mid x (solid angle of the cell-sized strip centred on the
    low face of d) / (arc length of that face). Coordinate lines of the
    equiangular panel are great circles, so both are exact; they match the
    metric widths to second order in h and dr/r."""
    h = 1.0 / na

    def unit(a, b):
        x, y = np.tan(-np.pi / 4 + np.pi / 2 * a), np.tan(-np.pi / 4 + np.pi / 2 * b)
        v = np.stack([np.ones_like(x), x, y], -1)
        return v / np.linalg.norm(v, axis=-1, keepdims=True)

    def arc(u, v):
        return np.arctan2(np.linalg.norm(np.cross(u, v), axis=-1), np.sum(u * v, axis=-1))

    def tri(a, b, c):
        num = np.abs(np.sum(a * np.cross(b, c), axis=-1))
        return 2 * np.arctan2(num, 1 + np.sum(a * b, -1) + np.sum(b * c, -1) + np.sum(c * a, -1))

    def quad(a0, a1, b0, b1):
        c00, c10, c11, c01 = unit(a0, b0), unit(a1, b0), unit(a1, b1), unit(a0, b1)
        return tri(c00, c10, c11) + tri(c00, c11, c01)

    K, J = np.meshgrid(np.arange(na) * h, np.arange(na) * h, indexing="ij")   # low corner (xi_a, xi_b) = (J, K)
    w1 = quad(J - h / 2, J + h / 2, K, K + h) / arc(unit(J, K), unit(J, K + h))
    w2 = quad(J, J + h, K - h / 2, K + h / 2) / arc(unit(J, K), unit(J + h, K))
    r = R0 + (r1 - R0) * np.arange(nr + 1) / nr
    rmid, dr = 0.5 * (r[:-1] + r[1:]), r[1:] - r[:-1]
    S = 1 / dr[None, None, :] + 1 / (rmid[None, None, :] * w1[..., None]) + 1 / (rmid[None, None, :] * w2[..., None])
    return 1.0 / (np.sqrt(GAMMA) * S.max())


def test_widths_converge_to_physical_widths():
    r1 = 60.0
    errs = []
    for n in (8, 16, 32):
        # uniform gas at rest: rho = 1, p = 1 -> c = sqrt(gamma)
        ju = np.zeros((6, 8, n, n, n))
        cfg = orc.OrcCfg(n, n, 5, R0, r1, 0.0, GAMMA, 1, 2, 0.3)
        r = R0 + (r1 - R0) * np.arange(n + 1) / n
        dvol = (r[1:] ** 3 - r[:-1] ** 3) / 3.0 * n
        det0 = np.array([[orc.lib().orc_cell_detA(ctypes.byref(cfg), j, k) for j in range(n)] for k in range(n)])
        J = det0[None, :, :, None] * dvol[None, None, None, :]
        ju[:, 0] = J
        ju[:, 4] = J / (GAMMA - 1)
        got = orc.dt_cfl(n, n, ju, order=4, r1=r1)
        ref = _spherical_dt(n, n, r1)
        errs.append(abs(got - ref) / ref)
    rates = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert errs[0] < 0.02, errs
    assert min(rates) > 1.7, (errs, rates)
