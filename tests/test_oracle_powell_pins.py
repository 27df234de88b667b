"""Pins of the Powell source (divergence_cleaning 1) of the second-order oracle.

The source is S = -J div(B) (0, B/4pi, v.B/4pi, v) with -J div(B) the flux
divergence of the face magnetic fluxes 1/2 (B_l + B_r).n A
(src/BoxOp_SecondOrder.H:317-369). The pins do not restate the oracle's code:
* a uniform field has zero divergence, and the discrete -J div(B) of a cell is
  a signed sum of B.n A over its faces, which vanishes by the closure of the
  integrated face normals (to round-off, seams included);
* B = c x e_x has div(B) = c, so -J div(B) / J -> -c at second order, with J
  the cell volume per unit mapped volume from the closed-form radial volume
  and the cell's solid angle;
* given -J div(B), the momentum, energy and induction sources take the
  textbook Powell form in Gaussian units.
"""
import numpy as np
import pytest

import oracle as orc

R0, R1 = 21.5, 215.0


def _stage(nr, na, field, v=(1.0, 0.5, -0.25)):
    """Ghosted Cartesian primitives (g = 2) with rho = p = 1, uniform v and
    B = field(X) at the cell centres (ghost cells included)."""
    g = 2
    W = np.zeros((6, 8, na + 2 * g, na + 2 * g, nr + 2 * g))
    r = R0 + (np.arange(-g, nr + g) + 0.5) * (R1 - R0) / nr
    for p in range(6):
        for kk in range(na + 2 * g):
            for jj in range(na + 2 * g):
                a = -np.pi / 4 + (np.pi / 2) * (jj - g + 0.5) / na
                b = -np.pi / 4 + (np.pi / 2) * (kk - g + 0.5) / na
                rhat = orc.basis(p, a, b)[:, 0]
                X = r[:, None] * rhat[None, :]
                B = field(X)
                W[p, 0, kk, jj] = 1.0
                W[p, 4, kk, jj] = 1.0
                for d in range(3):
                    W[p, 1 + d, kk, jj] = v[d]
                    W[p, 5 + d, kk, jj] = B[:, d]
    return W


def _source(nr, na, W, limiter=0):
    k1 = orc.Oracle(nr, na, limiter=limiter, riemann=1, divergence_cleaning=1).rhs(W)
    k0 = orc.Oracle(nr, na, limiter=limiter, riemann=1, divergence_cleaning=0).rhs(W)
    return k1 - k0


def _jacobian(nr, na):
    """J = dvol(i) detA(j,k): volume per unit mapped volume, [na][na][nr]."""
    dxi = 1.0 / nr
    e = np.arange(nr + 1) * dxi
    rr = R0 + (R1 - R0) * e
    dvol = (rr[1:] ** 3 - rr[:-1] ** 3) / 3.0 / dxi
    O = orc.Oracle(nr, na)
    detA = np.array([[O.cell_detA(j, k) for j in range(na)] for k in range(na)])
    return detA[:, :, None] * dvol[None, None, :]


def test_uniform_field_has_no_powell_source():
    nr, na = 8, 8
    B0 = np.array([0.3, -0.2, 0.5])
    W = _stage(nr, na, lambda X: np.broadcast_to(B0, X.shape))
    S = _source(nr, na, W, limiter=1)
    # -J div B is read off the x induction source (v_x = 1); its scale is the
    # magnitude of one radial face's term |B| r1^2 (pi/2)^2 nr
    scale = np.linalg.norm(B0) * R1 ** 2 * (np.pi / 2) ** 2 * nr
    assert np.abs(S[:, 0]).max() == 0.0
    assert np.abs(S[:, 5]).max() <= 1e-13 * scale


def test_linear_field_divergence_converges_at_second_order():
    c = 1.0 / R1
    errs = []
    for n in (8, 16):
        W = _stage(n, n, lambda X: np.stack([c * X[:, 0], 0 * X[:, 0], 0 * X[:, 0]], axis=1))
        S = _source(n, n, W)
        md = S[:, 5]                          # v_x = 1
        J = _jacobian(n, n)[None]
        errs.append(np.abs(md / (-c * J) - 1.0).mean())
    rate = np.log2(errs[0] / errs[1])
    assert errs[1] < 5e-3 and rate > 1.75, (errs, rate)


def test_source_components_take_the_powell_form():
    nr, na = 8, 8
    v = np.array([1.0, 0.5, -0.25])
    c = 1.0 / R1
    W = _stage(nr, na, lambda X: np.stack([c * X[:, 0], 0.2 + 0 * X[:, 0], -0.1 + c * X[:, 2]], axis=1), v=tuple(v))
    S = _source(nr, na, W)
    g = 2
    Wc = W[:, :, g:-g, g:-g, g:-g]
    md = S[:, 5] / v[0]
    assert np.abs(md).max() > 0
    scale = np.abs(md).max() * np.abs(Wc[:, 5:8]).max()
    for d in range(3):
        assert np.allclose(S[:, 5 + d], md * v[d], rtol=1e-13, atol=1e-13 * np.abs(md).max())
        assert np.allclose(S[:, 1 + d], md * Wc[:, 5 + d] / (4 * np.pi), rtol=1e-13, atol=1e-14 * scale)
    vb = sum(v[d] * Wc[:, 5 + d] for d in range(3))
    assert np.allclose(S[:, 4], md * vb / (4 * np.pi), rtol=1e-12, atol=1e-14 * scale)
    assert np.abs(S[:, 0]).max() == 0.0


# ----------------------------------------------------- fourth-order operator
# The fourth-order source (src/BoxOp_DivBcleaning.H:31-168) on one box whose
# ghosted input holds exact cell averages (4-point Gauss-Legendre per
# direction) of the native conserved state of rho = p = 1, a uniform velocity
# and B = B0 + c x e_x (div B = c). With J the volume per unit mapped volume:
#   induction: <-J div B v> = -c <J> v,  momentum: -c <J B>/(4 pi),
#   energy: -c <J v.B>/(4 pi),
# each converging at fourth order.
GL_X, GL_W = np.polynomial.legendre.leggauss(4)
GL_X, GL_W = 0.5 * (GL_X + 1.0), 0.5 * GL_W
V0 = np.array([1.0, 0.5, -0.25])
B0 = np.array([0.05, -0.02, 0.03])


def _field(X, c, b0=B0):
    B = np.broadcast_to(b0, X.shape).copy()
    B[..., 0] += c * X[..., 0]
    return B


def _box_averages(nr, na, p, lo, hi, c, g=5, b0=B0):
    """Ghosted native conserved averages [8][n2+2g][n1+2g][n0+2g] and, on the
    valid box, the cell averages of J, J B and J v.B."""
    gamma = 5.0 / 3.0
    n = [hi[d] - lo[d] + 1 + 2 * g for d in range(3)]
    L = [lo[d] - g for d in range(3)]
    U = np.zeros((8, n[2], n[1], n[0]))
    JB = np.zeros((3, n[2], n[1], n[0]))
    J = np.zeros((n[2], n[1], n[0]))
    JvB = np.zeros((n[2], n[1], n[0]))
    xr = (L[0] + np.arange(n[0])[:, None] + GL_X[None, :]) / nr          # [n0][4]
    r = R0 + (R1 - R0) * xr
    for kk in range(n[2]):
        for jj in range(n[1]):
            acc = np.zeros((8, n[0]))
            accJ = np.zeros(n[0]); accJB = np.zeros((3, n[0])); accJvB = np.zeros(n[0])
            for qa, wa in zip(GL_X, GL_W):
                for qb, wb in zip(GL_X, GL_W):
                    al = -np.pi / 4 + (np.pi / 2) * (L[1] + jj + qa) / na
                    be = -np.pi / 4 + (np.pi / 2) * (L[2] + kk + qb) / na
                    A = orc.basis(p, al, be)
                    Ai = np.linalg.inv(A)
                    X = r[..., None] * A[:, 0]                              # [n0][4][3]
                    B = _field(X, c, b0)
                    vn = Ai @ V0
                    Bn = np.einsum("ij,abj->abi", Ai, B)
                    E = 1.0 / (gamma - 1) + 0.5 * (V0 @ V0) + (B * B).sum(-1) / (8 * np.pi)
                    jac = r ** 2 * (R1 - R0) * np.linalg.det(A)              # [n0][4]
                    w = wa * wb * GL_W[None, :]
                    acc[0] += w.sum(1)
                    for d in range(3):
                        acc[1 + d] += (w * vn[d]).sum(1)
                        acc[5 + d] += (w * Bn[..., d]).sum(1)
                        accJB[d] += (w * jac * B[..., d]).sum(1)
                    acc[4] += (w * E).sum(1)
                    accJ += (w * jac).sum(1)
                    accJvB += (w * jac * (B @ V0)).sum(1)
            U[:, kk, jj] = acc
            J[kk, jj] = accJ; JB[:, kk, jj] = accJB; JvB[kk, jj] = accJvB
    v = (slice(g, -g),) * 3
    return U, J[v], JB[(slice(None),) + v], JvB[v]


def _o4_source(nr, na, p, lo, hi, U, op=1):
    box = [(p, tuple(lo), tuple(hi))]
    k1 = orc.o4_rhs_boxes(nr, na, box, U.ravel(), riemann=1, divergence_cleaning=1, spatial_operator=op)
    k0 = orc.o4_rhs_boxes(nr, na, box, U.ravel(), riemann=1, divergence_cleaning=0, spatial_operator=op)
    n = [hi[d] - lo[d] + 1 for d in range(3)]
    return (k1 - k0).reshape(8, n[2], n[1], n[0])


@pytest.mark.parametrize("op", [1, 2])
def test_fourth_order_powell_source_converges_at_fourth_order(op):
    # op 1: standard operator (face value from the cell-average primitives,
    # deconvolved coefficient); op 2: Berta (face value from <U>, point
    # primitives). For Berta the uniform part of the field is large, so the
    # magnetic pressure stays smooth and every Jameson flag is 1: the
    # product-rule branch is the one converging.
    c = 1.0 / R1
    b0 = B0 if op == 1 else np.array([3.0, -2.0, 2.5])
    errs = []
    for n, lo, hi in ((16, (4, 4, 4), (11, 11, 11)), (32, (8, 8, 8), (23, 23, 23))):
        U, J, JB, JvB = _box_averages(n, n, 2, lo, hi, c, b0=b0)
        if op == 2:
            assert orc.berta_theta(n, n, 2, lo, hi, U.ravel())[5:-5, 5:-5, 5:-5].min() == 1.0
        S = _o4_source(n, n, 2, lo, hi, U, op)
        assert np.abs(S[0]).max() == 0.0
        e = []
        for d in range(3):
            e.append(np.abs(S[5 + d] - (-c * J * V0[d])).max() / np.abs(c * J).max())
            e.append(np.abs(S[1 + d] - (-c * JB[d] / (4 * np.pi))).max() / np.abs(c * JB).max() * 4 * np.pi)
        e.append(np.abs(S[4] - (-c * JvB / (4 * np.pi))).max() / np.abs(c * JvB).max() * 4 * np.pi)
        errs.append(np.array(e))
    rate = np.log2(errs[0] / errs[1])
    assert (errs[1] < 1e-4).all() and (rate > 3.5).all(), (errs, rate)
