"""Exact (quadrature) data of a smooth analytic MHD state on the cubed-sphere
shell, written with numpy only, for the convergence pins of the oracle.

The panel map is the closed form of the equiangular cubed sphere: panel p
maps (xi1, xi2) in [0, 1]^2 to the unit vector M_p (1, tan a, tan b) / s,
a = -pi/4 + (pi/2) xi1, b = -pi/4 + (pi/2) xi2, s = |(1, tan a, tan b)|, with
M_p the signed permutation of the panel; the radius is linear in xi0. The
angular derivatives are taken by complex steps (exact to round-off).
Definitions follow the reference: native components are the coefficients
of a Cartesian vector in the basis A = (x_hat, d x_hat/d xi1, d x_hat/d xi2)
(src/MHD_BoundaryAverages.H:62-70); the level state is <J U> with
J = r^2 dr/dxi0 det A (src/BoxOp_Utilities.H:162-218); the RHS is
d<JU>/dt = -sum_d (<N_d . F>(hi) - <N_d . F>(lo)) / dxi_d with N_d the
cofactor rows of d x / d xi and F the ideal-MHD flux tensor in Gaussian units.
"""
import numpy as np

import oracle as orc

GAMMA = 5.0 / 3.0
PI = np.pi


def panel_matrix(p):
    """Signed permutation M_p of panel p, read from the basis at the panel
    centre (column 0 = centre direction, columns 1, 2 = the tangents over pi/2)."""
    A = orc.basis(p, 0.0, 0.0)
    M = np.column_stack([A[:, 0], A[:, 1] / (PI / 2), A[:, 2] / (PI / 2)])
    assert np.allclose(np.abs(M).sum(axis=0), 1.0, atol=1e-14) and np.allclose(M @ M.T, np.eye(3), atol=1e-14)
    return np.round(M)


def unit_dir(M, xi1, xi2):
    a = -PI / 4 + (PI / 2) * xi1
    b = -PI / 4 + (PI / 2) * xi2
    x, y = np.tan(a), np.tan(b)
    s = np.sqrt(1 + x * x + y * y)
    loc = np.stack([1 / s, x / s, y / s])
    return np.tensordot(M, loc, axes=(1, 0))


def basis(M, xi1, xi2):
    """(x_hat, d x_hat/d xi1, d x_hat/d xi2) as arrays [3, ...] each."""
    h = 1e-30
    n = unit_dir(M, xi1, xi2)
    d1 = unit_dir(M, xi1 + 1j * h, xi2).imag / h
    d2 = unit_dir(M, xi1, xi2 + 1j * h).imag / h
    return n.real, d1, d2


def state(X, r0=21.5):
    """Smooth primitive state at Cartesian points X [3, ...]: rho, v[3], p, B[3]."""
    r = np.sqrt((X * X).sum(axis=0))
    q = (r0 / r) ** 2
    rho = q * (1 + 0.1 * np.sin(0.05 * X[0]) * np.cos(0.04 * X[1]))
    omega = np.array([0.3, -0.2, 0.5])
    rot = np.stack([omega[1] * X[2] - omega[2] * X[1], omega[2] * X[0] - omega[0] * X[2],
                    omega[0] * X[1] - omega[1] * X[0]]) / r
    v = X / r * (1 + 0.1 * np.cos(0.03 * X[2])) + 0.2 * rot
    p = 0.1 * rho * (1 + 0.2 * np.sin(0.05 * X[1]))
    e = np.array([0.0, 0.6, 0.8])
    tor = np.stack([e[1] * X[2] - e[2] * X[1], e[2] * X[0] - e[0] * X[2], e[0] * X[1] - e[1] * X[0]]) / r
    B = q * (X / r + 0.3 * tor * (1 + 0.1 * np.sin(0.02 * X[0])))
    return rho, v, p, B


def conserved(rho, v, p, B):
    E = p / (GAMMA - 1) + 0.5 * rho * (v * v).sum(axis=0) + (B * B).sum(axis=0) / (8 * PI)
    return rho, rho * v, E, B


def flux_through(N, rho, v, p, B):
    """N . F of the ideal-MHD flux tensor for vectors N [3, ...] -> [8, ...]."""
    vn, bn = (v * N).sum(axis=0), (B * N).sum(axis=0)
    pt = p + (B * B).sum(axis=0) / (8 * PI)
    E = p / (GAMMA - 1) + 0.5 * rho * (v * v).sum(axis=0) + (B * B).sum(axis=0) / (8 * PI)
    vb = (v * B).sum(axis=0)
    out = np.empty((8,) + np.shape(rho))
    out[0] = rho * vn
    out[1:4] = rho * vn * v - bn * B / (4 * PI) + pt * N
    out[4] = (E + pt) * vn - bn * vb / (4 * PI)
    out[5:8] = vn * B - bn * v
    return out


def _gauss(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1), 0.5 * w


class Shell:
    def __init__(self, nr, na, r0=21.5, r1=215.0):
        self.nr, self.na, self.r0, self.r1 = nr, na, r0, r1

    def radius(self, xi0):
        return self.r0 + (self.r1 - self.r0) * xi0

    def native_averages(self, p, i_rng, j_rng, k_rng, quad=4):
        """Cell averages in xi of (rho, A^-1 rho v, E, A^-1 B) on the cells
        i_rng x j_rng x k_rng of panel p, as [8][k][j][i]."""
        M = panel_matrix(p)
        gx, gw = _gauss(quad)
        i = np.asarray(i_rng)[None, None, :, None, None, None]
        j = np.asarray(j_rng)[None, :, None, None, None, None]
        k = np.asarray(k_rng)[:, None, None, None, None, None]
        ta = gx[None, None, None, :, None, None]
        tb = gx[None, None, None, None, :, None]
        tr = gx[None, None, None, None, None, :]
        xi1 = (j + ta) / self.na
        xi2 = (k + tb) / self.na
        xi0 = (i + tr) / self.nr
        xi1, xi2, xi0 = np.broadcast_arrays(xi1, xi2, xi0)
        n, d1, d2 = basis(M, xi1, xi2)
        X = self.radius(xi0) * n
        rho, v, pr, B = state(X, self.r0)
        rho, m, E, B = conserved(rho, v, pr, B)
        A = np.stack([n, d1, d2], axis=1)                      # [3 (row), 3 (col), ...]
        Am = np.moveaxis(A, (0, 1), (-2, -1))
        Ai = np.linalg.inv(Am)
        mn = np.einsum("...ab,b...->a...", Ai, m)
        bn = np.einsum("...ab,b...->a...", Ai, B)
        u = np.concatenate([rho[None], mn, E[None], bn])
        w = gw[:, None, None] * gw[None, :, None] * gw[None, None, :]
        return (u * w).sum(axis=(-3, -2, -1))

    def ju_averages(self, p, quad=4):
        """<J U> of every valid cell of panel p (Cartesian vectors), [8][k][j][i]."""
        return self.ju_box(p, (0, 0, 0), (self.nr - 1, self.na - 1, self.na - 1), quad)

    def ju_box(self, p, lo, hi, quad=4):
        """<J U> of the cells lo..hi of panel p (Cartesian vectors), [8][k][j][i]."""
        M = panel_matrix(p)
        gx, gw = _gauss(quad)
        na, nr = self.na, self.nr
        i = np.arange(lo[0], hi[0] + 1)[None, None, :, None, None, None]
        j = np.arange(lo[1], hi[1] + 1)[None, :, None, None, None, None]
        k = np.arange(lo[2], hi[2] + 1)[:, None, None, None, None, None]
        xi1 = (j + gx[None, None, None, :, None, None]) / na
        xi2 = (k + gx[None, None, None, None, :, None]) / na
        xi0 = (i + gx[None, None, None, None, None, :]) / nr
        xi1, xi2, xi0 = np.broadcast_arrays(xi1, xi2, xi0)
        n, d1, d2 = basis(M, xi1, xi2)
        r = self.radius(xi0)
        X = r * n
        detA = np.einsum("a...,a...->...", n, np.cross(d1, d2, axis=0))
        J = r * r * (self.r1 - self.r0) * detA
        rho, m, E, B = conserved(*state(X, self.r0))
        u = np.concatenate([rho[None], m, E[None], B]) * J
        w = gw[:, None, None] * gw[None, :, None] * gw[None, None, :]
        return (u * w).sum(axis=(-3, -2, -1))

    def exact_rhs(self, p, lo, hi, quad=4):
        """-sum_d (<N_d . F>(hi face) - <N_d . F>(lo face)) / dxi_d on the
        cells lo..hi of panel p, [8][k][j][i]."""
        M = panel_matrix(p)
        gx, gw = _gauss(quad)
        n_cells = [hi[d] - lo[d] + 1 for d in range(3)]
        dxi = [1.0 / self.nr, 1.0 / self.na, 1.0 / self.na]
        rhs = np.zeros((8, n_cells[2], n_cells[1], n_cells[0]))
        for d in range(3):
            # faces lo_d .. hi_d+1 in direction d, quadrature over the transverse cell extents
            idx = [np.arange(lo[a], hi[a] + 1 + (a == d)) for a in range(3)]
            t = [(gx if a != d else np.zeros(1)) for a in range(3)]
            wt = [(gw if a != d else np.ones(1)) for a in range(3)]
            # xi[a] of shape [nk, nj, ni, qk, qj, qi]
            i = idx[0][None, None, :, None, None, None] + t[0][None, None, None, None, None, :]
            j = idx[1][None, :, None, None, None, None] + t[1][None, None, None, None, :, None]
            k = idx[2][:, None, None, None, None, None] + t[2][None, None, None, :, None, None]
            xi0, xi1, xi2 = np.broadcast_arrays(i / self.nr, j / self.na, k / self.na)
            n, d1, d2 = basis(M, xi1, xi2)
            r = self.radius(xi0)
            dr = self.r1 - self.r0
            X = r * n
            c0, c1, c2 = dr * n, r * d1, r * d2                  # columns of d x / d xi
            N = [np.cross(c1, c2, axis=0), np.cross(c2, c0, axis=0), np.cross(c0, c1, axis=0)][d]
            f = flux_through(N, *state(X, self.r0))
            w = wt[2][:, None, None] * wt[1][None, :, None] * wt[0][None, None, :]
            fa = (f * w).sum(axis=(-3, -2, -1))                  # face averages [8, nk, nj, ni]
            sl_hi = [slice(None)] * 4
            sl_lo = [slice(None)] * 4
            ax = 3 - d
            sl_hi[ax] = slice(1, None)
            sl_lo[ax] = slice(0, -1)
            rhs -= (fa[tuple(sl_hi)] - fa[tuple(sl_lo)]) / dxi[d]
        return rhs
