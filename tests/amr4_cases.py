"""Small fixed AMR hierarchies of the fourth-order operators (test data).

small_hierarchy(scale): level 0 = 6 x 24 x 16 x 16 (x scale), level 1 (ratio 2)
a 16^3 block of 8^3 patches in panel 1 plus an L-shaped pair in panel 2,
level 2 (ratio 2) a 16^3 block of 8^3 patches inside the panel-1 block: every
patch keeps 6 cells from the radial boundaries and 8 from the panel edges,
interpolates from cells its parent defines and is properly nested.
"""
import numpy as np

import heliocubed_amd as hc

R0, R1 = 21.5, 60.0


def small_hierarchy(scale=1, levels=3):
    s = scale
    shells = [(24 * s, 16 * s), (48 * s, 32 * s), (96 * s, 64 * s)][:levels]
    b = 8 * s
    patches = []
    for k in (8 * s, 16 * s):
        for j in (8 * s, 16 * s):
            for i in (16 * s, 24 * s):
                patches.append((1, (1, (i, j, k), (i + b - 1, j + b - 1, k + b - 1))))
    for j in (8 * s, 16 * s):
        patches.append((1, (2, (16 * s, j, 8 * s), (16 * s + b - 1, j + b - 1, 8 * s + b - 1))))
    if levels > 2:
        for k in (24 * s, 32 * s):
            for j in (24 * s, 32 * s):
                for i in (40 * s, 48 * s):
                    patches.append((2, (1, (i, j, k), (i + b - 1, j + b - 1, k + b - 1))))
    return shells, patches


def options(operator=1, riemann=2, precision=0, powell=0):
    return hc.Options(spatial_order=4, limiter=1, riemann_solver=riemann, spatial_operator=operator,
                      precision=precision, divergence_cleaning=powell)


def hc_shells(shells):
    return [hc.Shell(nr, na, R0, R1) for nr, na in shells]


def initial_state(shells, patches, opts, config=3, seed=11, quad=2):
    """Combined <JU> (level 0, then every patch in table order) and the
    level-0 inner ghosts."""
    sh = hc_shells(shells)
    ju0, inner = hc.init_level4(sh[0], opts, config, seed, quad)
    parts = [ju0.ravel()]
    for lev, box in patches:
        parts.append(hc.init_ju_boxes(sh[lev], opts, [box], config, seed, quad))
    return np.concatenate(parts), inner


def rest_state(shells, patches, rho=2.0, p=0.3, gamma=5.0 / 3.0):
    """<JU> of a uniform state at rest (no field): J <U> with <J> = dVolr <det A>."""
    import oracle as orc

    def level(nr, na, lo, hi, blk):
        xi = np.arange(lo[0], hi[0] + 2) / nr
        r = R0 + (R1 - R0) * xi
        dvol = (r[1:] ** 3 - r[:-1] ** 3) / 3.0 * nr
        cfg = orc.OrcCfg(nr, na, 5, R0, R1, 0.0, gamma, 1, 2, 0.3, 1, 0.02, 0, 0)
        det = np.array([[orc.lib().orc_cell_detA(cfg, j, k) for j in range(lo[1], hi[1] + 1)]
                        for k in range(lo[2], hi[2] + 1)])
        J = det[:, :, None] * dvol[None, None, :]
        u = np.zeros((8,) + J.shape)
        u[0] = rho * J
        u[4] = p / (gamma - 1) * J
        return u

    nr0, na0 = shells[0]
    parts = [np.stack([level(nr0, na0, (0, 0, 0), (nr0 - 1, na0 - 1, na0 - 1), p) for p in range(6)]).ravel()]
    for lev, (blk, lo, hi) in patches:
        nr, na = shells[lev]
        parts.append(level(nr, na, lo, hi, blk).ravel())
    inner = np.zeros((6, 8, na0 + 10, na0 + 10, 5))
    inner[:, 0] = rho
    inner[:, 4] = p / (gamma - 1)
    return np.concatenate(parts), inner


def split(x, shells, patches, stage=False):
    """Level-0 array and per-patch arrays of a combined vector."""
    nr0, na0 = shells[0]
    g = 5 if stage else 0
    n0 = 6 * 8 * (na0 + 2 * g) ** 2 * (nr0 + 2 * g)
    out0 = x[:n0].reshape(6, 8, na0 + 2 * g, na0 + 2 * g, nr0 + 2 * g)
    off, parts = n0, []
    for lev, (blk, lo, hi) in patches:
        n = [hi[d] - lo[d] + 1 + 2 * g for d in range(3)]
        m = 8 * n[0] * n[1] * n[2]
        parts.append(x[off:off + m].reshape(8, n[2], n[1], n[0]))
        off += m
    assert off == x.size
    return out0, parts


def rel_err_per_var(got, ref, shells, patches, stage=False):
    """max |got - ref| / max |ref| per component over the whole hierarchy."""
    g0, gp = split(got, shells, patches, stage)
    r0, rp = split(ref, shells, patches, stage)
    num = np.abs(g0 - r0).max(axis=(0, 2, 3, 4))
    den = np.abs(r0).max(axis=(0, 2, 3, 4))
    for a, b in zip(gp, rp):
        num = np.maximum(num, np.abs(a - b).max(axis=(1, 2, 3)))
        den = np.maximum(den, np.abs(b).max(axis=(1, 2, 3)))
    return num / den
