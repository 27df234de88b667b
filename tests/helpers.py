"""Shared helpers for the parity tests (relative L-infinity per variable)."""
import numpy as np


def rel_linf_per_var(test, ref, axis_var=1):
    """max |test-ref| / max |ref| for each of the 8 variables."""
    t = np.moveaxis(np.asarray(test), axis_var, 0).reshape(8, -1)
    r = np.moveaxis(np.asarray(ref), axis_var, 0).reshape(8, -1)
    scale = np.abs(r).max(axis=1)
    scale[scale == 0] = 1.0
    return np.abs(t - r).max(axis=1) / scale


def cfl_dt(n_ang, r0=21.5, cfl=0.25, speed=1.6):
    """Conservative fixed time step for the radial-wind problems."""
    return cfl * (r0 * np.pi / 2 / n_ang) / speed


def per_var(flat, boxes):
    """Concatenated box-table data (8 components per box) as (8, cells)."""
    out, off = [[] for _ in range(8)], 0
    for _, lo, hi in boxes:
        n = int(np.prod([hi[d] - lo[d] + 1 for d in range(3)]))
        blk = flat[off:off + 8 * n].reshape(8, n)
        for c in range(8):
            out[c].append(blk[c])
        off += 8 * n
    return np.array([np.concatenate(o) for o in out])
