"""Host-side argument checks of the Python mirror: every tensor handed to the
library is a contiguous float64 device tensor of the documented shape, or the
call raises ValueError before anything reaches the device. No device work is
issued (the objects are built without a library handle)."""
import pytest
import torch

import heliocubed_amd as hc


def _level(order=4, n_rad=16, n_ang=16):
    L = object.__new__(hc.Level)
    L.shell = hc.Shell(n_rad, n_ang)
    L.options = hc.Options(spatial_order=order, limiter=1, riemann_solver=2)
    L._handle = None
    g = 5 if order == 4 else 2
    L.ghost = lambda: g
    return L


def test_rk4_advance_checks_inner_ghosts():
    L = _level()
    U = torch.zeros(L.valid_shape, dtype=torch.float64)
    # U itself is a host tensor: rejected
    with pytest.raises(ValueError, match="U"):
        L.rk4_advance(U, torch.zeros(L.inner_shape, dtype=torch.float64), 0.1)


def test_inner_shape_follows_the_ghost_width():
    L4, L2 = _level(4), _level(2)
    assert L4.inner_shape == (6, 8, 26, 26, 5)
    assert L2.inner_shape == (6, 8, 20, 20, 2)


def test_require_rejects_host_wrong_dtype_and_shape():
    L = _level()
    good_shape = L.inner_shape
    with pytest.raises(ValueError, match="inner"):
        L._require(torch.zeros(good_shape, dtype=torch.float64), good_shape, "inner")   # host tensor
    with pytest.raises(ValueError, match="inner"):
        L._require(torch.zeros(good_shape, dtype=torch.float32), good_shape, "inner")
    # an O2-shaped ghost array (g = 2) on a fourth-order level (g = 5)
    with pytest.raises(ValueError, match="inner"):
        L._require(torch.zeros(_level(2).inner_shape, dtype=torch.float64), good_shape, "inner")
    L._require(None, good_shape, "inner")   # optional arguments may be None


def test_boxop_apply_checks_fluxes_before_the_device():
    op = object.__new__(hc.BoxOp)
    op._handle = None
    counts = {0: 10, 1: 4, 2: 5, 3: 5, 4: 5}
    op.count = lambda what: counts[what]
    U, R = torch.zeros(10, dtype=torch.float64), torch.zeros(4, dtype=torch.float64)
    with pytest.raises(ValueError, match="USph"):
        op.apply(U, R)    # host tensors
    with pytest.raises(ValueError, match="fluxes"):
        op.apply(U, R, fluxes=[None, None])
