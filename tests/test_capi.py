"""The C-ABI library loads and exports every function include/heliocubed.h declares.

No device work is issued here: only host-side helpers are called, and the
argument checks that reject bad input before any HIP call.
"""
import ctypes
import os
import re

import numpy as np

import heliocubed_amd as hc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "heliocubed.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hc_[A-Za-z0-9_]+)\s*\(", text)))


def test_every_declared_function_is_exported():
    lib = hc.lib()
    names = declared_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(hc.EXPORTED_SYMBOLS) == names


def test_create_rejects_bad_options_without_device():
    h = ctypes.c_void_p()
    sh = hc.Shell(8, 8).c()
    op = hc.Options(spatial_order=7).c()
    rc = hc.lib().hc_level_create(ctypes.byref(sh), ctypes.byref(op), ctypes.byref(h))
    assert rc != 0 and b"spatial_order" in hc.lib().hc_last_error()
    op = hc.Options(riemann_solver=5).c()
    assert hc.lib().hc_level_create(ctypes.byref(sh), ctypes.byref(op), ctypes.byref(h)) != 0
    sh = hc.Shell(8, 2).c()
    op = hc.Options().c()
    assert hc.lib().hc_level_create(ctypes.byref(sh), ctypes.byref(op), ctypes.byref(h)) != 0
    # Berta operator options: order 4 only, positive detector threshold, fallback 0/1
    sh = hc.Shell(16, 16).c()
    for bad in (dict(spatial_order=2, spatial_operator=2), dict(spatial_operator=3),
                dict(spatial_order=4, riemann_solver=2, spatial_operator=2, berta_eta_d=0.0),
                dict(spatial_order=4, riemann_solver=2, spatial_operator=2, berta_fallback=2)):
        op = hc.Options(**bad).c()
        rc = hc.lib().hc_level_create(ctypes.byref(sh), ctypes.byref(op), ctypes.byref(h))
        assert rc != 0, bad
        assert b"berta" in hc.lib().hc_last_error().lower() or b"operator" in hc.lib().hc_last_error()


def test_init_state_is_seeded_and_physical():
    shell = hc.Shell(8, 6)
    a, ia = hc.init_state(shell, hc.Options(), 0, 123)
    b, ib = hc.init_state(shell, hc.Options(), 0, 123)
    c, _ = hc.init_state(shell, hc.Options(), 0, 124)
    assert np.array_equal(a, b) and np.array_equal(ia, ib)
    assert not np.array_equal(a, c)
    assert (a[:, 0] > 0).all() and np.isfinite(a).all()
    # inner ghosts hold the background density (r0/r)^2 at the ghost centres
    r0, r1, nr = 21.5, 215.0, 8
    dr = (r1 - r0) / nr
    for ig, i in enumerate((-2, -1)):
        r = r0 + (i + 0.5) * dr
        assert np.allclose(ia[:, 0, :, :, ig], (r0 / r) ** 2, rtol=1e-14)


def test_multilevel_box_table_validation_without_device():
    h = ctypes.c_void_p()
    shells = (hc.HcShell * 2)(hc.Shell(32, 16).c(), hc.Shell(64, 32).c())
    op = hc.Options(spatial_order=4, limiter=1, riemann_solver=2).c()
    tab = hc.box_table([(0, (40, 0, 0), (47, 7, 7))])
    lev = (ctypes.c_int * 1)(0)          # radial 40..47 is outside level 0 (32 cells)
    assert hc.lib().hc_boxop_create_levels(shells, 2, ctypes.byref(op), tab, lev, 1, ctypes.byref(h)) != 0
    assert b"outside" in hc.lib().hc_last_error()
    lev = (ctypes.c_int * 1)(2)          # no level 2
    assert hc.lib().hc_boxop_create_levels(shells, 2, ctypes.byref(op), tab, lev, 1, ctypes.byref(h)) != 0
    assert b"level" in hc.lib().hc_last_error()


def test_tile_calls_reject_null_level_without_device():
    assert hc.lib().hc_level_set_tile(None, 32) != 0
    assert b"null" in hc.lib().hc_last_error()
    assert hc.lib().hc_level_tile(None) == 0
    t, ms = (ctypes.c_int * 3)(), ctypes.c_double()
    assert hc.lib().hc_level_autotune(None, None, None, t, ctypes.byref(ms), None) != 0
    assert hc.lib().hc_level_set_tile_shape(None, 32, 32, 32) != 0
    assert hc.lib().hc_level_tile_shape(None, t) != 0


def test_empty_box_table_is_rejected_without_device():
    h = ctypes.c_void_p()
    sh = hc.Shell(32, 16).c()
    op = hc.Options(spatial_order=4, limiter=1, riemann_solver=2).c()
    tab = hc.box_table([(0, (0, 0, 0), (7, 7, 7))])
    assert hc.lib().hc_boxop_create(ctypes.byref(sh), ctypes.byref(op), tab, 0, ctypes.byref(h)) != 0
    assert b"bad argument" in hc.lib().hc_last_error()


def test_box_table_rejects_small_angular_levels_without_device():
    """The metric tables of a box table span the angular cells -5 .. n_ang+5,
    which leave the panel's horizon below n_ang = 16: such levels are rejected."""
    h = ctypes.c_void_p()
    op = hc.Options(spatial_order=4, limiter=1, riemann_solver=2).c()
    tab = hc.box_table([(0, (0, 0, 0), (7, 7, 7))])
    for na in (8, 12, 15):
        sh = hc.Shell(32, na).c()
        assert hc.lib().hc_boxop_create(ctypes.byref(sh), ctypes.byref(op), tab, 1, ctypes.byref(h)) != 0
        assert b"n_ang >= 16" in hc.lib().hc_last_error()
