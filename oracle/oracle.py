"""ctypes wrapper of the CPU FP64 oracle (oracle/hc_oracle.cpp).

TEST INFRASTRUCTURE ONLY: tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py use it as the checker or the timed CPU
baseline. The product (heliocubed_amd/) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libhcoracle.so")


class OrcCfg(ctypes.Structure):
    _fields_ = [
        ("n_rad", ctypes.c_int),
        ("n_ang", ctypes.c_int),
        ("ghost", ctypes.c_int),
        ("r0", ctypes.c_double),
        ("r1", ctypes.c_double),
        ("radial_stretch", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("limiter", ctypes.c_int),
        ("riemann", ctypes.c_int),
        ("entropy_fix", ctypes.c_double),
        ("spatial_operator", ctypes.c_int),
        ("berta_eta_d", ctypes.c_double),
        ("berta_fallback", ctypes.c_int),
        ("divergence_cleaning", ctypes.c_int),
    ]


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P, D, I = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        cfgp = ctypes.POINTER(OrcCfg)
        sig = {
            "orc_create": (P, [cfgp]),
            "orc_destroy": (None, [P]),
            "orc_o2_prepare": (I, [P, P, P, P]),
            "orc_o2_rhs": (I, [P, P, P, P, P, P]),
            "orc_o2_rk4": (I, [P, P, P, D, I]),
            "orc_cell_detA": (D, [cfgp, I, I]),
            "orc_face_adjA": (None, [cfgp, I, I, I, I, P]),
            "orc_basis": (None, [I, D, D, P]),
            "orc_roe8": (I, [P, P, P, D, D, I]),
            "orc_rusanov": (None, [P, P, P, D]),
            "orc_plm": (None, [P, P, P, P, P, P, I]),
            "orc_face_flux": (None, [P, P, P, P, D, D, D, I]),
            "orc_cartesian_primitive": (None, [P, P, D, D]),
            "orc_ghost_donors": (I, [cfgp, I, I, I, P, P, P, P]),
            "orc_seam_edge": (None, [cfgp, I, I, I, P]),
            "orc_o4_rhs": (I, [cfgp, I, P, P, P, P, P, P, P]),
            "orc_o2_box_rhs": (I, [cfgp, I, P, P, P, P, P, P, P, D]),
            "orc_o4_rhs_table": (I, [cfgp, I, P, P, P]),
            "orc_limit_line": (None, [P, P]),
            "orc_eta_tilde": (D, [P, P]),
            "orc_minmod_cell": (None, [D, D, D, D, D, D, P]),
            "orc_minmod_width": (D, [cfgp, I, I, I, I, I]),
            "orc_prim_flux": (None, [P, P, D, I, P, P]),
            "orc_factored_roe": (None, [P, P, P, P, D, I, D, D]),
            "orc_face_average": (None, [cfgp, I, I, I, I, I, P]),
            "orc4_create": (P, [cfgp]),
            "orc4_destroy": (None, [P]),
            "orc4_prepare": (I, [P, P, P, P]),
            "orc4_rhs": (I, [P, P, P]),
            "orc4_rk4": (I, [P, P, P, D, I]),
            "orc4_rk4_cons": (I, [P, P, P, D, I, P]),
            "orc4_ghost_donors": (I, [cfgp, I, I, I, P, P, P, P]),
            "orc_wenoz": (None, [P, P]),
            "orc_berta_theta": (None, [cfgp, I, P, P, P, P]),
            "orc_weno3": (None, [P, P, D]),
            "orc_plm_berta": (None, [P, P]),
            "orc_dt_cfl": (D, [cfgp, I, P, D]),
            "orc_first_invalid": (ctypes.c_longlong, [P, I, I]),
            "orc_inner_ghosts_from_bc": (None, [cfgp, I, P, P]),
            "orc_inner_ghosts_outflow": (None, [cfgp, I, D, D, P]),
            "orc_amr4_weights": (None, [I, I, P]),
            "orc_amr_interp4": (None, [P, P, P, I, P, P, P, P]),
            "orc_amr_ghosts2": (None, [P, P, P, P, P, P, P, P, D, P]),
            "orc_amr_average_down": (None, [P, P, P, I, P, P, P, P]),
            "orc_amr_prolong4": (I, [P, P, P, P, P, P, P, P, P]),
            "orc_amr4_create": (P, [cfgp, I, P, I, ctypes.c_char_p, I]),
            "orc_bc_frame_at": (None, [P, I, I, I, D, I, D, P]),
            "orc_bc_points": (None, [cfgp, P, I, I, P, P, P]),
            "orc_bc_averages": (None, [cfgp, I, P, I, I, P, P, P]),
            "orc_amr4_destroy": (None, [P]),
            "orc_amr4_count": (ctypes.c_longlong, [P, I]),
            "orc_amr4_prepare": (I, [P, P, P, P]),
            "orc_amr4_stage_rhs": (I, [P, P, P, P]),
            "orc_amr4_rk4": (I, [P, P, P, D, I]),
            "orc_amr4_stage_rhs_opt": (I, [P, P, P, P, I]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray):
    """Data pointer of a C-contiguous float64 array. The caller must keep the
    array referenced until the library call returns (convert first, then
    pass the pointer)."""
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data


def _c(x):
    return np.ascontiguousarray(x, dtype=np.float64)


class Oracle:
    """O2 operator of one shell configuration on the CPU (FP64)."""

    def __init__(self, n_rad, n_ang, r0=21.5, r1=215.0, radial_stretch=0.0, gamma=5.0 / 3.0,
                 limiter=1, riemann=1, entropy_fix=0.3, ghost=2, divergence_cleaning=0):
        self.cfg = OrcCfg(n_rad, n_ang, ghost, r0, r1, radial_stretch, gamma, limiter, riemann, entropy_fix)
        self.cfg.divergence_cleaning = divergence_cleaning
        self.nr, self.na, self.g = n_rad, n_ang, ghost
        self._h = ctypes.c_void_p(lib().orc_create(ctypes.byref(self.cfg)))

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().orc_destroy(self._h)
            self._h = ctypes.c_void_p()

    @property
    def valid_shape(self):
        return (6, 8, self.na, self.na, self.nr)

    @property
    def stage_shape(self):
        g = self.g
        return (6, 8, self.na + 2 * g, self.na + 2 * g, self.nr + 2 * g)

    def face_shape(self, d):
        return (6, 8, self.na + (d == 2), self.na + (d == 1), self.nr + (d == 0))

    def prepare(self, ju, inner):
        W = np.zeros(self.stage_shape)
        lib().orc_o2_prepare(self._h, _p(ju), _p(inner), _p(W))
        return W

    def rhs(self, W, with_fluxes=False):
        k = np.zeros(self.valid_shape)
        if with_fluxes:
            F = [np.zeros(self.face_shape(d)) for d in range(3)]
            lib().orc_o2_rhs(self._h, _p(W), _p(k), _p(F[0]), _p(F[1]), _p(F[2]))
            return k, F
        lib().orc_o2_rhs(self._h, _p(W), _p(k), None, None, None)
        return k

    def rk4(self, ju, inner, dt, nsteps=1):
        u = np.ascontiguousarray(ju.copy())
        lib().orc_o2_rk4(self._h, _p(u), _p(inner), dt, nsteps)
        return u

    def cell_detA(self, j, k):
        return lib().orc_cell_detA(ctypes.byref(self.cfg), j, k)

    def face_adjA(self, p, d, j, k):
        n = np.zeros(3)
        lib().orc_face_adjA(ctypes.byref(self.cfg), p, d, j, k, _p(n))
        return n

    def ghost_donors(self, p, j, k):
        q = np.zeros(16, dtype=np.int32)
        jj = np.zeros(16, dtype=np.int32)
        kk = np.zeros(16, dtype=np.int32)
        w = np.zeros(16)
        n = lib().orc_ghost_donors(ctypes.byref(self.cfg), p, j, k, q.ctypes.data, jj.ctypes.data,
                                   kk.ctypes.data, _p(w))
        return [(int(q[x]), int(jj[x]), int(kk[x]), float(w[x])) for x in range(n)]

    def seam_edge(self, p, d, s):
        out = np.zeros(4, dtype=np.int32)
        lib().orc_seam_edge(ctypes.byref(self.cfg), p, d, s, out.ctypes.data)
        return tuple(int(x) for x in out)


def roe8(left, right, gamma=5.0 / 3.0, entropy_fix=0.3, check=True):
    out = np.zeros(8)
    l, r = _c(left), _c(right)
    ok = lib().orc_roe8(_p(out), _p(l), _p(r), gamma, entropy_fix, int(check))
    return bool(ok), out


def rusanov(left, right, gamma=5.0 / 3.0):
    out = np.zeros(8)
    l, r = _c(left), _c(right)
    lib().orc_rusanov(_p(out), _p(l), _p(r), gamma)
    return out


def plm(mm, m, c, p, limiter=1):
    lo, hi = np.zeros(8), np.zeros(8)
    args = [np.ascontiguousarray(x, dtype=np.float64) for x in (mm, m, c, p)]
    lib().orc_plm(_p(lo), _p(hi), *[_p(a) for a in args], limiter)
    return lo, hi


def face_flux(left, right, normal, area, gamma=5.0 / 3.0, entropy_fix=0.3, roe=False):
    out = np.zeros(8)
    l, r, n = _c(left), _c(right), _c(normal)
    lib().orc_face_flux(_p(out), _p(l), _p(r), _p(n), area, gamma, entropy_fix, int(roe))
    return out


def berta_theta(n_rad, n_ang, p, lo, hi, U, eta_d=0.02, gamma=5.0 / 3.0, r0=21.5, r1=215.0):
    """Berta Jameson flag on the box grown by 5 ([n2+10][n1+10][n0+10])."""
    cfg = OrcCfg(n_rad, n_ang, 5, r0, r1, 0.0, gamma, 1, 2, 0.3, 2, eta_d, 0)
    n = [hi[d] - lo[d] + 11 for d in range(3)]
    th = np.zeros((n[2], n[1], n[0]))
    u = _c(U)
    lo_a = np.array(lo, dtype=np.int32)
    hi_a = np.array(hi, dtype=np.int32)
    lib().orc_berta_theta(ctypes.byref(cfg), p, lo_a.ctypes.data, hi_a.ctypes.data,
                          _p(u), _p(th))
    return th


def berta_face(values, method="wenoz", dx2=0.0):
    """(left, right) point values of one component at the face between the
    third and fourth of six cells (WENOZ) or the second and third of four
    (WENO3, PLM)."""
    out = np.zeros(2)
    v = np.ascontiguousarray(values, dtype=np.float64)
    if method == "wenoz":
        lib().orc_wenoz(_p(out), _p(v))
    elif method == "weno3":
        lib().orc_weno3(_p(out), _p(v), dx2)
    else:
        lib().orc_plm_berta(_p(out), _p(v))
    return out[0], out[1]


def dt_cfl(n_rad, n_ang, JU, order=4, cone_density=0.0, gamma=5.0 / 3.0, r0=21.5, r1=215.0):
    """1 / max over valid cells of the CFL speed sum of a level state JU
    [6][8][na][na][nr] (dtCFLSecondOrderHost at order 2, dtCFL at order 4);
    the caller multiplies by its CFL number."""
    cfg = OrcCfg(n_rad, n_ang, 2 if order == 2 else 5, r0, r1, 0.0, gamma, 1, 2, 0.3, 1, 0.02, 0)
    ju = np.ascontiguousarray(JU, dtype=np.float64)
    assert ju.shape == (6, 8, n_ang, n_ang, n_rad), ju.shape
    return lib().orc_dt_cfl(ctypes.byref(cfg), order, _p(ju), cone_density)


def first_invalid(JU):
    """Flat index of the first non-finite value or non-positive density of a
    valid level state [6][8][na][na][nr] (validConservedState), or -1."""
    ju = np.ascontiguousarray(JU, dtype=np.float64)
    na, nr = ju.shape[2], ju.shape[4]
    return int(lib().orc_first_invalid(_p(ju), nr, na))


def inner_ghosts_from_bc(n_rad, n_ang, bc, order=4, gamma=5.0 / 3.0, r0=21.5, r1=215.0):
    """Inner radial ghosts [6][8][na+2g][na+2g][g] of a level (g = 2 at order 2,
    5 at order 4) from one strip of boundary-file primitives bc[6][8][na+2g][na+2g]."""
    g = 2 if order == 2 else 5
    cfg = OrcCfg(n_rad, n_ang, g, r0, r1, 0.0, gamma, 1, 2, 0.3, 1, 0.02, 0)
    n = n_ang + 2 * g
    b = np.ascontiguousarray(bc, dtype=np.float64)
    assert b.shape == (6, 8, n, n), b.shape
    out = np.zeros((6, 8, n, n, g))
    lib().orc_inner_ghosts_from_bc(ctypes.byref(cfg), order, _p(b), _p(out))
    return out


def inner_ghosts_outflow(n_rad, n_ang, strength, time, order=4, gamma=5.0 / 3.0, r0=21.5, r1=215.0):
    """Inner radial ghosts of the radial-outflow test boundary at `time` (s)."""
    g = 2 if order == 2 else 5
    cfg = OrcCfg(n_rad, n_ang, g, r0, r1, 0.0, gamma, 1, 2, 0.3, 1, 0.02, 0)
    n = n_ang + 2 * g
    out = np.zeros((6, 8, n, n, g))
    lib().orc_inner_ghosts_outflow(ctypes.byref(cfg), order, strength, time, _p(out))
    return out


def _ivec(v):
    return np.ascontiguousarray(v, dtype=np.int32)


def _box_shape(lo, hi, nc):
    return (nc, hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)


def amr4_weights(ratio, child):
    w = np.zeros(5)
    lib().orc_amr4_weights(ratio, child, _p(w))
    return w


def amr_interp4(coarse, coarse_lo, coarse_hi, ratio, fine_lo, fine_hi):
    """Fine averages [nc][k][j][i] on fine_lo..fine_hi from coarse averages
    [nc][k][j][i] on coarse_lo..coarse_hi (amrInterpolateFourthOrder)."""
    src = np.ascontiguousarray(coarse, dtype=np.float64)
    nc = src.shape[0]
    assert src.shape == _box_shape(coarse_lo, coarse_hi, nc)
    out = np.zeros(_box_shape(fine_lo, fine_hi, nc))
    args = [_ivec(x) for x in (coarse_lo, coarse_hi, ratio, fine_lo, fine_hi)]
    lib().orc_amr_interp4(_p(src), args[0].ctypes.data, args[1].ctypes.data, nc, args[2].ctypes.data,
                          args[3].ctypes.data, args[4].ctypes.data, _p(out))
    return out


def amr_ghosts2(coarse_ju_j, coarse_lo, coarse_hi, ratio, fine_lo, fine_hi, owned_lo, owned_hi, fine_init,
                gamma=5.0 / 3.0):
    """Second-order fine ghost cells (Cartesian primitives) written into a copy
    of fine_init [8][k][j][i] on fine_lo..fine_hi."""
    src = np.ascontiguousarray(coarse_ju_j, dtype=np.float64)
    assert src.shape == _box_shape(coarse_lo, coarse_hi, 9)
    out = np.ascontiguousarray(fine_init, dtype=np.float64).copy()
    assert out.shape == _box_shape(fine_lo, fine_hi, 8)
    args = [_ivec(x) for x in (coarse_lo, coarse_hi, ratio, fine_lo, fine_hi, owned_lo, owned_hi)]
    lib().orc_amr_ghosts2(_p(src), *[a.ctypes.data for a in args], gamma, _p(out))
    return out


def amr_average_down(fine, fine_lo, fine_hi, ratio, coarse_lo, coarse_hi):
    src = np.ascontiguousarray(fine, dtype=np.float64)
    nc = src.shape[0]
    out = np.zeros(_box_shape(coarse_lo, coarse_hi, nc))
    args = [_ivec(x) for x in (fine_lo, fine_hi, ratio, coarse_lo, coarse_hi)]
    lib().orc_amr_average_down(_p(src), args[0].ctypes.data, args[1].ctypes.data, nc, args[2].ctypes.data,
                               args[3].ctypes.data, args[4].ctypes.data, _p(out))
    return out


def amr_prolong4(interp, fine_j, fine_lo, fine_hi, coarse, coarse_lo, coarse_hi, ratio):
    """(fine J*U [8][..], invalid parent count) of the conservative fourth-order refinement."""
    it = np.ascontiguousarray(interp, dtype=np.float64)
    fj = np.ascontiguousarray(fine_j, dtype=np.float64)
    co = np.ascontiguousarray(coarse, dtype=np.float64)
    out = np.zeros(_box_shape(fine_lo, fine_hi, 8))
    args = [_ivec(x) for x in (fine_lo, fine_hi, coarse_lo, coarse_hi, ratio)]
    bad = lib().orc_amr_prolong4(_p(it), _p(fj), args[0].ctypes.data, args[1].ctypes.data, _p(co),
                                 args[2].ctypes.data, args[3].ctypes.data, args[4].ctypes.data, _p(out))
    return out, bad


def basis(p, alpha, beta):
    A = np.zeros(9)
    lib().orc_basis(p, alpha, beta, _p(A))
    return A.reshape(3, 3)


def o4_rhs_boxes(n_rad, n_ang, boxes, USph, limiter=1, entropy_fix=0.3, gamma=5.0 / 3.0, r0=21.5, r1=215.0,
                 with_fluxes=False, spatial_operator=1, eta_d=0.02, fallback=0, riemann=2, divergence_cleaning=0):
    """Fourth-order RHS (and optionally fluxes) of every box, concatenated like hc_boxop.

    spatial_operator 1 is the standard operator, 2 the Berta WENOZ operator
    (Jameson detector threshold eta_d, fallback 0 WENO3 / 1 PLM); riemann 2
    Roe 8-wave, 1 Rusanov."""
    cfg = OrcCfg(n_rad, n_ang, 5, r0, r1, 0.0, gamma, limiter, riemann, entropy_fix, spatial_operator, eta_d,
                 fallback, divergence_cleaning)
    g = 5
    rhs_parts, flux_parts, off = [], [[], [], []], 0
    for blk, lo, hi in boxes:
        n = [hi[d] - lo[d] + 1 for d in range(3)]
        size = 8 * (n[0] + 2 * g) * (n[1] + 2 * g) * (n[2] + 2 * g)
        u = np.ascontiguousarray(USph[off:off + size])
        off += size
        r = np.zeros(8 * n[0] * n[1] * n[2])
        lo_a = np.array(lo, dtype=np.int32)
        hi_a = np.array(hi, dtype=np.int32)
        F = [np.zeros(8 * (n[0] + (d == 0)) * (n[1] + (d == 1)) * (n[2] + (d == 2))) for d in range(3)]
        fp = [_p(F[d]) if with_fluxes else None for d in range(3)]
        lib().orc_o4_rhs(ctypes.byref(cfg), blk, lo_a.ctypes.data, hi_a.ctypes.data, _p(u), _p(r), *fp)
        rhs_parts.append(r)
        for d in range(3):
            flux_parts[d].append(F[d])
    rhs = np.concatenate(rhs_parts)
    if with_fluxes:
        return rhs, [np.concatenate(flux_parts[d]) for d in range(3)]
    return rhs


def o2_rhs_boxes(n_rad, n_ang, boxes, W, limiter=1, riemann=1, entropy_fix=0.3, gamma=5.0 / 3.0, r0=21.5, r1=215.0,
                 scale=1.0, with_fluxes=False, divergence_cleaning=0):
    """Second-order per-box RHS (secondOrderOperator with Cartesian primitives):
    W concatenates the boxes' Cartesian primitives on the box grown by 2; the
    RHS of every box (accumulated from zero) and optionally the face fluxes."""
    cfg = OrcCfg(n_rad, n_ang, 2, r0, r1, 0.0, gamma, limiter, riemann, entropy_fix, 1, 0.02, 0,
                 divergence_cleaning)
    g = 2
    rhs_parts, flux_parts, off = [], [[], [], []], 0
    for blk, lo, hi in boxes:
        n = [hi[d] - lo[d] + 1 for d in range(3)]
        size = 8 * (n[0] + 2 * g) * (n[1] + 2 * g) * (n[2] + 2 * g)
        w = np.ascontiguousarray(W[off:off + size], dtype=np.float64)
        off += size
        r = np.zeros(8 * n[0] * n[1] * n[2])
        lo_a, hi_a = np.array(lo, dtype=np.int32), np.array(hi, dtype=np.int32)
        F = [np.zeros(8 * (n[0] + (d == 0)) * (n[1] + (d == 1)) * (n[2] + (d == 2))) for d in range(3)]
        fp = [_p(F[d]) if with_fluxes else None for d in range(3)]
        lib().orc_o2_box_rhs(ctypes.byref(cfg), blk, lo_a.ctypes.data, hi_a.ctypes.data, _p(w), _p(r), *fp,
                             float(scale))
        rhs_parts.append(r)
        for d in range(3):
            flux_parts[d].append(F[d])
    rhs = np.concatenate(rhs_parts)
    if with_fluxes:
        return rhs, [np.concatenate(flux_parts[d]) for d in range(3)]
    return rhs


_opc = None


def opcount_box(n_rad, n_ang, box, USph, spatial_operator=1, limiter=1, riemann=2, eta_d=0.02, gamma=5.0 / 3.0,
                r0=21.5, r1=215.0):
    """(rhs, counts) of the box operator on one box from the operation-counting
    build: counts = {add, mul, div, sqrt, cmp, other} over the box (the
    geometry helpers excluded)."""
    global _opc
    if _opc is None:
        path = os.path.join(_HERE, "_build", "libhcopcount.so")
        if not os.path.exists(path):
            build()
        _opc = ctypes.CDLL(path)
        _opc.orc_opcount_box.argtypes = [ctypes.POINTER(OrcCfg), ctypes.c_int] + [ctypes.c_void_p] * 5
        _opc.orc_opcount_box.restype = None
    blk, lo, hi = box
    cfg = OrcCfg(n_rad, n_ang, 5, r0, r1, 0.0, gamma, limiter, riemann, 0.3, spatial_operator, eta_d, 0, 0)
    lo_a, hi_a = np.array(lo, dtype=np.int32), np.array(hi, dtype=np.int32)
    u = _c(USph)
    rhs = np.zeros(8 * np.prod([hi[d] - lo[d] + 1 for d in range(3)]))
    cnt = np.zeros(6, dtype=np.int64)
    _opc.orc_opcount_box(ctypes.byref(cfg), blk, lo_a.ctypes.data, hi_a.ctypes.data, _p(u), _p(rhs),
                         cnt.ctypes.data)
    return rhs, dict(zip(("add", "mul", "div", "sqrt", "cmp", "other"), cnt.tolist()))


def limit_line(w7):
    """(left state of the high face, right state of the low face) of the
    middle cell of a 7-cell line after Limiters_4O's limiter."""
    out, w = np.zeros(2), np.ascontiguousarray(w7, dtype=np.float64)
    lib().orc_limit_line(_p(w), _p(out))
    return out[0], out[1]


def minmod_cell(wb, w, wa, dxb, dx, dxa):
    """(W + slope, W - slope) of Limiters_minmod's minmod_calc for one component."""
    out = np.zeros(2)
    lib().orc_minmod_cell(wb, w, wa, dxb, dx, dxa, _p(out))
    return out[0], out[1]


def minmod_width(n_rad, n_ang, p, d, i, j, k, r0=21.5, r1=215.0, stretch=0.0):
    """Cell width along d of the minmod limiter at index (i, j, k) of panel p."""
    cfg = OrcCfg(n_rad, n_ang, 5, r0, r1, stretch, 5.0 / 3.0, 2, 2, 0.3, 1, 0.02, 0, 0)
    return lib().orc_minmod_width(ctypes.byref(cfg), p, d, i, j, k)


def eta_tilde(p5, v5):
    p, v = np.ascontiguousarray(p5, dtype=np.float64), np.ascontiguousarray(v5, dtype=np.float64)
    return lib().orc_eta_tilde(_p(p), _p(v))


def prim_flux(A, w, d, gamma=5.0 / 3.0):
    """(f[8], (ke, me, v.B)) of the fourth-order point flux of native state w."""
    f, kmv = np.zeros(8), np.zeros(3)
    a, s = np.ascontiguousarray(np.ravel(A), dtype=np.float64), np.ascontiguousarray(w, dtype=np.float64)
    lib().orc_prim_flux(_p(a), _p(s), gamma, d, _p(f), _p(kmv))
    return f, kmv


def factored_roe(wl, wr, A, radial_area, d, gamma=5.0 / 3.0, entropy_fix=0.3):
    out = np.zeros(8)
    l, r = np.ascontiguousarray(wl, dtype=np.float64), np.ascontiguousarray(wr, dtype=np.float64)
    a = np.ascontiguousarray(np.ravel(A), dtype=np.float64)
    lib().orc_factored_roe(_p(out), _p(l), _p(r), _p(a), radial_area, d, gamma, entropy_fix)
    return out


def o4_rhs_table(n_rad, n_ang, boxes, USph, limiter=1, entropy_fix=0.3, gamma=5.0 / 3.0, r0=21.5, r1=215.0,
                 spatial_operator=1, eta_d=0.02, fallback=0, riemann=2):
    """o4_rhs_boxes without fluxes, the boxes spread over the OpenMP threads."""
    cfg = OrcCfg(n_rad, n_ang, 5, r0, r1, 0.0, gamma, limiter, riemann, entropy_fix, spatial_operator, eta_d,
                 fallback, 0)
    tab = np.ascontiguousarray([[blk, *lo, *hi] for blk, lo, hi in boxes], dtype=np.int32)
    u = np.ascontiguousarray(USph, dtype=np.float64)
    rhs = np.zeros(sum(8 * np.prod([hi[d] - lo[d] + 1 for d in range(3)]) for _, lo, hi in boxes))
    lib().orc_o4_rhs_table(ctypes.byref(cfg), len(boxes), tab.ctypes.data, _p(u), _p(rhs))
    return rhs


class Oracle4:
    """Fourth-order level (standard or Berta operator + Roe, state_conversion 1) on the CPU."""

    def __init__(self, n_rad, n_ang, limiter=1, entropy_fix=0.3, gamma=5.0 / 3.0, r0=21.5, r1=215.0,
                 spatial_operator=1, eta_d=0.02, fallback=0, riemann=2, divergence_cleaning=0):
        self.cfg = OrcCfg(n_rad, n_ang, 5, r0, r1, 0.0, gamma, limiter, riemann, entropy_fix, spatial_operator,
                          eta_d, fallback, divergence_cleaning)
        self.nr, self.na = n_rad, n_ang
        self._h = ctypes.c_void_p(lib().orc4_create(ctypes.byref(self.cfg)))

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().orc4_destroy(self._h)
            self._h = ctypes.c_void_p()

    @property
    def stage_shape(self):
        return (6, 8, self.na + 10, self.na + 10, self.nr + 10)

    def prepare(self, ju, inner):
        W = np.zeros(self.stage_shape)
        a, b = _c(ju), _c(inner)
        lib().orc4_prepare(self._h, _p(a), _p(b), _p(W))
        return W

    def rhs(self, W):
        k = np.zeros((6, 8, self.na, self.na, self.nr))
        w = _c(W)
        lib().orc4_rhs(self._h, _p(w), _p(k))
        return k

    def rk4(self, ju, inner, dt, nsteps=1):
        u = np.ascontiguousarray(ju.copy())
        g = _c(inner)
        lib().orc4_rk4(self._h, _p(u), _p(g), dt, nsteps)
        return u

    def rk4_cons(self, ju, inner, dt, nsteps=1):
        """RK4 steps and the radialCons sums of the radial boundary fluxes."""
        u = np.ascontiguousarray(ju.copy())
        g = _c(inner)
        cons = np.zeros(8)
        lib().orc4_rk4_cons(self._h, _p(u), _p(g), dt, nsteps, _p(cons))
        return u, cons


class Amr4Oracle:
    """Fixed AMR hierarchy of fourth-order levels with synchronous classical
    RK4 (advanceRungeKuttaSynchronous) on the CPU. levels: [(n_rad, n_ang)]
    from the base shell up; patches: [(level, (panel, lo, hi))] for levels
    >= 1. Combined arrays: the level-0 array, then every patch in table order."""

    def __init__(self, levels, patches, limiter=1, entropy_fix=0.3, gamma=5.0 / 3.0, r0=21.5, r1=215.0,
                 spatial_operator=1, eta_d=0.02, fallback=0, riemann=2, divergence_cleaning=0):
        self.levels = [tuple(x) for x in levels]
        self.patches = [(int(l), (int(b), tuple(lo), tuple(hi))) for l, (b, lo, hi) in patches]
        arr = (OrcCfg * len(levels))()
        for i, (nr, na) in enumerate(levels):
            arr[i] = OrcCfg(nr, na, 5, r0, r1, 0.0, gamma, limiter, riemann, entropy_fix, spatial_operator, eta_d,
                            fallback, divergence_cleaning)
        self._cfgs = arr
        tab = np.ascontiguousarray([[l, b, *lo, *hi] for l, (b, lo, hi) in self.patches], dtype=np.int32)
        self._tab = tab
        err = ctypes.create_string_buffer(256)
        h = lib().orc_amr4_create(arr, len(levels), tab.ctypes.data if len(tab) else None, len(self.patches),
                                  err, 256)
        if not h:
            raise ValueError(err.value.decode())
        self._h = ctypes.c_void_p(h)
        self.state_size = int(lib().orc_amr4_count(self._h, 0))
        self.stage_size = int(lib().orc_amr4_count(self._h, 1))

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().orc_amr4_destroy(self._h)
            self._h = ctypes.c_void_p()

    def prepare(self, ju, inner):
        us = np.zeros(self.stage_size)
        a, b = _c(ju), _c(inner)
        lib().orc_amr4_prepare(self._h, _p(a), _p(b), _p(us))
        return us

    def stage_rhs(self, ju, inner, reflux=True):
        k = np.zeros(self.state_size)
        a, b = _c(ju), _c(inner)
        lib().orc_amr4_stage_rhs_opt(self._h, _p(a), _p(b), _p(k), int(reflux))
        return k

    def rk4(self, ju, inner, dt, nsteps=1):
        u = np.ascontiguousarray(np.array(ju, dtype=np.float64).ravel().copy())
        g = _c(inner)
        lib().orc_amr4_rk4(self._h, _p(u), _p(g), dt, nsteps)
        return u


def bc_frame_at(frames, cadence_hours, frame_rotate, time):
    """BoxData_to_BC's frame step on [nframes][8][nphi][ntheta] host frames."""
    f = _c(frames)
    nf, _, npf, nt = f.shape
    out = np.zeros((8, npf, nt))
    lib().orc_bc_frame_at(_p(f), nf, nt, npf, float(cadence_hours), int(frame_rotate), float(time), _p(out))
    return out


def _bc_cfg(n_rad, n_ang, ghost, r0, r1, gamma):
    return OrcCfg(n_rad, n_ang, ghost, r0, r1, 0.0, gamma, 1, 2, 0.3, 1, 0.02, 0, 0)


def bc_points(n_rad, n_ang, ghost, data0, theta, phi, r0=21.5, r1=215.0, gamma=5.0 / 3.0):
    """fillPoints onto the inner strip [6][8][na+2g][na+2g]."""
    d, t, p = _c(data0), _c(theta), _c(phi)
    n = n_ang + 2 * ghost
    out = np.zeros((6, 8, n, n))
    cfg = _bc_cfg(n_rad, n_ang, ghost, r0, r1, gamma)
    lib().orc_bc_points(ctypes.byref(cfg), _p(d), t.size, p.size, _p(t), _p(p), _p(out))
    return out


def bc_averages(n_rad, n_ang, ghost, order, data0, theta, phi, r0=21.5, r1=215.0, gamma=5.0 / 3.0):
    """fill: native conserved averages of the inner ghosts [6][8][na+2g][na+2g][g]."""
    d, t, p = _c(data0), _c(theta), _c(phi)
    n = n_ang + 2 * ghost
    out = np.zeros((6, 8, n, n, ghost))
    cfg = _bc_cfg(n_rad, n_ang, ghost, r0, r1, gamma)
    lib().orc_bc_averages(ctypes.byref(cfg), int(order), _p(d), t.size, p.size, _p(t), _p(p), _p(out))
    return out
