"""HelioCubed per-box MHD right-hand side on MI355X.

Python host side of the C ABI in ``include/heliocubed.h``. The compute path is
the HIP library ``heliocubed_amd/libheliocubed_amd.so`` (hand-written gfx950
kernels). Device memory and streams come from PyTorch; nothing here falls back
to a CPU implementation: if the library is missing the import of
:func:`lib` raises.

Reference interface mirrored (src/ = HelioCubed reference repository):

* :class:`Level` plays the role of ``BoxOp_EulerCubedSphere``'s level entry
  points ``preStageLevel`` (src/BoxOp_PreStage.H:427-501) and ``StageLevel``
  (src/BoxOp_EulerCubedSphereImplem.H:955-1012) plus ``MBLevelRK::advance``
  (src/MHD_LevelRK.H:22-29).
* ``Level.ghost()`` is ``BoxOp_EulerCubedSphere::ghost()``
  (src/BoxOp_EulerCubedSphere.H:376-390).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libheliocubed_amd.so")


class HcShell(ctypes.Structure):
    _fields_ = [
        ("n_rad", ctypes.c_int),
        ("n_ang", ctypes.c_int),
        ("inner_radius", ctypes.c_double),
        ("outer_radius", ctypes.c_double),
        ("radial_stretch", ctypes.c_double),
    ]


class HcOptions(ctypes.Structure):
    _fields_ = [
        ("spatial_order", ctypes.c_int),
        ("limiter", ctypes.c_int),
        ("riemann_solver", ctypes.c_int),
        ("entropy_fix", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("precision", ctypes.c_int),
        ("spatial_operator", ctypes.c_int),
        ("berta_eta_d", ctypes.c_double),
        ("berta_fallback", ctypes.c_int),
        ("divergence_cleaning", ctypes.c_int),
    ]


class HcBox(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int), ("lo", ctypes.c_int * 3), ("hi", ctypes.c_int * 3)]


# (name, restype, argtypes) of every exported C-ABI symbol.
_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int
_SIGNATURES = [
    ("hc_level_create", _I, [ctypes.POINTER(HcShell), ctypes.POINTER(HcOptions), ctypes.POINTER(_P)]),
    ("hc_level_destroy", None, [_P]),
    ("hc_level_ghost", _I, [_P]),
    ("hc_level_valid_count", ctypes.c_size_t, [_P]),
    ("hc_level_stage_count", ctypes.c_size_t, [_P]),
    ("hc_level_inner_count", ctypes.c_size_t, [_P]),
    ("hc_level_face_count", ctypes.c_size_t, [_P, _I]),
    ("hc_prepare_stage", _I, [_P, _P, _P, _D, _P, _P, _P]),
    ("hc_stage_rhs", _I, [_P, _P, _P, _P, _D, _P, _P, _P, _P]),
    ("hc_rk4_advance", _I, [_P, _P, _P, _D, _I, _P]),
    ("hc_level_use_graph", _I, [_P, _I]),
    ("hc_level_set_cons_sum", _I, [_P, _P]),
    ("hc_level_dt_cfl", _I, [_P, _P, _D, _P, _P]),
    ("hc_level_set_tile", _I, [_P, _I]),
    ("hc_level_set_tile_shape", _I, [_P, _I, _I, _I]),
    ("hc_level_tile", _I, [_P]),
    ("hc_level_tile_shape", _I, [_P, _P]),
    ("hc_level_autotune", _I, [_P, _P, _P, _P, _P, _P]),
    ("hc_level_valid_state", _I, [_P, _P, _P, _P, _P]),
    ("hc_level_inner_ghosts_from_bc", _I, [_P, _P, _P, _P]),
    ("hc_level_inner_ghosts_outflow", _I, [_P, _D, _D, _P, _P]),
    ("hc_bc_frame_at", _I, [_P, _I, _I, _I, _D, _I, _D, _P, _P]),
    ("hc_level_bc_points", _I, [_P, _P, _I, _I, _P, _P, _P, _P]),
    ("hc_level_bc_averages", _I, [_P, _P, _I, _I, _P, _P, _P, _P]),
    ("hc_amr_interp4", _I, [_P, _P, _I, _P, _P, _P, _P]),
    ("hc_amr_ghosts2", _I, [_P, _P, _P, _P, _P, _D, _P, _P]),
    ("hc_amr_average_down", _I, [_P, _P, _I, _P, _P, _P, _P]),
    ("hc_amr_prolong4", _I, [_P, _P, _P, _P, _P, _P, _D, _P, _P, _P]),
    ("hc_amr_prolong4_noalloc", _I, [_P, _P, _P, _P, _P, _P, _D, _P, _P, _P]),
    ("hc_init_state", _I, [ctypes.POINTER(HcShell), ctypes.POINTER(HcOptions), _I, ctypes.c_ulonglong, _P, _P]),
    ("hc_geom_cell_detA", _D, [_I, _I, _I]),
    ("hc_geom_face_normal", None, [_I, _I, _I, _I, _I, _P]),
    ("hc_geom_ghost_donors", _I, [_I, _I, _I, _I, _P, _P, _P, _P]),
    ("hc_boxop_create", _I, [ctypes.POINTER(HcShell), ctypes.POINTER(HcOptions), ctypes.POINTER(HcBox), _I,
                             ctypes.POINTER(_P)]),
    ("hc_boxop_create_levels", _I, [ctypes.POINTER(HcShell), _I, ctypes.POINTER(HcOptions), ctypes.POINTER(HcBox),
                                    ctypes.POINTER(ctypes.c_int), _I, ctypes.POINTER(_P)]),
    ("hc_boxop_destroy", None, [_P]),
    ("hc_boxop_count", ctypes.c_longlong, [_P, _I]),
    ("hc_boxop2_create", _I, [ctypes.POINTER(HcShell), ctypes.POINTER(HcOptions), ctypes.POINTER(HcBox), _I,
                              ctypes.POINTER(_P)]),
    ("hc_boxop2_destroy", None, [_P]),
    ("hc_boxop2_count", ctypes.c_longlong, [_P, _I]),
    ("hc_boxop2_apply", _I, [_P, _P, _P, _P, _P, _P, ctypes.c_double, _P]),
    ("hc_boxop_apply", _I, [_P, _P, _P, _P, _P, _P, _P]),
    ("hc_boxop_apply_scaled", _I, [_P, _P, _P, _P, _P, _P, ctypes.c_double, _P]),
    ("hc_init_native_boxes", _I, [ctypes.POINTER(HcShell), ctypes.POINTER(HcOptions), _I, ctypes.c_ulonglong,
                                  ctypes.POINTER(HcBox), _I, _I, _P]),
    ("hc_init_level4", _I, [ctypes.POINTER(HcShell), ctypes.POINTER(HcOptions), _I, ctypes.c_ulonglong, _I, _P, _P]),
    ("hc_init_ju_boxes", _I, [ctypes.POINTER(HcShell), ctypes.POINTER(HcOptions), _I, ctypes.c_ulonglong,
                              ctypes.POINTER(HcBox), _I, _I, _P]),
    ("hc_amr4_create", _I, [ctypes.POINTER(HcShell), _I, ctypes.POINTER(HcOptions), ctypes.POINTER(HcBox),
                            ctypes.POINTER(ctypes.c_int), _I, ctypes.POINTER(_P)]),
    ("hc_amr4_destroy", None, [_P]),
    ("hc_amr4_count", ctypes.c_longlong, [_P, _I]),
    ("hc_amr4_set_tiles", _I, [_P, _P, _P]),
    ("hc_amr4_prepare", _I, [_P, _P, _P, _P, _P]),
    ("hc_amr4_stage_rhs", _I, [_P, _P, _P, _P, _P]),
    ("hc_amr4_rk4_advance", _I, [_P, _P, _P, _D, _I, _P]),
    ("hc_level_profile", _I, [_P, _I]),
    ("hc_level_profile_read", _I, [_P, _I, ctypes.c_char_p, _I, ctypes.POINTER(_D), ctypes.POINTER(ctypes.c_longlong)]),
    ("hc_last_error", ctypes.c_char_p, []),
]
EXPORTED_SYMBOLS = [s[0] for s in _SIGNATURES]

_lib = None


def lib() -> ctypes.CDLL:
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(the HIP path has no CPU fallback)")
        handle = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGNATURES:
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


class HcError(RuntimeError):
    pass


def _check(rc: int) -> None:
    if rc != 0:
        raise HcError(lib().hc_last_error().decode())


@dataclass
class Shell:
    """Shell geometry (ParseInputs inner_radius/outer_radius/radial_stretch)."""

    n_rad: int
    n_ang: int
    inner_radius: float = 21.5
    outer_radius: float = 215.0
    radial_stretch: float = 0.0

    def c(self) -> HcShell:
        return HcShell(self.n_rad, self.n_ang, self.inner_radius, self.outer_radius, self.radial_stretch)


@dataclass
class Options:
    """Hot-path option subset of ParseInputs (src/Inputs_Parsing.H:71-123)."""

    spatial_order: int = 2
    limiter: int = 1
    riemann_solver: int = 1
    entropy_fix: float = 0.3
    gamma: float = 5.0 / 3.0
    precision: int = 0  # 0 fp64, 1 mixed (fp32 slopes / reconstruction weights and Riemann dissipation)
    spatial_operator: int = 1  # fourth order: 1 standard, 2 Berta WENOZ
    berta_eta_d: float = 0.02  # Berta Jameson detector threshold
    berta_fallback: int = 0  # Berta fallback: 0 WENO3, 1 PLM
    divergence_cleaning: int = 0  # 1: Powell source terms

    def c(self) -> HcOptions:
        return HcOptions(self.spatial_order, self.limiter, self.riemann_solver, self.entropy_fix, self.gamma,
                         self.precision, self.spatial_operator, self.berta_eta_d, self.berta_fallback,
                         self.divergence_cleaning)


def init_state(shell: Shell, options: Options, config: int = 0, seed: int = 20261203):
    """Seeded initial J*U level and inner ghost primitives as numpy arrays."""
    import numpy as np

    g = 2
    na, nr = shell.n_ang, shell.n_rad
    ju = np.zeros((6, 8, na, na, nr))
    inner = np.zeros((6, 8, na + 2 * g, na + 2 * g, g))
    sh, op = shell.c(), options.c()
    _check(lib().hc_init_state(ctypes.byref(sh), ctypes.byref(op), config, seed,
                               ju.ctypes.data, inner.ctypes.data))
    return ju, inner


def init_level4(shell: Shell, options: Options, config: int = 3, seed: int = 20261203, quad: int = 3):
    """<J U> of the valid level and native inner-ghost averages of a fourth-order level (numpy)."""
    import numpy as np

    g = 5
    na, nr = shell.n_ang, shell.n_rad
    ju = np.zeros((6, 8, na, na, nr))
    inner = np.zeros((6, 8, na + 2 * g, na + 2 * g, g))
    sh, op = shell.c(), options.c()
    _check(lib().hc_init_level4(ctypes.byref(sh), ctypes.byref(op), config, seed, quad,
                                ju.ctypes.data, inner.ctypes.data))
    return ju, inner


def _ptr(t) -> int | None:
    return None if t is None else t.data_ptr()


class Level:
    """A six-panel shell level on the current HIP device."""

    def __init__(self, shell: Shell, options: Options):
        import torch

        self.shell = shell
        self.options = options
        self._handle = ctypes.c_void_p()
        sh, op = shell.c(), options.c()
        _check(lib().hc_level_create(ctypes.byref(sh), ctypes.byref(op), ctypes.byref(self._handle)))
        self.device = torch.device("cuda", torch.cuda.current_device())

    def __del__(self):
        if getattr(self, "_handle", None) and self._handle.value and _lib is not None:
            _lib.hc_level_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def ghost(self) -> int:
        return lib().hc_level_ghost(self._handle)

    @property
    def valid_shape(self):
        s = self.shell
        return (6, 8, s.n_ang, s.n_ang, s.n_rad)

    @property
    def stage_shape(self):
        s, g = self.shell, self.ghost()
        return (6, 8, s.n_ang + 2 * g, s.n_ang + 2 * g, s.n_rad + 2 * g)

    def face_shape(self, d: int):
        s = self.shell
        return (6, 8, s.n_ang + (d == 2), s.n_ang + (d == 1), s.n_rad + (d == 0))

    @staticmethod
    def _stream():
        import torch

        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    @staticmethod
    def _require(t, shape, name):
        import torch

        if t is None:
            return
        if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous() or tuple(t.shape) != tuple(shape):
            raise ValueError(f"{name}: need a contiguous float64 device tensor of shape {tuple(shape)}")

    def prepare_stage(self, U0, inner, W, K=None, a_dt: float = 0.0):
        """preStageLevel: W <- ghosted Cartesian primitives of U0 + a_dt*K."""
        self._require(U0, self.valid_shape, "U0")
        self._require(K, self.valid_shape, "K")
        self._require(W, self.stage_shape, "W")
        self._require(inner, self.inner_shape, "inner")
        if inner is None:
            raise ValueError("inner: the inner radial ghosts are required")
        _check(lib().hc_prepare_stage(self._handle, _ptr(U0), _ptr(K), a_dt, _ptr(inner), _ptr(W), self._stream()))

    def stage_rhs(self, W, K, U=None, b_dt: float = 0.0, fluxes=None):
        """StageLevel: K <- d<JU>/dt; optionally U += b_dt*K and the face fluxes."""
        self._require(W, self.stage_shape, "W")
        self._require(K, self.valid_shape, "K")
        self._require(U, self.valid_shape, "U")
        f = list(fluxes) if fluxes is not None else [None, None, None]
        for d in range(3):
            self._require(f[d], self.face_shape(d), f"F{d}")
        _check(lib().hc_stage_rhs(self._handle, _ptr(W), _ptr(K), _ptr(U), b_dt,
                                  _ptr(f[0]), _ptr(f[1]), _ptr(f[2]), self._stream()))

    def profile(self, enable: bool = True):
        """Turn per-kernel HIP-event timing on (clearing old records) or off."""
        _check(lib().hc_level_profile(self._handle, int(enable)))

    def profile_read(self):
        """{kernel name: (total ms, launches)} of the recorded launches."""
        out = {}
        slot = 0
        name = ctypes.create_string_buffer(128)
        while True:
            ms, n = ctypes.c_double(), ctypes.c_longlong()
            rc = lib().hc_level_profile_read(self._handle, slot, name, 128, ctypes.byref(ms), ctypes.byref(n))
            if rc == 1:
                return out
            _check(rc)
            out[name.value.decode()] = (ms.value, n.value)
            slot += 1

    @property
    def inner_shape(self):
        s, g = self.shell, self.ghost()
        return (6, 8, s.n_ang + 2 * g, s.n_ang + 2 * g, g)

    def rk4_advance(self, U, inner, dt: float, nsteps: int = 1):
        """MBLevelRK4::advance: nsteps classical RK4 steps in place on U
        (replayed from a captured hipGraph of one step unless use_graph(False))."""
        self._require(U, self.valid_shape, "U")
        self._require(inner, self.inner_shape, "inner")
        if inner is None:
            raise ValueError("inner: the inner radial ghosts are required")
        _check(lib().hc_rk4_advance(self._handle, _ptr(U), _ptr(inner), dt, nsteps, self._stream()))

    def use_graph(self, enable: bool = True):
        _check(lib().hc_level_use_graph(self._handle, int(bool(enable))))

    def set_cons_sum(self, cons):
        """Conservation sums (radialCons): `cons` a float64 device tensor of 8
        elements that every stage RHS adds b_dt x the net radial boundary
        inflow to (fourth order); None turns it off. Keep the tensor alive
        while it is set."""
        import torch

        if cons is not None and (cons.dtype != torch.float64 or not cons.is_cuda or cons.numel() != 8
                                 or not cons.is_contiguous()):
            raise ValueError("cons: need a contiguous float64 device tensor of 8 elements")
        self._cons = cons
        _check(lib().hc_level_set_cons_sum(self._handle, _ptr(cons)))

    def set_tile(self, tile):
        """Fourth-order box-table tiles (results unchanged): an edge for cubes,
        or a (radial, alpha, beta) shape."""
        if isinstance(tile, (tuple, list)):
            _check(lib().hc_level_set_tile_shape(self._handle, *[int(t) for t in tile]))
        else:
            _check(lib().hc_level_set_tile(self._handle, int(tile)))

    @property
    def tile(self) -> int:
        """Edge of cubic tiles (-1 for another shape)."""
        return lib().hc_level_tile(self._handle)

    @property
    def tile_shape(self):
        t = (ctypes.c_int * 3)()
        _check(lib().hc_level_tile_shape(self._handle, t))
        return tuple(t)

    def autotune(self, U, inner):
        """Times one stage RHS of U for the candidate tile shapes (cubes and
        radial pencils) and keeps the fastest; returns ((t0, t1, t2), ms per
        stage). U is not modified."""
        self._require(U, self.valid_shape, "U")
        self._require(inner, self.inner_shape, "inner")
        if inner is None:
            raise ValueError("inner: the inner radial ghosts are required")
        t, ms = (ctypes.c_int * 3)(), ctypes.c_double()
        _check(lib().hc_level_autotune(self._handle, _ptr(U), _ptr(inner), t, ctypes.byref(ms), self._stream()))
        return tuple(t), ms.value

    def dt_cfl(self, U, cone_boundary_density: float = 0.0) -> float:
        """1 / max CFL speed sum of the valid state U (BoxOp dtCFL); the caller
        applies its CFL number."""
        self._require(U, self.valid_shape, "U")
        out = ctypes.c_double()
        _check(lib().hc_level_dt_cfl(self._handle, _ptr(U), float(cone_boundary_density), ctypes.byref(out),
                                     self._stream()))
        return out.value

    def inner_ghosts_from_bc(self, bc, inner):
        """Inner radial ghosts (the rk4_advance / prepare_stage `inner` array)
        from one radial strip of boundary-file primitives bc[6][8][na+2g][na+2g]."""
        g = self.ghost()
        n = self.shell.n_ang + 2 * g
        self._require(bc, (6, 8, n, n), "bc")
        self._require(inner, (6, 8, n, n, g), "inner")
        _check(lib().hc_level_inner_ghosts_from_bc(self._handle, _ptr(bc), _ptr(inner), self._stream()))

    def bc_points(self, data0, theta, phi, bc):
        """fillPoints: the boundary map data0[8][nphi][ntheta] (hc_bc_frame_at)
        at the inner strip's cell centres, bc[6][8][na+2g][na+2g]."""
        g = self.ghost()
        n = self.shell.n_ang + 2 * g
        nt, npf = theta.numel(), phi.numel()
        self._require(data0, (8, npf, nt), "data0")
        self._require(bc, (6, 8, n, n), "bc")
        _check(lib().hc_level_bc_points(self._handle, _ptr(data0), nt, npf, _ptr(theta), _ptr(phi), _ptr(bc),
                                        self._stream()))

    def bc_averages(self, data0, theta, phi, inner):
        """fill (BC_conserved_averages 1): native conserved averages of the
        inner ghosts inner[6][8][na+2g][na+2g][g] from the map data0."""
        g = self.ghost()
        n = self.shell.n_ang + 2 * g
        nt, npf = theta.numel(), phi.numel()
        self._require(data0, (8, npf, nt), "data0")
        self._require(inner, (6, 8, n, n, g), "inner")
        _check(lib().hc_level_bc_averages(self._handle, _ptr(data0), nt, npf, _ptr(theta), _ptr(phi), _ptr(inner),
                                          self._stream()))

    def inner_ghosts_outflow(self, inner, strength: float, time: float):
        """Inner radial ghosts of the radial-outflow test boundary at `time` (s)."""
        g = self.ghost()
        n = self.shell.n_ang + 2 * g
        self._require(inner, (6, 8, n, n, g), "inner")
        _check(lib().hc_level_inner_ghosts_outflow(self._handle, float(strength), float(time), _ptr(inner),
                                                   self._stream()))

    def valid_state(self, U):
        """(valid, first_invalid): every value finite and every density > 0
        (BoxOp validConservedState); first_invalid is the flat index of the
        first offending value in the valid layout, or -1."""
        self._require(U, self.valid_shape, "U")
        ok, first = ctypes.c_int(), ctypes.c_longlong()
        _check(lib().hc_level_valid_state(self._handle, _ptr(U), ctypes.byref(ok), ctypes.byref(first),
                                          self._stream()))
        return bool(ok.value), first.value


def _patch(lo, hi):
    b = HcBox()
    b.block = 0
    for d in range(3):
        b.lo[d] = int(lo[d])
        b.hi[d] = int(hi[d])
    return b


def _patch_shape(lo, hi, nc):
    return (nc, hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)


def _check_patch(t, shape, name):
    import torch

    if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous() or tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: need a contiguous float64 device tensor of shape {tuple(shape)}")


def amr_interp4(coarse, coarse_lo, coarse_hi, ratio, fine, fine_lo, fine_hi):
    """Fourth-order fine averages fine[nc][k][j][i] (fine_lo..fine_hi) from the
    coarse averages coarse[nc][k][j][i] (coarse_lo..coarse_hi), refinement
    ratio per direction; cells whose stencil leaves the coarse patch are NaN."""
    import torch

    nc = coarse.shape[0]
    _check_patch(coarse, _patch_shape(coarse_lo, coarse_hi, nc), "coarse")
    _check_patch(fine, _patch_shape(fine_lo, fine_hi, nc), "fine")
    cb, fb = _patch(coarse_lo, coarse_hi), _patch(fine_lo, fine_hi)
    r = (ctypes.c_int * 3)(*[int(x) for x in ratio])
    _check(lib().hc_amr_interp4(_ptr(coarse), ctypes.byref(cb), nc, r, ctypes.byref(fb), _ptr(fine),
                                ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))


def amr_ghosts2(coarse_ju_j, coarse_lo, coarse_hi, ratio, fine_w, fine_lo, fine_hi, owned_lo, owned_hi,
                gamma=5.0 / 3.0):
    """Second-order fine ghost cells (Cartesian primitives) of fine_w[8][..]
    outside owned_lo..owned_hi from coarse J*U and J (coarse_ju_j[9][..])."""
    import torch

    _check_patch(coarse_ju_j, _patch_shape(coarse_lo, coarse_hi, 9), "coarse_ju_j")
    _check_patch(fine_w, _patch_shape(fine_lo, fine_hi, 8), "fine_w")
    cb, fb, ob = _patch(coarse_lo, coarse_hi), _patch(fine_lo, fine_hi), _patch(owned_lo, owned_hi)
    r = (ctypes.c_int * 3)(*[int(x) for x in ratio])
    _check(lib().hc_amr_ghosts2(_ptr(coarse_ju_j), ctypes.byref(cb), r, ctypes.byref(fb), ctypes.byref(ob),
                                float(gamma), _ptr(fine_w), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))


def amr_average_down(fine, fine_lo, fine_hi, ratio, coarse, coarse_lo, coarse_hi):
    """coarse[nc][..] = mean of each coarse cell's children in fine[nc][..]."""
    import torch

    nc = fine.shape[0]
    _check_patch(fine, _patch_shape(fine_lo, fine_hi, nc), "fine")
    _check_patch(coarse, _patch_shape(coarse_lo, coarse_hi, nc), "coarse")
    fb, cb = _patch(fine_lo, fine_hi), _patch(coarse_lo, coarse_hi)
    r = (ctypes.c_int * 3)(*[int(x) for x in ratio])
    _check(lib().hc_amr_average_down(_ptr(fine), ctypes.byref(fb), nc, r, ctypes.byref(cb), _ptr(coarse),
                                     ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))


def amr_prolong4(interp, fine_j, fine_lo, fine_hi, coarse, coarse_lo, coarse_hi, ratio, fine_ju,
                 gamma=5.0 / 3.0):
    """Conservative fourth-order refinement (fine J*U in fine_ju[8][..]) from
    interp = amr_interp4 of the coarse J*U and J (9 components), the fine
    Jacobian fine_j[1][..] and the coarse J*U and J coarse[9][..]. Returns the
    number of invalid parents (0 on success; HcError is raised otherwise)."""
    import torch

    _check_patch(interp, _patch_shape(fine_lo, fine_hi, 9), "interp")
    _check_patch(fine_j, _patch_shape(fine_lo, fine_hi, 1), "fine_j")
    _check_patch(coarse, _patch_shape(coarse_lo, coarse_hi, 9), "coarse")
    _check_patch(fine_ju, _patch_shape(fine_lo, fine_hi, 8), "fine_ju")
    fb, cb = _patch(fine_lo, fine_hi), _patch(coarse_lo, coarse_hi)
    r = (ctypes.c_int * 3)(*[int(x) for x in ratio])
    bad = ctypes.c_int()
    _check(lib().hc_amr_prolong4(_ptr(interp), _ptr(fine_j), ctypes.byref(fb), _ptr(coarse), ctypes.byref(cb), r,
                                 float(gamma), _ptr(fine_ju), ctypes.byref(bad),
                                 ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return bad.value


def box_table(boxes):
    """ctypes array of HcBox from (block, lo, hi) triples."""
    arr = (HcBox * len(boxes))()
    for n, (blk, lo, hi) in enumerate(boxes):
        arr[n].block = blk
        for d in range(3):
            arr[n].lo[d] = lo[d]
            arr[n].hi[d] = hi[d]
    return arr


def shell_boxes(shell: Shell, size=(32, 32, 32)):
    """Uniform decomposition of all six panels into boxes of the given size."""
    out = []
    for p in range(6):
        for k in range(0, shell.n_ang, size[2]):
            for j in range(0, shell.n_ang, size[1]):
                for i in range(0, shell.n_rad, size[0]):
                    out.append((p, (i, j, k), (min(i + size[0], shell.n_rad) - 1,
                                               min(j + size[1], shell.n_ang) - 1,
                                               min(k + size[2], shell.n_ang) - 1)))
    return out


def init_native_boxes(shell: Shell, options: Options, boxes, config: int = 3, seed: int = 20261203, quad: int = 3):
    """Native conserved cell averages (ghost 5) of every box, concatenated, as numpy."""
    import numpy as np

    g = 5
    total = sum(8 * (hi[0] - lo[0] + 1 + 2 * g) * (hi[1] - lo[1] + 1 + 2 * g) * (hi[2] - lo[2] + 1 + 2 * g)
                for _, lo, hi in boxes)
    out = np.zeros(total)
    tab = box_table(boxes)
    sh, op = shell.c(), options.c()
    _check(lib().hc_init_native_boxes(ctypes.byref(sh), ctypes.byref(op), config, seed, tab, len(boxes), quad,
                                      out.ctypes.data))
    return out


def bc_frame_at(frames, cadence_hours: float, frame_rotate: int, time: float, data0):
    """BoxData_to_BC's frame step: frames [nframes][8][nphi][ntheta] (device)
    -> the rotated, time-interpolated map data0 [8][nphi][ntheta] at `time` (s)."""
    import torch

    if frames.dim() != 4 or frames.shape[1] != 8:
        raise ValueError("frames: need a [nframes][8][nphi][ntheta] tensor")
    nf, _, npf, nt = frames.shape
    for t, shape, name in ((frames, tuple(frames.shape), "frames"), (data0, (8, npf, nt), "data0")):
        if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous() or tuple(t.shape) != shape:
            raise ValueError(f"{name}: need a contiguous float64 device tensor of shape {shape}")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _check(lib().hc_bc_frame_at(_ptr(frames), nf, nt, npf, float(cadence_hours), int(frame_rotate), float(time),
                                _ptr(data0), stream))


def init_ju_boxes(shell: Shell, options: Options, boxes, config: int = 3, seed: int = 20261203, quad: int = 3):
    """<J U> of every box (valid cells, [8][n2][n1][n0] each), concatenated, as numpy."""
    import numpy as np

    total = sum(8 * (hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1) for _, lo, hi in boxes)
    out = np.zeros(total)
    tab = box_table(boxes)
    sh, op = shell.c(), options.c()
    _check(lib().hc_init_ju_boxes(ctypes.byref(sh), ctypes.byref(op), config, seed, tab, len(boxes), quad,
                                  out.ctypes.data))
    return out


class Amr4:
    """Fixed AMR hierarchy of fourth-order levels with the synchronous
    classical RK4 of CubedSphereAMR::advanceRungeKuttaSynchronous
    (src/MHD_AMRRungeKutta.H:5-41) on the current HIP device.

    shells: one Shell per level (level 0 = the whole shell); patches: the
    refined patches as (level, (block, lo, hi)). State tensors hold the
    level-0 valid level followed by every patch in table order."""

    def __init__(self, shells, options: Options, patches):
        self.options = options
        self.shells = list(shells)
        self.patches = [(int(l), (int(b), tuple(lo), tuple(hi))) for l, (b, lo, hi) in patches]
        self._handle = ctypes.c_void_p()
        op = options.c()
        boxes = [b for _, b in self.patches]
        tab = box_table(boxes) if boxes else None
        shs = (HcShell * len(self.shells))(*[s.c() for s in self.shells])
        lev = (ctypes.c_int * max(1, len(boxes)))(*[l for l, _ in self.patches])
        _check(lib().hc_amr4_create(shs, len(self.shells), ctypes.byref(op), tab, lev, len(boxes),
                                    ctypes.byref(self._handle)))

    def __del__(self):
        if getattr(self, "_handle", None) and self._handle.value and _lib is not None:
            _lib.hc_amr4_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    @property
    def state_count(self) -> int:
        return lib().hc_amr4_count(self._handle, 0)

    @property
    def stage_count(self) -> int:
        return lib().hc_amr4_count(self._handle, 1)

    @property
    def level0_count(self) -> int:
        return lib().hc_amr4_count(self._handle, 2)

    @property
    def tile_count(self) -> int:
        """Tiles over the refined levels (patches grouped into rectangular tiles)."""
        return lib().hc_amr4_count(self._handle, 3)

    def set_tiles(self, level0_tile=None, refined_max=None):
        """Level-0 tile shape (radial, alpha, beta) and the largest refined tile
        (cells per direction; the patch size gives one tile per patch). Results
        do not depend on either; synchronizes the device."""
        a = (ctypes.c_int * 3)(*level0_tile) if level0_tile is not None else None
        b = (ctypes.c_int * 3)(*refined_max) if refined_max is not None else None
        _check(lib().hc_amr4_set_tiles(self._handle, a, b))

    def _check_args(self, pairs):
        import torch

        for t, n, name in pairs:
            if t is None or t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous() or t.numel() != n:
                raise ValueError(f"{name}: need a contiguous float64 device tensor of {n} elements")

    def _inner_count(self):
        sh = self.shells[0]
        return 6 * 8 * (sh.n_ang + 10) ** 2 * 5

    def prepare(self, JU, inner, USph):
        import torch

        self._check_args([(JU, self.state_count, "JU"), (inner, self._inner_count(), "inner"),
                          (USph, self.stage_count, "USph")])
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _check(lib().hc_amr4_prepare(self._handle, _ptr(JU), _ptr(inner), _ptr(USph), stream))

    def stage_rhs(self, JU, inner, K):
        import torch

        self._check_args([(JU, self.state_count, "JU"), (inner, self._inner_count(), "inner"),
                          (K, self.state_count, "K")])
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _check(lib().hc_amr4_stage_rhs(self._handle, _ptr(JU), _ptr(inner), _ptr(K), stream))

    def rk4_advance(self, JU, inner, dt: float, nsteps: int = 1):
        import torch

        self._check_args([(JU, self.state_count, "JU"), (inner, self._inner_count(), "inner")])
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _check(lib().hc_amr4_rk4_advance(self._handle, _ptr(JU), _ptr(inner), float(dt), int(nsteps), stream))


class BoxOp:
    """Fourth-order BoxOp (BoxOp_EulerCubedSphere::operator()) over a box table.

    apply(USph, rhs) evaluates every box in the same launches; rhs is
    accumulated into, as the reference's a_Rhs (the caller zeroes it).
    """

    def __init__(self, shell, options: Options, boxes, levels=None):
        """shell: a Shell, or a list of Shells (one per refinement level) with
        levels[n] the level of box n (hc_boxop_create_levels)."""
        self.options, self.boxes = options, list(boxes)
        self.shells = list(shell) if isinstance(shell, (list, tuple)) else [shell]
        self.shell = self.shells[0]
        self.levels = list(levels) if levels is not None else [0] * len(self.boxes)
        self._handle = ctypes.c_void_p()
        op = options.c()
        tab = box_table(self.boxes)
        shells = (HcShell * len(self.shells))(*[s.c() for s in self.shells])
        lev = (ctypes.c_int * len(self.boxes))(*self.levels)
        _check(lib().hc_boxop_create_levels(shells, len(self.shells), ctypes.byref(op), tab, lev, len(self.boxes),
                                            ctypes.byref(self._handle)))

    def __del__(self):
        if getattr(self, "_handle", None) and self._handle.value and _lib is not None:
            _lib.hc_boxop_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def count(self, what: int) -> int:
        return lib().hc_boxop_count(self._handle, what)

    def apply(self, USph, rhs, fluxes=None, scale=1.0):
        """rhs += scale * (mapped flux divergence) of every box; face fluxes
        into `fluxes` (three tensors or None entries)."""
        import torch

        f = list(fluxes) if fluxes is not None else [None, None, None]
        if len(f) != 3:
            raise ValueError("fluxes: need three face-flux tensors (or None entries)")
        checks = [(USph, 0, "USph"), (rhs, 1, "rhs")] + [(f[d], 2 + d, f"F{d}") for d in range(3) if f[d] is not None]
        for t, what, name in checks:
            if t is None or t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous() \
                    or t.numel() != self.count(what):
                raise ValueError(f"{name}: need a contiguous float64 device tensor of {self.count(what)} elements")
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _check(lib().hc_boxop_apply_scaled(self._handle, _ptr(USph), _ptr(rhs), _ptr(f[0]), _ptr(f[1]), _ptr(f[2]),
                                           float(scale), stream))


class BoxOp2:
    """Second-order per-box operator (secondOrderOperator with Cartesian
    primitives) over a box table. apply(W, rhs) evaluates every box in the
    same launches from the boxes' Cartesian primitives on the box grown by 2;
    rhs is accumulated into (rhs += scale * (-div F), plus the Powell source)."""

    def __init__(self, shell, options: Options, boxes):
        self.shell, self.options, self.boxes = shell, options, list(boxes)
        self._handle = ctypes.c_void_p()
        sh, op = shell.c(), options.c()
        _check(lib().hc_boxop2_create(ctypes.byref(sh), ctypes.byref(op), box_table(self.boxes), len(self.boxes),
                                      ctypes.byref(self._handle)))

    def __del__(self):
        if getattr(self, "_handle", None) and self._handle.value and _lib is not None:
            _lib.hc_boxop2_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def count(self, what: int) -> int:
        return lib().hc_boxop2_count(self._handle, what)

    def apply(self, W, rhs, fluxes=None, scale=1.0):
        import torch

        f = list(fluxes) if fluxes is not None else [None, None, None]
        if len(f) != 3:
            raise ValueError("fluxes: need three face-flux tensors (or None entries)")
        checks = [(W, 0, "W"), (rhs, 1, "rhs")] + [(f[d], 2 + d, f"F{d}") for d in range(3) if f[d] is not None]
        for t, what, name in checks:
            if t is None or t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous() \
                    or t.numel() != self.count(what):
                raise ValueError(f"{name}: need a contiguous float64 device tensor of {self.count(what)} elements")
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _check(lib().hc_boxop2_apply(self._handle, _ptr(W), _ptr(rhs), _ptr(f[0]), _ptr(f[1]), _ptr(f[2]),
                                     float(scale), stream))
