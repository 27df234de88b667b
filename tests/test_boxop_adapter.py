"""The C++ adapter with the reference's per-box operator signature
(include/heliocubed_boxop.hpp; BoxOp_EulerCubedSphere::operator() and
BertaOperator, src/BoxOp_EulerCubedSphere.H:412-430).

CPU: the header compiles on its own with a plain C++17 compiler (no HIP
headers) and the host program that drives it is built. GPU: that program
evaluates one box through the adapter (a_Rhs accumulated with a_scale, the
three flux arrays written) and the result is compared with the CPU oracle.
"""
import os
import subprocess

import numpy as np
import pytest

import heliocubed_amd as hc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "heliocubed_amd", "hc_boxop_adapter_check")


def test_adapter_header_compiles_standalone(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text('#include "heliocubed_boxop.hpp"\n'
                   'static_assert(heliocubed::BoxOp_EulerCubedSphere::numState() == 8, "");\n'
                   'int main() { heliocubed::DeviceBoxData b; return (int)b.size(0); }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_adapter_program_is_built():
    assert os.access(CHECK, os.X_OK), "build with __graft_entry__.build()"


@pytest.mark.gpu
@pytest.mark.parametrize("operator,scale", [(1, 1.0), (1, 0.5), (2, 2.0)])
def test_adapter_matches_oracle(tmp_path, operator, scale):
    import oracle as orc
    from helpers import rel_linf_per_var

    n_rad, n_ang = 32, 16
    block, lo, hi = 3, (8, 0, 4), (15, 7, 11)
    opts = hc.Options(spatial_order=4, limiter=1, riemann_solver=2, spatial_operator=operator)
    cfg = 4 if operator == 2 else 3
    U = hc.init_native_boxes(hc.Shell(n_rad, n_ang), opts, [(block, lo, hi)], config=cfg, seed=11, quad=2)
    fin, fout = tmp_path / "u.bin", tmp_path / "out.bin"
    U.astype(np.float64).tofile(fin)
    args = [CHECK, "4", str(n_rad), str(n_ang), str(operator), "1", "2", "0", str(block), *map(str, lo), *map(str, hi),
            repr(scale), str(fin), str(fout)]
    print("adapter case:", operator, scale, flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(fout, dtype=np.float64)
    n = [hi[d] - lo[d] + 1 for d in range(3)]
    nr = 8 * int(np.prod(n))
    rhs, flux = orc.o4_rhs_boxes(n_rad, n_ang, [(block, lo, hi)], U, spatial_operator=operator, with_fluxes=True)
    err = rel_linf_per_var(got[:nr].reshape(8, -1), (scale * rhs).reshape(8, -1), axis_var=0)
    assert err.max() <= 1e-12, err
    off = nr
    for d in range(3):
        m = 8 * int(np.prod([n[q] + (q == d) for q in range(3)]))
        ef = rel_linf_per_var(got[off:off + m].reshape(8, -1), np.asarray(flux[d]).reshape(8, -1), axis_var=0)
        assert ef.max() <= 1e-12, (d, ef)
        off += m


@pytest.mark.gpu
@pytest.mark.parametrize("limiter,riemann,scale", [(1, 1, 1.0), (3, 2, 0.5)])
def test_adapter_second_order_matches_oracle(tmp_path, limiter, riemann, scale):
    """The adapter at spatial_order 2 with a_cartesianPrimitives: the box's
    Cartesian primitives (ghost 2) cut out of a prepared stage level."""
    import oracle as orc
    from helpers import rel_linf_per_var

    n_rad, n_ang = 12, 10
    block, lo, hi = 4, (6, 3, 1), (11, 8, 7)
    opts = hc.Options(limiter=limiter, riemann_solver=riemann)
    ju, inner = hc.init_state(hc.Shell(n_rad, n_ang), opts, 0, 7)
    W = orc.Oracle(n_rad, n_ang, limiter=limiter, riemann=riemann).prepare(ju, inner)
    g = 2
    Wb = W[block, :, lo[2]:hi[2] + 1 + 2 * g, lo[1]:hi[1] + 1 + 2 * g, lo[0]:hi[0] + 1 + 2 * g].copy()
    fin, fout = tmp_path / "w.bin", tmp_path / "out.bin"
    Wb.astype(np.float64).tofile(fin)
    args = [CHECK, "2", str(n_rad), str(n_ang), "1", str(limiter), str(riemann), "0", str(block), *map(str, lo),
            *map(str, hi), repr(scale), str(fin), str(fout)]
    print("adapter case: order 2", limiter, riemann, scale, flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(fout, dtype=np.float64)
    n = [hi[d] - lo[d] + 1 for d in range(3)]
    nr = 8 * int(np.prod(n))
    rhs, flux = orc.o2_rhs_boxes(n_rad, n_ang, [(block, lo, hi)], Wb.ravel(), limiter=limiter, riemann=riemann,
                                 scale=scale, with_fluxes=True)
    assert rel_linf_per_var(got[:nr].reshape(8, -1), rhs.reshape(8, -1), axis_var=0).max() <= 1e-12
    off = nr
    for d in range(3):
        m = 8 * int(np.prod([n[q] + (q == d) for q in range(3)]))
        ef = rel_linf_per_var(got[off:off + m].reshape(8, -1), flux[d].reshape(8, -1), axis_var=0)
        assert ef.max() <= 1e-12, (d, ef)
        off += m
