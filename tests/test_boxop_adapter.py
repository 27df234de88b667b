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
                   'static_assert(heliocubed::BoxOp_EulerCubedSphere::ghost() == 5, "");\n'
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
    args = [CHECK, str(n_rad), str(n_ang), str(operator), "1", "2", "0", str(block), *map(str, lo), *map(str, hi),
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
