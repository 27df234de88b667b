"""The operation-counting build of the oracle (oracle/hc_opcount.cpp: the same
source with a counting scalar in place of double) reproduces the oracle's
box RHS bit for bit, and its counts scale with the work: twice the radial
extent of a box costs about twice the operations (the ghost layers add a
fixed part), and the Berta operator costs less than the standard one on the
same smooth state."""
import numpy as np

import heliocubed_amd as hc
import oracle as orc


def _count(op, box, nr=32, na=16):
    shell = hc.Shell(nr, na)
    opts = hc.Options(spatial_order=4, limiter=1, riemann_solver=2, spatial_operator=op)
    U = hc.init_native_boxes(shell, opts, [box], config=3, seed=5, quad=1)
    rhs, cnt = orc.opcount_box(nr, na, box, U, spatial_operator=op)
    ref = orc.o4_rhs_boxes(nr, na, [box], U, spatial_operator=op)
    assert np.array_equal(rhs, ref)
    return cnt


def test_counting_build_reproduces_the_oracle_and_scales():
    a = _count(1, (2, (8, 0, 4), (15, 7, 11)))
    b = _count(1, (2, (8, 0, 4), (23, 7, 11)))
    fa = a["add"] + a["mul"] + a["div"] + a["sqrt"]
    fb = b["add"] + b["mul"] + b["div"] + b["sqrt"]
    assert a["div"] > 0 and a["sqrt"] > 0 and a["cmp"] > 0
    # 8x8x8 -> 16x8x8 valid cells: the ghosted volume grows from 18^3 to 26*18*18
    assert 1.2 < fb / fa < 2.0, fb / fa
    berta = _count(2, (2, (8, 0, 4), (15, 7, 11)))
    assert berta["add"] + berta["mul"] + berta["div"] + berta["sqrt"] < fa
