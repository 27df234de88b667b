"""The allocation-free conservative refinement hc_amr_prolong4_noalloc (its
per-parent flags live in a static device array) against hc_amr_prolong4 on
the same inputs: identical refined values and invalid-parent counts on two
calls in a row, the second with fewer parents, and on an inadmissible parent."""
import ctypes

import numpy as np
import pytest
import torch

import heliocubed_amd as hc
from test_oracle_amr_pins import _prolong_case


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _box(lo, hi):
    b = hc.HcBox()
    for d in range(3):
        b.lo[d], b.hi[d] = int(lo[d]), int(hi[d])
    return b


def _call(name, it, fj, flo, fhi, co, clo, chi, r, out):
    fn = getattr(hc.lib(), name)   # signatures set by heliocubed_amd.lib()
    bad = ctypes.c_int(-1)
    fb, cb = _box(flo, fhi), _box(clo, chi)
    ra = (ctypes.c_int * 3)(*r)
    rc = fn(ctypes.c_void_p(it.data_ptr()), ctypes.c_void_p(fj.data_ptr()), ctypes.byref(fb),
            ctypes.c_void_p(co.data_ptr()), ctypes.byref(cb), ra, ctypes.c_double(5.0 / 3.0),
            ctypes.c_void_p(out.data_ptr()), ctypes.byref(bad), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return rc, bad.value


def _inputs(seed, jump, flo, fhi, device):
    coarse, interp, fine_j, _, clo, chi, _, _, r, _ = _prolong_case(seed, jump=jump)
    # the case's interpolation covers the fine cells 0..7: take the requested sub-box of it
    sl = tuple(slice(flo[d], fhi[d] + 1) for d in (2, 1, 0))
    it = _dev(interp[(slice(None),) + sl])
    fj = _dev(fine_j[(slice(None),) + sl])
    return it, fj, _dev(coarse), clo, chi, r


@pytest.mark.gpu
def test_noalloc_refinement_matches_the_allocating_entry(device):
    print("noalloc refinement: two calls, 4^3 then 2x1x2 parents", flush=True)
    for seed, jump, flo, fhi in ((31, True, (0, 0, 0), (7, 7, 7)), (37, False, (2, 4, 0), (5, 5, 3))):
        it, fj, co, clo, chi, r = _inputs(seed, jump, flo, fhi, device)
        shape = (8,) + tuple(fhi[d] - flo[d] + 1 for d in (2, 1, 0))
        a = torch.full(shape, 7.0, dtype=torch.float64, device=device)
        b = torch.full(shape, 7.0, dtype=torch.float64, device=device)
        rca, bada = _call("hc_amr_prolong4", it, fj, flo, fhi, co, clo, chi, r, a)
        rcb, badb = _call("hc_amr_prolong4_noalloc", it, fj, flo, fhi, co, clo, chi, r, b)
        print("prolong4", flo, fhi, "rc", rca, rcb, "invalid", bada, badb, flush=True)
        assert rca == rcb == 0 and bada == badb == 0
        assert torch.equal(a, b)


@pytest.mark.gpu
def test_noalloc_refinement_reports_an_inadmissible_parent(device):
    print("noalloc refinement: inadmissible parent", flush=True)
    coarse, interp, fine_j, _, clo, chi, flo, fhi, r, _ = _prolong_case(29)
    coarse[0, 3, 3, 3] = -1.0
    it, fj, co = _dev(interp), _dev(fine_j), _dev(coarse)
    a = torch.zeros((8, 8, 8, 8), dtype=torch.float64, device=device)
    b = torch.zeros_like(a)
    rca, bada = _call("hc_amr_prolong4", it, fj, flo, fhi, co, clo, chi, r, a)
    rcb, badb = _call("hc_amr_prolong4_noalloc", it, fj, flo, fhi, co, clo, chi, r, b)
    assert rca != 0 and rcb != 0 and bada == badb == 1
    assert b"invalid" in hc.lib().hc_last_error()
