"""GPU parity of the AMR coarse-to-fine transfers with the oracle:
hc_amr_interp4 (amrInterpolateFourthOrder, src/MHD_AMRInterpolation.H:39-161)
and hc_amr_ghosts2 (second-order fine ghosts, src/MHD_AMR.H:1282-1393), on
patches with negative ghost indices, mixed ratios, missing donors (NaN) and
positivity fallbacks; relative 1e-12, NaN cells identical."""
import numpy as np
import pytest
import torch

import heliocubed_amd as hc
import oracle as orc
from test_oracle_amr_pins import _coarse_state, FLO, FHI, OLO, OHI, R2


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.gpu
@pytest.mark.parametrize("ratio", [(2, 2, 2), (2, 1, 2), (4, 4, 4), (3, 2, 1), (1, 1, 1)])
def test_interp4_matches_oracle(device, ratio):
    rng = np.random.default_rng(sum(ratio) * 7)
    clo, chi = (-3, -2, 0), (6, 5, 7)
    coarse = rng.random((8,) + tuple(chi[d] - clo[d] + 1 for d in (2, 1, 0))) + 0.5
    coarse[3, 2, 4, 5] = np.nan                   # a NaN donor propagates
    flo = tuple(ratio[d] * clo[d] - 1 for d in range(3))   # reaches past the donors: NaN edges
    fhi = tuple(ratio[d] * chi[d] + ratio[d] for d in range(3))
    ref = orc.amr_interp4(coarse, clo, chi, ratio, flo, fhi)
    fine = torch.full(ref.shape, 12345.0, dtype=torch.float64, device=device)
    hc.amr_interp4(_dev(coarse), clo, chi, ratio, fine, flo, fhi)
    torch.cuda.synchronize()
    got = fine.cpu().numpy()
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    ok = ~np.isnan(ref)
    assert ok.any()
    err = np.abs(got[ok] - ref[ok]).max() / np.abs(ref[ok]).max()
    print("interp4", ratio, "nan cells", int((~ok).sum()), "err", err)
    assert err <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["smooth", "hole", "nan_parent"])
def test_ghosts2_match_oracle(device, case):
    coarse, _, clo, chi = _coarse_state(seed=11)
    idx = lambda c: (c[2] - clo[2], c[1] - clo[1], c[0] - clo[0])
    if case == "hole":
        coarse[(0,) + idx((0, 2, 2))] = -20.0 * coarse[8, 0, 0, 0]
        coarse[(4,) + idx((5, 4, 1))] = -50.0
    if case == "nan_parent":
        coarse[(8,) + idx((4, 2, 2))] = np.nan
    init = np.full((8, 12, 12, 12), -7.0)
    ref = orc.amr_ghosts2(coarse, clo, chi, R2, FLO, FHI, OLO, OHI, init)
    fine = _dev(init)
    hc.amr_ghosts2(_dev(coarse), clo, chi, R2, fine, FLO, FHI, OLO, OHI)
    torch.cuda.synchronize()
    got = fine.cpu().numpy()
    assert np.array_equal(got == -7.0, ref == -7.0)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("ghosts2", case, "err", err)
    assert np.isfinite(got).all() and err <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("ratio", [(2, 2, 2), (2, 3, 1), (4, 2, 2)])
def test_average_down_matches_oracle(device, ratio):
    rng = np.random.default_rng(sum(ratio))
    clo, chi = (-2, 1, 0), (3, 4, 5)
    flo = tuple(ratio[d] * clo[d] - 1 for d in range(3))
    fhi = tuple(ratio[d] * chi[d] + ratio[d] for d in range(3))
    fine = rng.random((8,) + tuple(fhi[d] - flo[d] + 1 for d in (2, 1, 0)))
    ref = orc.amr_average_down(fine, flo, fhi, ratio, clo, chi)
    out = torch.zeros(ref.shape, dtype=torch.float64, device=device)
    hc.amr_average_down(_dev(fine), flo, fhi, ratio, out, clo, chi)
    torch.cuda.synchronize()
    err = np.abs(out.cpu().numpy() - ref).max() / np.abs(ref).max()
    print("average_down", ratio, err)
    assert err <= 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("jump", [False, True])
def test_prolong4_matches_oracle(device, jump):
    from test_oracle_amr_pins import _prolong_case
    coarse, interp, fine_j, _, clo, chi, flo, fhi, r, J = _prolong_case(23, jump=jump)
    ref, bad = orc.amr_prolong4(interp, fine_j, flo, fhi, coarse, clo, chi, r)
    assert bad == 0
    # the device chain: interpolation then conservative refinement
    it = torch.zeros((9, 8, 8, 8), dtype=torch.float64, device=device)
    hc.amr_interp4(_dev(coarse), clo, chi, r, it, flo, fhi)
    out = torch.zeros((8, 8, 8, 8), dtype=torch.float64, device=device)
    nbad = hc.amr_prolong4(it, _dev(fine_j), flo, fhi, _dev(coarse), clo, chi, r, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("prolong4", jump, nbad, err)
    assert nbad == 0 and err <= 1e-12


@pytest.mark.gpu
def test_prolong4_rejects_inadmissible_parent(device):
    from test_oracle_amr_pins import _prolong_case
    coarse, interp, fine_j, _, clo, chi, flo, fhi, r, J = _prolong_case(29)
    coarse[0, 3, 3, 3] = -1.0
    it = torch.zeros((9, 8, 8, 8), dtype=torch.float64, device=device)
    hc.amr_interp4(_dev(coarse), clo, chi, r, it, flo, fhi)
    out = torch.zeros((8, 8, 8, 8), dtype=torch.float64, device=device)
    with pytest.raises(hc.HcError, match="invalid"):
        hc.amr_prolong4(it, _dev(fine_j), flo, fhi, _dev(coarse), clo, chi, r, out)
