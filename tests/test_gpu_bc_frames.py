"""GPU parity of the boundary-file path of MHDReader::BoxData_to_BC
(src/MHDReader.H:221-331: Carrington rotation and time blend of the frames)
and of MHD_BoundaryAverages (src/MHD_BoundaryAverages.H: fillPoints onto the
inner strip, fill of the native conserved averages of the inner ghosts) with
the oracle's restatement, within relative 1e-12 per component; the frame step
feeds the maps as BoxData_to_BC does."""
import numpy as np
import pytest
import torch

import heliocubed_amd as hc
import oracle as orc
from helpers import rel_linf_per_var

R1 = 60.0
CARR = 25.38 * 24 * 3600


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _map(nphi, ntheta, seed):
    """A smooth positive boundary map [8][nphi][ntheta] with its node angles."""
    rng = np.random.default_rng(seed)
    theta = np.linspace(0.05, np.pi - 0.05, ntheta)
    phi = np.linspace(0.0, 2 * np.pi, nphi, endpoint=False) + 0.1
    T, P = np.meshgrid(theta, phi, indexing="xy")
    base = np.array([300, 4e7, 2e6, -1e6, 10, 150, 20, -30.0])
    d = np.zeros((8, nphi, ntheta))
    for c in range(8):
        a = rng.uniform(0.1, 0.3, 3)
        d[c] = base[c] * (1 + a[0] * np.sin(T) * np.cos(P + a[1]) + a[2] * np.cos(2 * T))
    return d, theta, phi


@pytest.mark.gpu
@pytest.mark.parametrize("nphi,ntheta,rotate", [(360, 181, 0), (256, 64, 1), (90, 45, 0), (100, 33, 1)])
def test_frame_step_matches_oracle(device, nphi, ntheta, rotate):
    print("case bc frame", nphi, ntheta, rotate, flush=True)
    rng = np.random.default_rng(nphi + ntheta)
    frames = rng.standard_normal((3, 8, nphi, ntheta))
    fr = _dev(frames)
    out = torch.zeros((8, nphi, ntheta), dtype=torch.float64, device=device)
    cad = 1.5
    for t in (0.0, 0.37 * cad * 3600, 1.2 * cad * 3600, 2.0 * cad * 3600, 5.5 * cad * 3600, 0.31 * CARR):
        hc.bc_frame_at(fr, cad, rotate, t, out)
        torch.cuda.synchronize()
        ref = orc.bc_frame_at(frames, cad, rotate, t)
        err = np.abs(out.cpu().numpy() - ref).max() / np.abs(ref).max()
        assert err <= 1e-12, (t, err)


@pytest.mark.gpu
def test_frame_step_rejects_bad_arguments(device):
    fr = torch.zeros((2, 8, 720, 10), dtype=torch.float64, device=device)
    out = torch.zeros((8, 720, 10), dtype=torch.float64, device=device)
    with pytest.raises(RuntimeError, match="nphi"):
        hc.bc_frame_at(fr, 1.0, 0, 0.0, out)
    fr = torch.zeros((2, 8, 36, 10), dtype=torch.float64, device=device)
    out = torch.zeros((8, 36, 10), dtype=torch.float64, device=device)
    for cad, rot, t in ((0.0, 0, 1.0), (1.0, 2, 1.0), (1.0, 0, -1.0), (1.0, 0, float("inf"))):
        with pytest.raises(RuntimeError):
            hc.bc_frame_at(fr, cad, rot, t, out)


@pytest.mark.gpu
@pytest.mark.parametrize("order,nr,na,nphi,ntheta", [(2, 12, 8, 72, 37), (2, 16, 16, 180, 91),
                                                     (4, 16, 16, 128, 64), (4, 12, 20, 360, 181)])
def test_bc_maps_match_oracle(device, order, nr, na, nphi, ntheta):
    print("case bc maps", order, nr, na, nphi, ntheta, flush=True)
    shell = hc.Shell(nr, na, 21.5, R1)
    opts = hc.Options(spatial_order=order, limiter=1, riemann_solver=2)
    L = hc.Level(shell, opts)
    g = L.ghost()
    d0, theta, phi = _map(nphi, ntheta, order * 100 + na)
    frames = np.stack([d0, 1.1 * d0])
    data0 = torch.zeros((8, nphi, ntheta), dtype=torch.float64, device=device)
    hc.bc_frame_at(_dev(frames), 2.0, 1, 0.4 * 2.0 * 3600 + 0.05 * CARR, data0)
    th, ph = _dev(theta), _dev(phi)
    bc = torch.zeros((6, 8, na + 2 * g, na + 2 * g), dtype=torch.float64, device=device)
    inner = torch.zeros((6, 8, na + 2 * g, na + 2 * g, g), dtype=torch.float64, device=device)
    L.bc_points(data0, th, ph, bc)
    L.bc_averages(data0, th, ph, inner)
    torch.cuda.synchronize()
    host0 = orc.bc_frame_at(frames, 2.0, 1, 0.4 * 2.0 * 3600 + 0.05 * CARR)
    assert np.abs(data0.cpu().numpy() - host0).max() <= 1e-12 * np.abs(host0).max()
    ref_p = orc.bc_points(nr, na, g, host0, theta, phi, r1=R1)
    ref_a = orc.bc_averages(nr, na, g, order, host0, theta, phi, r1=R1)
    ep = rel_linf_per_var(bc.cpu().numpy(), ref_p)
    ea = rel_linf_per_var(inner.cpu().numpy(), ref_a)
    print("bc points err", ep.max(), "averages err", ea.max())
    assert np.isfinite(bc.cpu().numpy()).all() and ep.max() <= 1e-12, ep
    assert np.isfinite(inner.cpu().numpy()).all() and ea.max() <= 1e-12, ea
