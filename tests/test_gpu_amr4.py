"""GPU parity of the fixed-hierarchy AMR RK4 of the fourth-order operators
(advanceRungeKuttaSynchronous, src/MHD_AMRRungeKutta.H:5-41): the stage
input of every level (prepareFourthOrder, src/MHD_AMRFourthOrder.H:5-187),
one stage RHS with the flux registers and the restriction (src/MHD_AMR.H:
597-607, 651-663, 916-921, 1417-1499) and whole RK4 steps, on a three-level
hierarchy (level 0 = 6 x 24 x 16 x 16, 10 patches of 8^3 at level 1, 8 at
level 2) against the CPU oracle: FP64 within relative L-infinity 1e-12 per
variable over the hierarchy, mixed precision within 1e-6 on the stage RHS."""

import numpy as np
import pytest
import torch

import amr4_cases as C
import heliocubed_amd as hc
import oracle as orc

TOL = 1e-12


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())).cuda()


def _case(operator, riemann, config, precision=0):
    shells, patches = C.small_hierarchy()
    opts = C.options(operator=operator, riemann=riemann, precision=precision)
    ju, inner = C.initial_state(shells, patches, opts, config=config)
    A = hc.Amr4(C.hc_shells(shells), opts, patches)
    O = orc.Amr4Oracle(shells, patches, r0=C.R0, r1=C.R1, spatial_operator=operator, riemann=riemann)
    return shells, patches, ju, inner, A, O


CASES = [(1, 2, 3), (2, 2, 4), (2, 1, 3)]


@pytest.mark.gpu
@pytest.mark.parametrize("operator,riemann,config", CASES)
def test_amr4_prepare_matches_oracle(device, operator, riemann, config):
    print("case amr4 prepare", operator, riemann, config, flush=True)
    shells, patches, ju, inner, A, O = _case(operator, riemann, config)
    assert A.state_count == ju.size and A.stage_count == O.stage_size
    US = torch.zeros(A.stage_count, dtype=torch.float64, device=device)
    A.prepare(_dev(ju), _dev(inner), US)
    torch.cuda.synchronize()
    got, ref = US.cpu().numpy(), O.prepare(ju, inner)
    assert np.isfinite(got).all()
    err = C.rel_err_per_var(got, ref, shells, patches, stage=True)
    print("amr4 prepare err", err)
    assert err.max() <= TOL, err


@pytest.mark.gpu
@pytest.mark.parametrize("operator,riemann,config", CASES)
def test_amr4_stage_rhs_matches_oracle(device, operator, riemann, config):
    print("case amr4 stage rhs", operator, riemann, config, flush=True)
    shells, patches, ju, inner, A, O = _case(operator, riemann, config)
    K = torch.zeros(A.state_count, dtype=torch.float64, device=device)
    A.stage_rhs(_dev(ju), _dev(inner), K)
    torch.cuda.synchronize()
    got, ref = K.cpu().numpy(), O.stage_rhs(ju, inner)
    assert np.isfinite(got).all()
    err = C.rel_err_per_var(got, ref, shells, patches)
    print("amr4 stage rhs err", err)
    assert err.max() <= TOL, err


@pytest.mark.gpu
@pytest.mark.parametrize("operator,config", [(1, 3), (2, 4)])
def test_amr4_mixed_stage_rhs_within_1e6(device, operator, config):
    print("case amr4 mixed stage rhs", operator, config, flush=True)
    shells, patches, ju, inner, A, O = _case(operator, 2, config, precision=1)
    K = torch.zeros(A.state_count, dtype=torch.float64, device=device)
    A.stage_rhs(_dev(ju), _dev(inner), K)
    torch.cuda.synchronize()
    err = C.rel_err_per_var(K.cpu().numpy(), O.stage_rhs(ju, inner), shells, patches)
    print("amr4 mixed stage rhs err", err)
    assert err.max() <= 1e-6, err


@pytest.mark.gpu
@pytest.mark.parametrize("operator,config", [(1, 3), (2, 4)])
def test_amr4_rk4_matches_oracle(device, operator, config):
    print("case amr4 rk4", operator, config, flush=True)
    shells, patches, ju, inner, A, O = _case(operator, 2, config)
    dt = 0.2 * (C.R0 * np.pi / 2 / shells[-1][1]) / 1.6
    U = _dev(ju)
    A.rk4_advance(U, _dev(inner), dt, 2)
    torch.cuda.synchronize()
    got, ref = U.cpu().numpy(), O.rk4(ju, inner, dt, 2)
    assert np.isfinite(got).all()
    err = C.rel_err_per_var(got, ref, shells, patches)
    inc = C.rel_err_per_var(got - ju, ref - ju, shells, patches)
    print("amr4 rk4 err", err, "increment err", inc)
    assert err.max() <= TOL, err
    assert inc.max() <= 1e-10, inc
    # covered coarse cells hold the mean of their children after the step
    g0, gp = C.split(got, shells, patches)
    l, (b, lo, hi) = patches[0]
    avg = gp[0].reshape(8, 4, 2, 4, 2, 4, 2).mean(axis=(2, 4, 6))
    c = [x // 2 for x in lo]
    cov = g0[b, :, c[2]:c[2] + 4, c[1]:c[1] + 4, c[0]:c[0] + 4]
    scale = np.abs(avg).max(axis=(1, 2, 3))[:, None, None, None]
    assert (np.abs(cov - avg) <= 1e-14 * scale).all()


@pytest.mark.gpu
@pytest.mark.parametrize("operator", [1, 2])
def test_amr4_tilings_are_identical(device, operator):
    """Refined patches prepared and evaluated per patch (the reference's
    structure), in tiles of whole patches, and with other level-0 tiles give
    the same stage RHS and RK4 step bit for bit."""
    print("case amr4 tilings", operator, flush=True)
    shells, patches, ju, inner, A, O = _case(operator, 2, 3 if operator == 1 else 4)
    dt = 0.2 * (C.R0 * np.pi / 2 / shells[-1][1]) / 1.6
    out = []
    for level0, refined in ((None, None), (None, (8, 8, 8)), ((24, 16, 16), (16, 8, 16)), ((16, 32, 16), None)):
        A.set_tiles(level0, refined)
        K = torch.zeros(A.state_count, dtype=torch.float64, device=device)
        A.stage_rhs(_dev(ju), _dev(inner), K)
        U = _dev(ju)
        A.rk4_advance(U, _dev(inner), dt, 1)
        torch.cuda.synchronize()
        print("tiles", level0, refined, "refined tiles", A.tile_count)
        out.append((K.cpu().numpy(), U.cpu().numpy()))
    for k, u in out[1:]:
        assert np.array_equal(k, out[0][0]) and np.array_equal(u, out[0][1])
