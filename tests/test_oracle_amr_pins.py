"""CPU pins of the oracle's AMR coarse-to-fine transfers
(amrInterpolateFourthOrder, src/MHD_AMRInterpolation.H:39-161; second-order
fine ghosts, src/MHD_AMR.H:1282-1393).

Expectations come from exact cell averages of polynomials (primitive
functions), not from the oracle's code: the fourth-order rule reproduces
tensor quartics exactly and conserves each parent's average; the quadratic
ghost rule reproduces tensor quadratics; invalid averages fall back to the
parent; owned cells and cells with a non-finite parent are left alone.
"""
import numpy as np
import pytest

import oracle as orc

GAMMA = 5.0 / 3.0


def _avg1(c, a, b):
    """Average over [a, b] of sum_n c[n] x^n."""
    return sum(cn * (b ** (n + 1) - a ** (n + 1)) / ((n + 1) * (b - a)) for n, cn in enumerate(c))


def _box_avgs(terms, lo, hi, scale):
    """Averages over cells [p/scale, (p+1)/scale] of sum_t prod_d poly_t,d."""
    axes = [np.arange(lo[d], hi[d] + 1) for d in range(3)]
    out = np.zeros((hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1))
    for t in terms:
        f = [np.array([_avg1(t[d], p / scale[d], (p + 1) / scale[d]) for p in axes[d]]) for d in range(3)]
        out += f[2][:, None, None] * f[1][None, :, None] * f[0][None, None, :]
    return out


def test_weights_are_conservative_and_exact_for_quartics():
    for r in (2, 3, 4, 8):
        W = np.array([orc.amr4_weights(r, c) for c in range(r)])
        assert np.allclose(W.sum(axis=1), 1.0, atol=1e-14)
        assert np.allclose(W.mean(axis=0), [0, 0, 1, 0, 0], atol=1e-14)
        for n in range(5):
            coarse = np.array([_avg1([0] * n + [1], k - 0.5, k + 0.5) for k in range(-2, 3)])
            for c in range(r):
                x = (c + 0.5) / r - 0.5
                assert abs(W[c] @ coarse - _avg1([0] * n + [1], x - 0.5 / r, x + 0.5 / r)) <= 1e-13
    assert np.array_equal(orc.amr4_weights(1, 0), [0, 0, 1, 0, 0])


@pytest.mark.parametrize("ratio", [(2, 2, 2), (2, 1, 2), (3, 3, 1), (1, 1, 1)])
def test_interp4_reproduces_tensor_quartics(ratio):
    rng = np.random.default_rng(sum(ratio))
    terms = [[rng.standard_normal(5) for _ in range(3)] for _ in range(2)]
    clo, chi = (-3, -2, 0), (6, 5, 7)
    coarse = np.stack([_box_avgs(terms, clo, chi, (1, 1, 1)) * (1 + c) for c in range(3)])
    # fine target whose stencils stay inside the coarse patch
    flo = tuple(ratio[d] * (clo[d] + 2) for d in range(3))
    fhi = tuple(ratio[d] * (chi[d] - 2) + ratio[d] - 1 for d in range(3))
    fine = orc.amr_interp4(coarse, clo, chi, ratio, flo, fhi)
    exact = np.stack([_box_avgs(terms, flo, fhi, ratio) * (1 + c) for c in range(3)])
    assert np.abs(fine - exact).max() <= 1e-11 * np.abs(exact).max()
    if ratio == (1, 1, 1):
        sub = coarse[:, flo[2] - clo[2]:fhi[2] - clo[2] + 1, flo[1] - clo[1]:fhi[1] - clo[1] + 1,
                     flo[0] - clo[0]:fhi[0] - clo[0] + 1]
        assert np.array_equal(fine, sub)


def test_interp4_conserves_parent_averages_and_marks_missing_donors():
    rng = np.random.default_rng(7)
    clo, chi, r = (0, 0, 0), (7, 7, 7), (2, 2, 2)
    coarse = rng.random((2, 8, 8, 8)) + 1
    fine = orc.amr_interp4(coarse, clo, chi, r, (0, 0, 0), (15, 15, 15))
    # children of coarse cells with the full stencil inside average to the parent
    inner = fine[:, 4:12, 4:12, 4:12].reshape(2, 4, 2, 4, 2, 4, 2).mean(axis=(2, 4, 6))
    assert np.allclose(inner, coarse[:, 2:6, 2:6, 2:6], rtol=1e-14)
    # stencils reaching past the coarse patch give NaN
    assert np.isnan(fine[:, 0, 0, 0]).all() and np.isnan(fine[:, 15, 8, 8]).all()
    assert np.isfinite(fine[:, 4:12, 4:12, 4:12]).all()


def _state_terms(rng):
    # conserved components as tensor quadratics, gentle enough over the patch
    # (|x| <= 6) that density and thermal energy stay positive
    return [[[np.array([1.0, 0.02 * rng.standard_normal(), 0.004 * rng.standard_normal()]) for _ in range(3)]]
            for _ in range(8)]


SCALARS = [1.0, 0.1, -0.2, 0.3, 10.0, 0.5, -0.4, 0.2]


def _coarse_state(J=2.5, seed=3):
    rng = np.random.default_rng(seed)
    clo, chi = (-2, -2, -2), (5, 5, 5)
    terms = _state_terms(rng)
    coarse = np.zeros((9, 8, 8, 8))
    for c in range(8):
        coarse[c] = J * SCALARS[c] * _box_avgs(terms[c], clo, chi, (1, 1, 1))
    coarse[8] = J
    return coarse, terms, clo, chi


def _primitives(u):
    rho = u[0]
    v = u[1:4] / rho
    p = (GAMMA - 1) * (u[4] - 0.5 * rho * (v ** 2).sum(0) - (u[5:8] ** 2).sum(0) / (8 * np.pi))
    return np.concatenate([rho[None], v, p[None], u[5:8]])


FLO, FHI, OLO, OHI, R2 = (-2, -2, -2), (9, 9, 9), (0, 0, 0), (7, 7, 7), (2, 2, 2)


def test_ghosts2_reproduce_quadratics():
    coarse, terms, clo, chi = _coarse_state()
    init = np.full((8, 12, 12, 12), -7.0)
    got = orc.amr_ghosts2(coarse, clo, chi, R2, FLO, FHI, OLO, OHI, init)
    u = np.stack([SCALARS[c] * _box_avgs(terms[c], FLO, FHI, R2) for c in range(8)])
    exp = _primitives(u)
    ghost = np.ones((12, 12, 12), bool)
    ghost[2:10, 2:10, 2:10] = False
    assert np.abs(got[:, ghost] - exp[:, ghost]).max() <= 1e-12 * np.abs(exp).max()
    assert (got[:, ~ghost] == -7.0).all()


def test_ghosts2_falls_back_to_the_parent_and_skips_missing_parents():
    coarse, terms, clo, chi = _coarse_state()
    J = coarse[8, 0, 0, 0]
    init = np.full((8, 12, 12, 12), -7.0)
    # parent (-1, 2, 2); its +radial neighbour (0, 2, 2) gets a deep density hole
    par, nb = (-1, 2, 2), (0, 2, 2)
    idx = lambda c: (c[2] - clo[2], c[1] - clo[1], c[0] - clo[0])
    coarse[(0,) + idx(nb)] = -20.0 * J
    got = orc.amr_ghosts2(coarse, clo, chi, R2, FLO, FHI, OLO, OHI, init)
    parent_u = coarse[(slice(0, 8),) + idx(par)] / J
    # the child on the hole's side (x = +1/4) falls back, the other (x = -1/4) does not
    fine = lambda p: (p[2] - FLO[2], p[1] - FLO[1], p[0] - FLO[0])
    near = got[(slice(None),) + fine((-1, 4, 4))]
    far = got[(slice(None),) + fine((-2, 4, 4))]
    assert np.allclose(near, _primitives(parent_u[:, None])[:, 0], rtol=1e-14)
    assert not np.allclose(far, near)
    # a parent with non-finite J leaves its children untouched
    coarse2, _, _, _ = _coarse_state()
    coarse2[(8,) + idx((4, 2, 2))] = np.nan
    got2 = orc.amr_ghosts2(coarse2, clo, chi, R2, FLO, FHI, OLO, OHI, init)
    assert (got2[(slice(None),) + fine((8, 4, 4))] == -7.0).all()
    assert (got2[(slice(None),) + fine((9, 5, 5))] == -7.0).all()


def test_average_down_inverts_interp_and_conserves():
    """Restriction (Stencil::AvgDown): the mean of the children. The fourth-order
    interpolation is conservative, so restricting its output returns the coarse
    data; the fine total equals the coarse total times the ratio product."""
    rng = np.random.default_rng(5)
    clo, chi, r = (0, 0, 0), (7, 7, 7), (2, 3, 2)
    coarse = rng.random((2, 8, 8, 8)) + 1
    flo = tuple(r[d] * 2 for d in range(3))
    fhi = tuple(r[d] * 6 - 1 for d in range(3))
    fine = orc.amr_interp4(coarse, clo, chi, r, flo, fhi)
    back = orc.amr_average_down(fine, flo, fhi, r, (2, 2, 2), (5, 5, 5))
    assert np.allclose(back, coarse[:, 2:6, 2:6, 2:6], rtol=1e-14)
    f2 = rng.random((3, 6, 9, 4))
    c2 = orc.amr_average_down(f2, (0, 0, 0), (3, 8, 5), (2, 3, 3), (0, 0, 0), (1, 2, 1))
    assert np.allclose(c2.sum(axis=(1, 2, 3)) * 18, f2.sum(axis=(1, 2, 3)), rtol=1e-13)
    assert np.allclose(c2[:, 1, 2, 0], f2[:, 3:6, 6:9, 0:2].mean(axis=(1, 2, 3)), rtol=1e-14)


def _prolong_case(seed, jump=False):
    rng = np.random.default_rng(seed)
    clo, chi, r = (-2, -2, -2), (5, 5, 5), (2, 2, 2)
    J = 1.5
    terms = _state_terms(rng)
    coarse = np.zeros((9, 8, 8, 8))
    for c in range(8):
        coarse[c] = J * SCALARS[c] * _box_avgs(terms[c], clo, chi, (1, 1, 1))
    coarse[8] = J
    if jump:
        # a density and energy cliff: the quartic undershoots next to it
        coarse[0:5, :, :, 4:] *= 1e-3
        coarse[5:8, :, :, 4:] *= np.sqrt(1e-3)
    flo, fhi = (0, 0, 0), (7, 7, 7)                     # parents 0..3, donors -2..5
    interp = orc.amr_interp4(coarse, clo, chi, r, flo, fhi)
    fine_j = np.full((1, 8, 8, 8), J)
    return coarse, interp, fine_j, terms, clo, chi, flo, fhi, r, J


def _child_sums(fine, r=2):
    return fine.reshape(fine.shape[0], 4, r, 4, r, 4, r).sum(axis=(2, 4, 6))


def test_prolong4_is_conservative_and_exact_for_smooth_data():
    coarse, interp, fine_j, terms, clo, chi, flo, fhi, r, J = _prolong_case(13)
    fine, bad = orc.amr_prolong4(interp, fine_j, flo, fhi, coarse, clo, chi, r)
    assert bad == 0
    parents = coarse[:8, 2:6, 2:6, 2:6]
    assert np.allclose(_child_sums(fine), 8 * parents, rtol=1e-13)
    exact = np.stack([J * SCALARS[c] * _box_avgs(terms[c], flo, fhi, r) for c in range(8)])
    assert np.abs(fine - exact).max() <= 1e-12 * np.abs(exact).max()


def test_prolong4_limits_toward_the_parent_to_stay_positive():
    coarse, interp, fine_j, _, clo, chi, flo, fhi, r, J = _prolong_case(17, jump=True)
    phys = interp[:8] / interp[8]
    assert (phys[0] < 0).any() or ((phys[4] - 0.5 * (phys[1:4] ** 2).sum(0) / phys[0]
                                    - (phys[5:8] ** 2).sum(0) / (8 * np.pi)) < 0).any()
    fine, bad = orc.amr_prolong4(interp, fine_j, flo, fhi, coarse, clo, chi, r)
    assert bad == 0
    u = fine / J
    thermal = u[4] - 0.5 * (u[1:4] ** 2).sum(0) / u[0] - (u[5:8] ** 2).sum(0) / (8 * np.pi)
    assert (u[0] >= 1.67262192e-26).all() and (thermal >= 0).all()
    assert np.allclose(_child_sums(fine), 8 * coarse[:8, 2:6, 2:6, 2:6], rtol=1e-12)


def test_prolong4_reports_inadmissible_parents():
    coarse, interp, fine_j, _, clo, chi, flo, fhi, r, J = _prolong_case(19)
    coarse[0, 3, 3, 3] = -1.0          # parent (1, 1, 1): negative density
    interp = orc.amr_interp4(coarse, clo, chi, r, flo, fhi)
    _, bad = orc.amr_prolong4(interp, fine_j, flo, fhi, coarse, clo, chi, r)
    assert bad >= 1
