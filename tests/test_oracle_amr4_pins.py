"""Pins of the oracle's fixed-hierarchy AMR RK4 (advanceRungeKuttaSynchronous,
src/MHD_AMRRungeKutta.H:5-41, with prepareFourthOrder,
src/MHD_AMRFourthOrder.H:5-187, and the flux registers of
src/MHD_AMR.H:651-663, 916-921):

* conservation: with the flux registers, the level-0 sum of the stage RHS
  (covered cells hold the restricted fine RHS) equals that of the bare
  level-0 operator on the restricted state, i.e. minus the same radial
  boundary fluxes, to round-off; without the registers it does not (the
  control shows the check has power);
* the prepared stage input of the refined patches converges to the exact
  native averages of an analytic state (numpy quadrature of the closed-form
  panel map, tests/exact_shell.py): fourth order inside the patches, at
  least third order in the interpolated ghosts, second order on the first
  valid layer next to a coarse-fine face (the reference computes the first
  ghost layer's average-product correction while the donor's own cells are
  still zero, src/MHD_AMRFourthOrder.H:10,37-63);
* a hierarchy without refined levels prepares the level by the one-pass
  <JU> -> <U> of the AMR path, fourth-order convergent, and differs from the
  single-level two-pass form (src/MHD_StateConversion.H:153-246) only at
  truncation level;
* the table order of the patches does not change the result.
"""
import numpy as np
import pytest

import amr4_cases as C
import exact_shell as X
import oracle as orc


def _restricted_level0(ju, shells, patches):
    s0, sp = C.split(ju.copy(), shells, patches)
    s0 = s0.copy()
    sp = [x.copy() for x in sp]
    for lev in range(len(shells) - 1, 0, -1):
        for (l, (b, lo, hi)), arr in zip(patches, sp):
            if l != lev:
                continue
            avg = arr.reshape(8, arr.shape[1] // 2, 2, arr.shape[2] // 2, 2, arr.shape[3] // 2, 2).mean(axis=(2, 4, 6))
            c = [x // 2 for x in lo]
            if lev == 1:
                s0[b, :, c[2]:c[2] + avg.shape[1], c[1]:c[1] + avg.shape[2], c[0]:c[0] + avg.shape[3]] = avg
                continue
            for (l2, (b2, lo2, _)), arr2 in zip(patches, sp):
                if l2 == lev - 1 and b2 == b and all(lo2[d] <= c[d] < lo2[d] + arr2.shape[3 - d] for d in range(3)):
                    o = [c[d] - lo2[d] for d in range(3)]
                    arr2[:, o[2]:o[2] + avg.shape[1], o[1]:o[1] + avg.shape[2], o[0]:o[0] + avg.shape[3]] = avg
    return s0


@pytest.mark.parametrize("operator,config", [(1, 3), (2, 4)])
def test_flux_registers_conserve(operator, config):
    shells, patches = C.small_hierarchy()
    opts = C.options(operator=operator)
    ju, inner = C.initial_state(shells, patches, opts, config=config)
    O = orc.Amr4Oracle(shells, patches, r0=C.R0, r1=C.R1, spatial_operator=operator)
    k = O.stage_rhs(ju, inner)
    k_bare = O.stage_rhs(ju, inner, reflux=False)
    O1 = orc.Amr4Oracle(shells[:1], [], r0=C.R0, r1=C.R1, spatial_operator=operator)
    ku = O1.stage_rhs(_restricted_level0(ju, shells, patches).ravel(), inner)
    n0 = ku.size
    tot = lambda x: x[:n0].reshape(6, 8, -1).sum(axis=(0, 2))
    scale = np.abs(ku.reshape(6, 8, -1)).sum(axis=(0, 2))
    err = np.abs(tot(k) - tot(ku)) / scale
    err_bare = np.abs(tot(k_bare) - tot(ku)) / scale
    print("conservation with registers", err, "without", err_bare)
    assert err.max() < 1e-13
    assert err_bare[[0, 4]].min() > 1e-6


def _exact_inputs(shells, patches, quad=3):
    E = [X.Shell(nr, na, C.R0, C.R1) for nr, na in shells]
    na0 = shells[0][1]
    ju0 = np.stack([E[0].ju_averages(p, quad=quad) for p in range(6)])
    parts = [ju0.ravel()] + [E[l].ju_box(b, lo, hi, quad=quad).ravel() for l, (b, lo, hi) in patches]
    inner = np.stack([E[0].native_averages(p, range(-5, 0), range(-5, na0 + 5), range(-5, na0 + 5), quad=quad)
                      for p in range(6)])
    return E, np.concatenate(parts), np.ascontiguousarray(inner)


def _patch_errors(scale):
    shells, patches = C.small_hierarchy(scale, 2)
    E, ju, inner = _exact_inputs(shells, patches)
    O = orc.Amr4Oracle(shells, patches, r0=C.R0, r1=C.R1)
    _, up = C.split(O.prepare(ju, inner), shells, patches, stage=True)
    interior = first = ghost = 0
    for (l, (b, lo, hi)), u in zip(patches, up):
        ex = E[l].native_averages(b, range(lo[0] - 5, hi[0] + 6), range(lo[1] - 5, hi[1] + 6),
                                  range(lo[2] - 5, hi[2] + 6), quad=3)
        d = np.abs(u - ex) / np.abs(ex).max(axis=(1, 2, 3))[:, None, None, None]
        inner_m = np.zeros(d.shape[1:], bool)
        inner_m[6:-6, 6:-6, 6:-6] = True
        valid_m = np.zeros(d.shape[1:], bool)
        valid_m[5:-5, 5:-5, 5:-5] = True
        interior = np.maximum(interior, d[:, inner_m].max(axis=1))
        first = np.maximum(first, d[:, valid_m & ~inner_m].max(axis=1))
        ghost = np.maximum(ghost, d[:, ~valid_m].max(axis=1))
    return interior, first, ghost


def test_refined_patch_preparation_converges():
    a, b = _patch_errors(1), _patch_errors(2)
    rates = [np.log2(x / y) for x, y in zip(a, b)]
    print("rates interior", rates[0], "first layer", rates[1], "ghosts", rates[2])
    print("errors at scale 2", b)
    # patch interiors: densities and energies (the well-resolved scalars) at
    # fourth order, every component above 3
    assert rates[0][[0, 4]].min() > 3.5 and rates[0].min() > 3.0
    assert rates[2].min() > 2.5
    assert rates[1].min() > 1.2
    assert max(x.max() for x in b) < 5e-6


def test_single_level_one_pass_preparation_converges():
    errs, diffs = [], []
    for scale in (1, 2):
        shells, _ = C.small_hierarchy(scale, 1)
        E, ju, inner = _exact_inputs(shells, [])
        nr, na = shells[0]
        O = orc.Amr4Oracle(shells, [], r0=C.R0, r1=C.R1)
        us = O.prepare(ju, inner).reshape(6, 8, na + 10, na + 10, nr + 10)
        ex = np.stack([E[0].native_averages(p, range(nr), range(-5, na + 5), range(-5, na + 5), quad=3)
                       for p in (0, 3)])
        u = us[[0, 3], :, :, :, 5:-5]
        errs.append(np.abs(u - ex).max(axis=(0, 2, 3, 4)) / np.abs(ex).max(axis=(0, 2, 3, 4)))
        single = orc.Oracle4(nr, na, r1=C.R1).prepare(ju.reshape(6, 8, na, na, nr), inner)
        diffs.append(np.abs(us - single).max() / np.abs(single).max())
    rates = np.log2(errs[0] / errs[1])
    print("one-pass rates", rates, "differences to the two-pass form", diffs)
    assert rates[[0, 4]].min() > 3.5 and rates.min() > 3.0
    assert 0 < diffs[1] < diffs[0] < 1e-4


def test_patch_order_does_not_matter():
    shells, patches = C.small_hierarchy()
    opts = C.options()
    ju, inner = C.initial_state(shells, patches, opts, config=3)
    O = orc.Amr4Oracle(shells, patches, r0=C.R0, r1=C.R1)
    k = O.stage_rhs(ju, inner)
    perm = np.random.default_rng(3).permutation(len(patches))
    p2 = [patches[i] for i in perm]
    _, parts = C.split(ju, shells, patches)
    k0 = C.split(k, shells, patches)
    ju2 = np.concatenate([ju[:k0[0].size]] + [parts[i].ravel() for i in perm])
    k2 = orc.Amr4Oracle(shells, p2, r0=C.R0, r1=C.R1).stage_rhs(ju2, inner)
    a0, ap = C.split(k2, shells, p2)
    assert np.array_equal(a0, k0[0])
    for i, x in zip(perm, ap):
        assert np.array_equal(x, k0[1][i])
