"""Host-side checks of hc_amr4_create: hierarchies outside the built form are
rejected with a message before any device work (no GPU needed)."""
import pytest

import amr4_cases as C
import heliocubed_amd as hc


def _create(shells, patches):
    return hc.Amr4(C.hc_shells(shells), C.options(), patches)


@pytest.mark.parametrize("mutate,message", [
    (lambda sh, p: (sh, p + [(1, (1, (0, 8, 8), (7, 15, 15)))]), "6 cells from the radial boundaries"),
    (lambda sh, p: (sh, p + [(1, (1, (16, 0, 8), (23, 7, 15)))]), "8 cells from the panel edges"),
    (lambda sh, p: (sh, p + [(1, (1, (16, 8, 8), (31, 23, 23)))]), "one box size"),
    (lambda sh, p: (sh, p + [(1, (1, (20, 8, 8), (27, 15, 15)))]), "aligned"),
    (lambda sh, p: (sh, p + [p[0]]), "duplicate"),
    (lambda sh, p: (sh, [(2, x[1]) if x[0] == 1 else x for x in p]), "no patches"),
    (lambda sh, p: ([sh[0], (40, 32), sh[2]], p), "multiples"),
    (lambda sh, p: (sh, [x for x in p if x[0] == 1][:1] + [x for x in p if x[0] == 2]), "inside one patch"),
])
def test_bad_hierarchies_are_rejected(mutate, message):
    shells, patches = C.small_hierarchy()
    sh, pt = mutate(shells, patches)
    with pytest.raises(hc.HcError, match=message):
        _create(sh, pt)


def test_nesting_is_checked():
    # a level-2 patch at the edge of the level-1 block: its coarse-fine face
    # borders level-0 cells
    shells, patches = C.small_hierarchy()
    bad = [x for x in patches if x[0] == 1] + [(2, (1, (32, 32, 16), (39, 39, 23)))]
    with pytest.raises(hc.HcError, match="interpolation source|nested"):
        _create(shells, bad)
