"""GPU parity of the batched fourth-order box operator over an uneven,
multi-level box table (the batched form of the AMR stage evaluation,
src/MHD_AMR.H:1417-1499, which applies BoxOp_EulerCubedSphere::operator() to
every patch of every level): boxes of three refinement levels and several
sizes, interleaved in the table, evaluated in the same launches; each box must
match the oracle evaluated at its own level within relative L-infinity 1e-12."""
import numpy as np
import pytest
import torch

import heliocubed_amd as hc
import oracle as orc
from helpers import per_var, rel_linf_per_var

TOL = 1e-12
SHELLS = [hc.Shell(32, 16), hc.Shell(64, 32), hc.Shell(128, 64)]
TABLE = [
    # (level, (block, lo, hi))
    (1, (2, (20, 8, 10), (35, 23, 25))),
    (0, (3, (8, 0, 4), (15, 7, 11))),
    (2, (2, (40, 16, 20), (47, 23, 27))),
    (2, (2, (48, 16, 20), (55, 23, 27))),
    (0, (0, (16, 4, 8), (23, 11, 15))),
    (1, (5, (8, 4, 0), (39, 19, 7))),
    (2, (1, (60, 30, 36), (67, 37, 43))),
]


def _inputs(opts, config, seed=7):
    parts = []
    for lev, box in TABLE:
        parts.append(hc.init_native_boxes(SHELLS[lev], opts, [box], config=config, seed=seed, quad=2))
    return np.concatenate(parts)


def _reference(U, opts):
    out, off = [], 0
    for lev, (blk, lo, hi) in TABLE:
        n = [hi[d] - lo[d] + 1 for d in range(3)]
        size = 8 * (n[0] + 10) * (n[1] + 10) * (n[2] + 10)
        sh = SHELLS[lev]
        out.append(orc.o4_rhs_boxes(sh.n_rad, sh.n_ang, [(blk, lo, hi)], U[off:off + size], limiter=opts.limiter,
                                    riemann=opts.riemann_solver, spatial_operator=opts.spatial_operator,
                                    eta_d=opts.berta_eta_d))
        off += size
    return np.concatenate(out)


@pytest.mark.gpu
@pytest.mark.parametrize("operator,riemann,config", [(1, 2, 3), (2, 2, 4), (2, 1, 3)])
def test_multilevel_box_table_matches_oracle(device, operator, riemann, config):
    opts = hc.Options(spatial_order=4, limiter=1, riemann_solver=riemann, spatial_operator=operator)
    U = _inputs(opts, config)
    ref = _reference(U, opts)
    boxes = [b for _, b in TABLE]
    op = hc.BoxOp(SHELLS, opts, boxes, levels=[lev for lev, _ in TABLE])
    assert op.count(0) == U.size and op.count(1) == ref.size
    R = torch.zeros(op.count(1), dtype=torch.float64, device=device)
    op.apply(torch.from_numpy(U).to(device), R)
    torch.cuda.synchronize()
    got = R.cpu().numpy()
    assert np.isfinite(got).all()
    err = rel_linf_per_var(per_var(got, boxes), per_var(ref, boxes), axis_var=0)
    print("multilevel table err", operator, riemann, err)
    assert err.max() <= TOL, err
