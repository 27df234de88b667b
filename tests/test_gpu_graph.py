"""The hipGraph-replayed RK4 step equals the eagerly launched step bit for
bit, across replays separated by unrelated eager device work and
synchronizations, and after a re-capture for a new dt (second- and
fourth-order levels)."""
import numpy as np
import pytest
import torch

import heliocubed_amd as hc


def _level_inputs(order):
    if order == 4:
        shell = hc.Shell(16, 16, 21.5, 60.0)
        opts = hc.Options(spatial_order=4, limiter=1, riemann_solver=2)
        ju, inner = hc.init_level4(shell, opts, 3, 4, 2)
    else:
        shell = hc.Shell(16, 8)
        opts = hc.Options(spatial_order=2, limiter=1, riemann_solver=2)
        ju, inner = hc.init_state(shell, opts, 0, 4)
    return shell, opts, ju, inner


@pytest.mark.gpu
@pytest.mark.parametrize("order", [2, 4])
def test_graph_replay_matches_eager(device, order):
    shell, opts, ju, inner = _level_inputs(order)
    dt = 0.2 * (21.5 * np.pi / 2 / shell.n_ang) / 1.6
    G = torch.from_numpy(inner).to(device)
    eager = hc.Level(shell, opts)
    eager.use_graph(False)
    graph = hc.Level(shell, opts)
    Ue = torch.from_numpy(ju).to(device)
    Ug = torch.from_numpy(ju).to(device)
    noise = torch.randn(1 << 20, device=device, dtype=torch.float64)
    for rep in range(4):
        eager.rk4_advance(Ue, G, dt, 1)
        graph.rk4_advance(Ug, G, dt, 1)
        noise = torch.sin(noise) + 1.0      # unrelated eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(Ue, Ug), rep
    dt2 = 0.5 * dt                          # new dt: re-captured graph
    eager.rk4_advance(Ue, G, dt2, 2)
    graph.rk4_advance(Ug, G, dt2, 2)
    torch.cuda.synchronize()
    assert torch.equal(Ue, Ug)
    assert torch.isfinite(Ug).all()


@pytest.mark.gpu
def test_box_table_graph_replay_matches_eager(device):
    """A multi-level box table evaluation (hc_boxop_apply) captured in a
    hipGraph through torch's stream capture replays bit for bit like the eager
    call, across replays separated by eager evaluations and unrelated work."""
    print("box-table graph replay", flush=True)
    levels = [(32, 16), (64, 32)]
    table = [(0, (0, (0, 0, 0), (15, 7, 7))), (0, (3, (16, 8, 0), (31, 15, 7))),
             (1, (1, (8, 8, 8), (15, 15, 15))), (1, (1, (16, 8, 8), (23, 15, 15)))]
    shells = [hc.Shell(nr, na) for nr, na in levels]
    opts = hc.Options(spatial_order=4, limiter=1, riemann_solver=2)
    boxes = [b for _, b in table]
    U = np.concatenate([hc.init_native_boxes(shells[l], opts, [b], config=3, seed=9, quad=1) for l, b in table])
    op = hc.BoxOp(shells, opts, boxes, levels=[l for l, _ in table])
    Ud = torch.from_numpy(U).to(device)
    Rg = torch.zeros(op.count(1), dtype=torch.float64, device=device)
    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(graph, stream=side):
            op.apply(Ud, Rg)
    torch.cuda.current_stream().wait_stream(side)
    noise = torch.randn(1 << 20, device=device, dtype=torch.float64)
    for rep in range(3):
        Rg.zero_()
        graph.replay()
        Re = torch.zeros_like(Rg)
        op.apply(Ud, Re)
        noise = torch.sin(noise) + 1.0
        torch.cuda.synchronize()
        assert torch.isfinite(Rg).all()
        assert torch.equal(Rg, Re), rep
