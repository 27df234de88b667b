"""Box layouts of the seeded test problems (not on the hot path).

amr_rk4_hierarchy(): the three-level hierarchy of the AMR RK4 workload: a base
level of 6 x 64 radial x 48x48 angular cells (884,736 cells; 108 boxes of
32x16x16 in the reference's layout), level 1 (ratio 2) in 16^3 patches over a
30-degree wedge of panel 1 (angular [32, 64)^2, radial [16, 112) of level 1:
24 patches), level 2 (ratio 4) in 8^3 patches over the inner part of the
wedge (radial [64, 192), angular [72, 120)^2 of level 2: 576 patches); 708
boxes, 1,277,952 cells. Every refined patch keeps its ghost regions inside its panel and the
radial domain and the levels are properly nested, as the fixed-hierarchy RK4
step requires.

amr_3level_table(): the three-level AMR hierarchy of the batched workload: a
base level of 6 x 64 radial x 48x48 angular cells in 32x16x16 boxes, level 1
(ratio 2) in 16^3 boxes over a 30-degree wedge (the central 32x32 angular
cells of panel 1 at level 1), level 2 (ratio 4) in 8^3 boxes at its shock
front (the radial band [96, 160) of level 2 inside the wedge): 652 boxes of
uneven sizes, the batched form of the AMR stage evaluation
(src/MHD_AMR.H:1417-1499).
"""

AMR_LEVELS = [(64, 48), (128, 96), (256, 192)]


def amr_3level_table():
    """Levels (n_rad, n_ang) and the (level, (block, lo, hi)) box table."""
    table = []
    for p in range(6):
        for k in range(0, 48, 16):
            for j in range(0, 48, 16):
                for i in range(0, 64, 32):
                    table.append((0, (p, (i, j, k), (i + 31, j + 15, k + 15))))
    for k in range(32, 64, 16):
        for j in range(32, 64, 16):
            for i in range(0, 128, 16):
                table.append((1, (1, (i, j, k), (i + 15, j + 15, k + 15))))
    for k in range(64, 128, 8):
        for j in range(64, 128, 8):
            for i in range(96, 160, 8):
                table.append((2, (1, (i, j, k), (i + 7, j + 7, k + 7))))
    return list(AMR_LEVELS), table


AMR4_LEVELS = [(64, 48), (128, 96), (256, 192)]


def amr_rk4_hierarchy():
    """Levels (n_rad, n_ang) and the refined patches (level, (block, lo, hi))."""
    patches = []
    for k in range(32, 64, 16):
        for j in range(32, 64, 16):
            for i in range(16, 112, 16):
                patches.append((1, (1, (i, j, k), (i + 15, j + 15, k + 15))))
    for k in range(72, 120, 8):
        for j in range(72, 120, 8):
            for i in range(64, 192, 8):
                patches.append((2, (1, (i, j, k), (i + 7, j + 7, k + 7))))
    return list(AMR4_LEVELS), patches
