"""Box layouts of the seeded test problems (not on the hot path).

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
