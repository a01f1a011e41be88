"""Synthetic pressure systems for the U-Net checks (dense [z][y][x] grids)."""
import numpy as np


def synthetic_system(n: int, seed: int = 0, fill: float = 0.55):
    """(b float32, coef uint16) of a pool of fluid below height fill*n with a
    few air pockets; walls are solid.  coef = diag | fluid-neighbour mask << 3
    (bits -x, +x, -y, +y, -z, +z), diag = number of non-solid neighbours, as
    the solver assembles it."""
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    fluid = y < int(fill * n)
    fluid &= rng.random((n, n, n)) > 0.05
    coef = np.zeros((n, n, n), np.uint16)
    offs = [(0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0)]
    diag = np.zeros((n, n, n), np.int64)
    mask = np.zeros((n, n, n), np.int64)
    for d, (dz, dy, dx) in enumerate(offs):
        zz, yy, xx = z + dz, y + dy, x + dx
        inside = (zz >= 0) & (zz < n) & (yy >= 0) & (yy < n) & (xx >= 0) & (xx < n)
        nb_fluid = np.zeros((n, n, n), bool)
        nb_fluid[inside] = fluid[zz[inside], yy[inside], xx[inside]]
        diag += inside
        mask |= nb_fluid.astype(np.int64) << d
    coef[fluid] = (diag[fluid] | (mask[fluid] << 3)).astype(np.uint16)
    b = np.where(fluid, rng.normal(size=(n, n, n)), 0.0).astype(np.float32)
    return b, coef
