"""Seeded synthetic particle scenes.

All randomness comes from Philox4x32-10 (Salmon et al., SC'11) keyed by
(seed, stream) and counted by particle slot, so a scene is a pure function of
its parameters on any machine.  Jittered seeding places ppc = 8 particles per
cell at the 2x2x2 sub-positions x = (i + (a + u)/2) * h, u ~ U[0,1).
Positions are generated in FP32 (widened exactly to FP64 where the
reference-compatible path needs doubles).
"""
from __future__ import annotations

from typing import Callable, Optional, Tuple

import numpy as np

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = np.uint32(0x9E3779B9)
_W1 = np.uint32(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32-10; counters are uint32 arrays, key two uint32."""
    c = [np.asarray(x, dtype=np.uint32).copy() for x in (c0, c1, c2, c3)]
    k0 = np.uint32(k0)
    k1 = np.uint32(k1)
    with np.errstate(over="ignore"):
        for r in range(10):
            p0 = _M0 * c[0].astype(np.uint64)
            p1 = _M1 * c[2].astype(np.uint64)
            hi0 = (p0 >> np.uint64(32)).astype(np.uint32)
            lo0 = (p0 & _MASK).astype(np.uint32)
            hi1 = (p1 >> np.uint64(32)).astype(np.uint32)
            lo1 = (p1 & _MASK).astype(np.uint32)
            c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
            if r != 9:
                k0 = np.uint32(k0 + _W0)
                k1 = np.uint32(k1 + _W1)
    return c


def uniform01(seed: int, stream: int, count: int) -> np.ndarray:
    """count x 4 FP32 uniforms in [0, 1) (24-bit mantissa from each word)."""
    idx = np.arange(count, dtype=np.uint64)
    lo = (idx & _MASK).astype(np.uint32)
    hi = (idx >> np.uint64(32)).astype(np.uint32)
    words = philox4x32_10(lo, hi, np.full(count, stream, np.uint32), np.zeros(count, np.uint32),
                          seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
    return np.stack([(w >> np.uint32(8)).astype(np.float32) * np.float32(1.0 / 16777216.0) for w in words], axis=1)


def jittered_cells(n: Tuple[int, int, int], h: float, seed: int,
                   keep: Optional[Callable[[np.ndarray, np.ndarray, np.ndarray], np.ndarray]] = None,
                   origin=(0.0, 0.0, 0.0), cell_mask: Optional[np.ndarray] = None):
    """8 particles per cell at jittered 2x2x2 sub-positions; FP32 positions.

    cell_mask (nx, ny, nz bool, indexed [i, j, k]) selects seeded cells;
    keep(x, y, z) filters particles.  Slots are counted over the full
    grid x 8 so a particle's jitter does not depend on the filters.
    """
    nx, ny, nz = n
    if cell_mask is None:
        cells = np.arange(nx * ny * nz, dtype=np.int64)
    else:
        cells = np.flatnonzero(cell_mask.transpose(2, 1, 0).reshape(-1))  # x fastest
    i = cells % nx
    j = (cells // nx) % ny
    k = cells // (nx * ny)
    slot = (cells[:, None] * 8 + np.arange(8)[None, :]).reshape(-1)
    sub = np.tile(np.arange(8), cells.size)
    a = (sub & 1).astype(np.float32)
    b = ((sub >> 1) & 1).astype(np.float32)
    c = ((sub >> 2) & 1).astype(np.float32)
    u = _uniform_at(seed, 0, slot)
    hf = np.float32(h)
    ci = np.repeat(i, 8).astype(np.float32)
    cj = np.repeat(j, 8).astype(np.float32)
    ck = np.repeat(k, 8).astype(np.float32)
    x = np.float32(origin[0]) + (ci + (a + u[:, 0]) * np.float32(0.5)) * hf
    y = np.float32(origin[1]) + (cj + (b + u[:, 1]) * np.float32(0.5)) * hf
    z = np.float32(origin[2]) + (ck + (c + u[:, 2]) * np.float32(0.5)) * hf
    if keep is not None:
        m = keep(x, y, z)
        x, y, z = x[m], y[m], z[m]
    return x.astype(np.float32), y.astype(np.float32), z.astype(np.float32)


def _uniform_at(seed: int, stream: int, slots: np.ndarray) -> np.ndarray:
    slots = slots.astype(np.uint64)
    lo = (slots & _MASK).astype(np.uint32)
    hi = (slots >> np.uint64(32)).astype(np.uint32)
    words = philox4x32_10(lo, hi, np.full(slots.size, stream, np.uint32), np.zeros(slots.size, np.uint32),
                          seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
    return np.stack([(w >> np.uint32(8)).astype(np.float32) * np.float32(1.0 / 16777216.0) for w in words], axis=1)


def uniform_random(count: int, seed: int, lo=0.0, hi=1.0) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """count particles uniform in [lo, hi)^3, FP32 (primitive microbench)."""
    u = uniform01(seed, 0, count)
    s = np.float32(hi - lo)
    return (np.float32(lo) + u[:, 0] * s, np.float32(lo) + u[:, 1] * s, np.float32(lo) + u[:, 2] * s)


def uniform_velocities(count: int, seed: int, amp=1.0):
    """count FP32 velocities uniform in [-amp, amp)^3."""
    u = uniform01(seed, 1, count)
    a = np.float32(amp)
    return tuple((u[:, d] * np.float32(2.0) - np.float32(1.0)) * a for d in range(3))


def dam_break(n: int = 64, seed: int = 1, width: float = 0.4, height: float = 0.6):
    """Config-1 dam break: unit cube, n^3 cells, fluid x < width, y < height."""
    h = 1.0 / n
    return jittered_cells((n, n, n), h, seed,
                          keep=lambda x, y, z: (x < np.float32(width)) & (y < np.float32(height)))


def rc_block_scene(n: int, h: float, count_cells: Tuple[int, int, int], seed: int, origin_cell=(0, 0, 0),
                   vel_amp: float = 0.0, vel_seed: int = 7):
    """A block of jittered particles for the reference-compatible path:
    FP64 positions (exact widening of FP32) and optional random velocities."""
    nx, ny, nz = count_cells
    mask = np.zeros((n, n, n), dtype=bool)
    ox, oy, oz = origin_cell
    mask[ox:ox + nx, oy:oy + ny, oz:oz + nz] = True
    x, y, z = jittered_cells((n, n, n), h, seed, cell_mask=mask)
    pos = [a.astype(np.float64) for a in (x, y, z)]
    if vel_amp:
        v = uniform_velocities(x.size, vel_seed, vel_amp)
        vel = [a.astype(np.float64) for a in v]
    else:
        vel = [np.zeros(x.size) for _ in range(3)]
    return pos, vel


def _sphere_cells(n: int, center, radius: float) -> np.ndarray:
    """Bool [i, j, k] mask of cells whose 8 jitter sub-cells may fall in the sphere
    (cell centre within radius + half a cell diagonal)."""
    h = 1.0 / n
    lo = [max(0, int((c - radius) / h) - 1) for c in center]
    hi = [min(n, int((c + radius) / h) + 2) for c in center]
    m = np.zeros((n, n, n), dtype=bool)
    ii, jj, kk = np.meshgrid(*(np.arange(l, u) for l, u in zip(lo, hi)), indexing="ij")
    d2 = sum(((a + 0.5) * h - c) ** 2 for a, c in zip((ii, jj, kk), center))
    m[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = d2 <= (radius + 0.87 * h) ** 2
    return m


def dam_break_sphere(n: int = 128, seed: int = 5, center=(0.6, 0.2, 0.5), radius: float = 0.1):
    """Dam break (x < 0.4, y < 0.6) with a static solid sphere.  Returns the
    particles and the solid mask (n^3 bytes indexed x + y*n + z*n*n, cells
    whose centre lies inside the sphere)."""
    x, y, z = dam_break(n, seed)
    c = np.float32(center)
    r2 = np.float32(radius * radius)
    out = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 > r2
    h = 1.0 / n
    a = (np.arange(n) + 0.5) * h
    zz, yy, xx = np.meshgrid(a, a, a, indexing="ij")
    solid = ((xx - center[0]) ** 2 + (yy - center[1]) ** 2 + (zz - center[2]) ** 2 <= radius * radius)
    return (x[out], y[out], z[out]), solid.astype(np.uint8).reshape(-1)


def colliding_spheres(n: int = 512, seed: int = 4, radius: float = 0.15, speed: float = 2.0):
    """Two fluid spheres at (0.3, .5, .5) moving +x and (0.7, .5, .5) moving -x."""
    xs, ys, zs, us = [], [], [], []
    h = 1.0 / n
    for cx, u in ((0.3, speed), (0.7, -speed)):
        center = (cx, 0.5, 0.5)
        mask = _sphere_cells(n, center, radius)
        c = np.float32(center)
        r2 = np.float32(radius * radius)
        x, y, z = jittered_cells((n, n, n), h, seed, cell_mask=mask,
                                 keep=lambda x, y, z: (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 <= r2)
        xs.append(x); ys.append(y); zs.append(z)
        us.append(np.full(x.size, np.float32(u)))
    x, y, z, vx = (np.concatenate(a) for a in (xs, ys, zs, us))
    zero = np.zeros_like(x)
    return (x, y, z), (vx, zero, zero.copy())


def wave_surface(x, z, seed: int, depth: float = 0.35, amp: float = 0.02, modes: int = 4):
    """Seeded surface height depth + amp * sum_k a_k sin(2 pi (f_k x + g_k z) + phi_k)."""
    u = uniform01(seed, 2, modes)
    y = np.full(np.shape(x), depth, dtype=np.float64)
    for k in range(modes):
        a = 0.5 + 0.5 * float(u[k, 0])
        f = 1 + int(3 * float(u[k, 1]))
        g = 1 + int(3 * float(u[k, 2]))
        phi = 2 * np.pi * float(u[k, 3])
        y += amp * a * np.sin(2 * np.pi * (f * np.asarray(x, np.float64) + g * np.asarray(z, np.float64)) + phi) / modes
    return y


def wave_tank(n: int = 256, seed: int = 6, depth: float = 0.35, paddle_x: float = 0.05):
    """Wave tank: fluid below a seeded perturbed surface around `depth`, right
    of the paddle plane x = paddle_x (the paddle's rest position)."""
    h = 1.0 / n
    top = int(np.ceil((depth + 0.03) / h)) + 1
    mask = np.zeros((n, n, n), dtype=bool)
    mask[:, :min(n, top), :] = True
    x, y, z = jittered_cells((n, n, n), h, seed, cell_mask=mask)
    keep = (y < wave_surface(x, z, seed, depth).astype(np.float32)) & (x >= np.float32(paddle_x))
    return x[keep], y[keep], z[keep]


def paddle_position(t: float, rest: float = 0.05, amp: float = 0.05, freq: float = 1.0):
    """Paddle plane x(t) = rest + amp sin(2 pi f t) and its velocity."""
    w = 2 * np.pi * freq
    return rest + amp * np.sin(w * t), amp * w * np.cos(w * t)
