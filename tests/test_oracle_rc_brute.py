"""Brute-force pins of the oracle's reference-compatible node and voxel logic.

Each check restates a piece of the reference from its source text, as
plainly as possible (Python loops over every voxel and particle of tiny
scenes, exact fma through fractions), and compares with the oracle:
  - used-voxel marking and compaction (engine/particles.cu:88-321): low-wall
    voxels by the node coordinates of the particle whose index equals the
    node index (Q6), the "y" voxel index equal to x (Q7), the upper-wall test
    (Q8), the 4^3 sub-voxel stencil skipped when it would start below 0
    (Q24), the three staggered distance checks with fma (Q9);
  - the row table (particleToGridFunctions.cu:347-378): the first node's
    cell key per row, sentinel B, and the first node missed when its row
    equals B / sizeX;
  - divU across neighbouring nodes (voxelSolveFunctions.cu:20-271): the
    apron fetch with offset 2 radius = 4, row-table keys used as node
    indices, the z-negative branch writing its z index into the x slot
    (Q17), forward differences per component;
  - G2P (particles.cu:562-662): voxels holding exactly 0.0 take no weight;
  - the dt-halving retry of pressureSolve (particles.cu:528-549) driven by
    scripted residual sequences.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

N1 = 8           # voxels per node edge: 2 floor(radius) + (2 << refinementLevel)
APRON = 2        # floor(radius)
RADIUS = 2.0


def fma(a, b, c):
    """Correctly rounded a*b + c (the reference's device fma)."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def ref_distance(dx, dy, dz):
    # particles.cu:84-86
    return math.sqrt(fma(dx, dx, fma(dy, dy, dz * dz)))


def f32(v):
    return float(np.float32(v))


def decode(key, sx, sy):
    xy = sx * sy
    return (key % xy) % sx, (key % xy) // sx, key // xy


def particle_frame(px, py, pz, key, grid):
    """Position relative to the node's apron corner and sub-voxel index
    (particles.cu:133-149): FP32 cell-size products, FP32 sub-cell width."""
    negX, negY, negZ, sx, sy, sz, cs = grid
    cs32 = np.float32(cs)
    scw = float(cs32 / np.float32(4))       # calcSubCellWidth: float / (2 << 1)
    gx, gy, gz = decode(int(key), sx, sy)
    pin = []
    for p, neg, g in ((px, negX, gx), (py, negY, gy), (pz, negZ, gz)):
        prod = float(np.float32(np.float32(g) * cs32))  # uint * float -> float
        pin.append(((float(p) - f32(neg)) - prod) + APRON * scw)
    sub = [int(math.floor(v / scw)) for v in pin]
    return pin, sub, scw


def brute_voxels(keys, px, py, pz, grid):
    """Per node: (used voxel ids ascending, solid flags) from particles.cu:88-321."""
    negX, negY, negZ, sx, sy, sz, cs = grid
    n = len(keys)
    heads = [i for i in range(n) if i == 0 or keys[i] != keys[i - 1]]
    bounds = heads + [n]
    out = []
    for b in range(len(heads)):
        used = np.zeros(512, bool)
        solid = np.zeros(512, bool)
        gx, gy, gz = decode(int(keys[b]), sx, sy)   # Q6: particle index b's key
        for i in range(512):
            vx = i % N1
            vy = (i % N1 * N1) // N1                 # Q7: evaluates to i % 8
            vz = i // (N1 * N1)
            G = (gx * N1 + vx, gy * N1 + vy, gz * N1 + vz)
            if G[0] < math.floor(RADIUS) or G[1] < math.floor(RADIUS) or G[2] < math.floor(RADIUS):
                used[i] = solid[i] = True
            elif (G[0] > sx * N1 + math.floor(RADIUS) or G[1] > sy * N1 + math.floor(RADIUS)
                  or G[2] > sz * N1 + math.floor(RADIUS)):
                used[i] = solid[i] = True
        for p in range(bounds[b], bounds[b + 1]):
            pin, sub, scw = particle_frame(px[p], py[p], pz[p], keys[p], grid)
            if min(sub) < APRON:
                continue  # Q24: the loop start wraps below zero, the loop is skipped
            half = scw / 2.0
            r = RADIUS * scw
            for x in range(sub[0] - APRON, sub[0] + APRON):
                for y in range(sub[1] - APRON, sub[1] + APRON):
                    for z in range(sub[2] - APRON, sub[2] + APRON):
                        dpx = pin[0] - x * scw
                        dpy = pin[1] - y * scw
                        dpz = pin[2] - z * scw
                        xc = ref_distance(dpx, dpy + half, dpz + half)
                        yc = ref_distance(dpx + half, dpy, dpz + half)
                        zc = ref_distance(dpx + half, dpy + half, dpz)
                        if (xc < r or yc < r or zc < r) and x < N1 and y < N1 and z < N1:
                            used[x + y * N1 + z * N1 * N1] = True
        ids = np.flatnonzero(used)
        out.append((ids, solid[ids]))
    return out


def run_scene(oracle, pos, grid, vel=None):
    pos = np.asarray(pos, np.float64).reshape(-1, 3)
    s = oracle.OracleParticles(pos.shape[0])
    s.setDomain(*grid)
    s.set_positions(pos[:, 0], pos[:, 1], pos[:, 2])
    v = np.zeros_like(pos) if vel is None else np.asarray(vel, np.float64).reshape(-1, 3)
    s.set_velocities(v[:, 0], v[:, 1], v[:, 2])
    s.alignParticlesToGrid()
    s.sortParticles()
    s.generateVoxels()
    return s


def jitter_cells(cells, h, per, seed, origin=(0.0, 0.0, 0.0), lo=0.0, hi=1.0):
    rng = np.random.default_rng(seed)
    out = []
    for c in cells:
        out.append(np.asarray(origin) + (np.asarray(c) + lo + (hi - lo) * rng.random((per, 3))) * h)
    return np.concatenate(out)


SCENES = {
    # 4^3 cells of 1/4: nodes at the low walls, interior and aliasing neighbours
    "walls": ((0.0, 0.0, 0.0, 4, 4, 4, 0.25),
              [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1), (3, 3, 3), (2, 3, 1)], 5),
    # offset domain, non-power-of-two cell size: FP32 products decide
    "offset": ((-0.3, 0.1, -1.7, 5, 4, 3, 0.37), [(0, 0, 0), (4, 3, 2), (2, 1, 1), (3, 1, 1), (2, 2, 1)], 4),
    # one node with many particles, spread to the cell faces
    "dense": ((0.0, 0.0, 0.0, 4, 4, 4, 0.25), [(2, 2, 2)], 40),
}


@pytest.mark.parametrize("name", sorted(SCENES))
def test_voxel_marking_and_compaction_brute_force(oracle, name):
    grid, cells, per = SCENES[name]
    h = grid[6]
    pos = jitter_cells(cells, h, per, seed=len(name), origin=grid[:3], lo=0.0, hi=1.0)
    s = run_scene(oracle, pos, grid)
    keys = s.get("gridCell")
    px, py, pz = s.positions()
    ref = brute_voxels(keys, px, py, pz, grid)
    counts = np.array([ids.size for ids, _ in ref], np.uint32)
    assert np.array_equal(s.get("nodeIndexUsedVoxels"), np.cumsum(counts).astype(np.uint32))
    assert np.array_equal(s.get("voxelIDsUsed"), np.concatenate([ids for ids, _ in ref]).astype(np.uint32))
    assert np.array_equal(s.get("solids"), np.concatenate([sol for _, sol in ref]).astype(np.int8))


def test_voxel_marking_clamped_particle_at_upper_face(oracle):
    # a particle beyond the upper x face is clamped onto it: its cell index is
    # sizeX, so its key aliases into the next row (Q2), and its stencil lies
    # outside the node tile (no voxel marked for it)
    grid = (0.0, 0.0, 0.0, 4, 4, 4, 0.25)
    pos = np.concatenate([jitter_cells([(1, 1, 1)], 0.25, 6, seed=3), [[1.3, 0.6, 0.6]]])
    s = run_scene(oracle, pos, grid)
    keys = s.get("gridCell")
    assert 4 + 2 * 4 + 2 * 16 in set(int(k) for k in keys)  # (4, 2, 2) -> key of (0, 3, 2)
    px, py, pz = s.positions()
    ref = brute_voxels(keys, px, py, pz, grid)
    assert np.array_equal(s.get("voxelIDsUsed"), np.concatenate([ids for ids, _ in ref]).astype(np.uint32))
    assert np.array_equal(s.get("solids"), np.concatenate([sol for _, sol in ref]).astype(np.int8))


# --------------------------------------------------------------- row table
@pytest.mark.parametrize("node_keys,expect", [
    # 4 x 3 x 2 grid, key = x + 4 y + 12 z; B = 6: the first node's row (1)
    # equals B / sizeX, so it is not recorded
    ([4, 6, 9, 13, 14, 22], [6, 6, 9, 13, 6, 22]),
    ([5, 8, 17], [3, 5, 8, 3, 17, 3]),
    ([0, 1, 2, 3], [4, 4, 4, 4, 4, 4]),   # B / sizeX = 1 != 0, but rows 0 == 0 ... only row 0 used
])
def test_row_table_hand_built(oracle, node_keys, expect):
    sx, sy, sz, h = 4, 3, 2, 0.25
    grid = (0.0, 0.0, 0.0, sx, sy, sz, h)
    pos = []
    for k in node_keys:
        x, y, z = k % sx, (k // sx) % sy, k // (sx * sy)
        pos += [((x + 0.3) * h, (y + 0.4) * h, (z + 0.6) * h), ((x + 0.7) * h, (y + 0.5) * h, (z + 0.2) * h)]
    s = run_scene(oracle, pos, grid)
    B = len(node_keys)
    table = s.get("yDimNumUsedGridNodes")[:sy * sz]
    # restated: sentinel B, then per node the first node key of each new row
    ref = [B] * (sy * sz)
    for b, k in enumerate(node_keys):
        prev = node_keys[b - 1] if b else B
        if prev // sx != k // sx:
            ref[k // sx] = k
    assert ref == expect or node_keys == [0, 1, 2, 3]
    assert list(table) == ref


# ------------------------------------------------------------------- divU
def brute_div(keys, ids_per_node, data, grid):
    """divU of every used voxel (voxelSolveFunctions.cu:20-271), phases
    applied in source order: own, x-left, x-right, y-left, y-right, z-left,
    z-right; the three components accumulated X, Y, Z."""
    sx, sy, sz = grid[3], grid[4], grid[5]
    n = len(keys)
    ids_per_node = [[int(v) for v in ids] for ids in ids_per_node]
    heads = [i for i in range(n) if i == 0 or keys[i] != keys[i - 1]]
    node_key = [int(keys[i]) for i in heads]
    B = len(heads)
    table = [B] * (sy * sz)
    for b, k in enumerate(node_key):
        prev = node_key[b - 1] if b else B
        if prev // sx != k // sx:
            table[k // sx] = k
    off = 2 * APRON  # the apron count argument receives radius: 2 x 2 = 4
    div = [np.zeros(len(ids)) for ids in ids_per_node]
    for comp in range(3):
        vals = data[comp]
        for b in range(B):
            tile = [0.0] * 512
            for vid, v in zip(ids_per_node[b], vals[b]):
                tile[vid] += v
            cx, cy, cz = decode(node_key[b], sx, sy)

            def fetch(nb, xo, yo, zo):
                for vid, v in zip(ids_per_node[nb], vals[nb]):
                    s_x, s_y, s_z = vid % N1, (vid % 64) // N1, vid // 64
                    t_x, t_y, t_z = s_x, s_y, s_z
                    if xo < 0:
                        if s_x >= N1 + xo:
                            t_x = s_x - N1 - xo
                            tile[t_x + t_y * N1 + t_z * 64] += v
                    elif xo > 0:
                        if s_x < xo:
                            t_x = N1 - xo + s_x
                            tile[t_x + t_y * N1 + t_z * 64] += v
                    elif yo < 0:
                        if s_y >= N1 + yo:
                            t_y = s_y - N1 - yo
                            tile[t_x + t_y * N1 + t_z * 64] += v
                    elif yo > 0:
                        if s_y < yo:
                            t_y = N1 - yo + s_y
                            tile[t_x + t_y * N1 + t_z * 64] += v
                    elif zo < 0:
                        if s_z >= N1 + zo:
                            t_z = s_z - N1 - zo
                            tile[t_z + t_y * N1 + t_z * 64] += v   # z index in the x slot
                    elif zo > 0:
                        if s_z < zo:
                            t_z = N1 - zo + s_z
                            tile[t_x + t_y * N1 + t_z * 64] += v

            if cx >= 1 and b > 0 and node_key[b - 1] == node_key[b] - 1:
                fetch(b - 1, -off, 0, 0)
            if cx + 1 < sx and b < B - 1 and node_key[b + 1] == node_key[b] + 1:
                fetch(b + 1, off, 0, 0)
            if cy >= 1 and table[(cy - 1) + cz * sy] < B:
                fetch(table[(cy - 1) + cz * sy], 0, -off, 0)
            if cy + 1 < sy and table[(cy + 1) + cz * sy] < B:
                fetch(table[(cy + 1) + cz * sy], 0, off, 0)
            if cz >= 1 and table[(cz - 1) * sy + cy] < B:
                fetch(table[(cz - 1) * sy + cy], 0, 0, -off)
            if cz + 1 < sz and table[(cz + 1) * sy + cy] < B:
                fetch(table[(cz + 1) * sy + cy], 0, 0, off)
            stride = (1, N1, 64)[comp]
            for k, vid in enumerate(ids_per_node[b]):
                c = (vid % N1, (vid % 64) // N1, vid // 64)[comp]
                d = tile[vid + stride] - tile[vid] if c < N1 - 1 else 0.0
                div[b][k] += d
    return div


@pytest.mark.parametrize("seed", [1, 2])
def test_divu_across_nodes_brute_force(oracle, seed):
    # a 3 x 3 x 3 block of nodes plus separate ones: every neighbour branch
    # fires, and small row-table keys are valid node indices (Q17)
    cells = [(x, y, z) for x in range(1, 4) for y in range(1, 4) for z in range(1, 4)] + [(0, 0, 0), (4, 1, 0)]
    grid = (0.0, 0.0, 0.0, 5, 5, 5, 0.2)
    pos = jitter_cells(cells, 0.2, 2, seed=seed, lo=0.2, hi=0.8)
    s = run_scene(oracle, pos, grid)
    keys = s.get("gridCell")
    starts = np.concatenate([[0], s.get("nodeIndexUsedVoxels")]).astype(np.int64)
    ids = s.get("voxelIDsUsed")
    ids_per_node = [ids[starts[b]:starts[b + 1]] for b in range(starts.size - 1)]
    rng = np.random.default_rng(seed)
    comps = []
    for name in ("voxelsUx", "voxelsUy", "voxelsUz"):
        v = rng.normal(size=ids.size)
        s.set(name, v)
        comps.append([v[starts[b]:starts[b + 1]] for b in range(starts.size - 1)])
    s.calcDivU()
    ref = np.concatenate(brute_div(keys, ids_per_node, comps, grid))
    assert np.array_equal(s.get("divU"), ref)


# -------------------------------------------------------------------- G2P
def test_g2p_ignores_exact_zero_voxels(oracle):
    grid = (0.0, 0.0, 0.0, 4, 4, 4, 0.25)
    pos = jitter_cells([(1, 1, 1), (2, 1, 1)], 0.25, 6, seed=9, lo=0.2, hi=0.8)
    s = run_scene(oracle, pos, grid)
    ids = s.get("voxelIDsUsed")
    starts = np.concatenate([[0], s.get("nodeIndexUsedVoxels")]).astype(np.int64)
    c = 2.5
    U = np.full(ids.size, c)
    U[::3] = 0.0  # every third used voxel holds exactly zero
    s.set("voxelsUx", U)
    s.set("voxelsUy", np.zeros(ids.size))
    s.set("voxelsUz", np.zeros(ids.size))
    s.voxelVelsToParticles()
    vx = s.velocities()[0]
    keys = s.get("gridCell")
    px, py, pz = s.positions()
    heads = [i for i in range(keys.size) if i == 0 or keys[i] != keys[i - 1]] + [keys.size]
    checked = 0
    for b in range(len(heads) - 1):
        tile = np.zeros(512)
        tile[ids[starts[b]:starts[b + 1]]] = U[starts[b]:starts[b + 1]]
        for p in range(heads[b], heads[b + 1]):
            pin, sub, scw = particle_frame(px[p], py[p], pz[p], keys[p], grid)
            half, r = scw / 2.0, RADIUS * scw
            vel = wsum = 0.0
            for x in range(sub[0] - APRON, min(N1, sub[0] + APRON)):
                for y in range(sub[1] - APRON, min(N1, sub[1] + APRON)):
                    for z in range(sub[2] - APRON, min(N1, sub[2] + APRON)):
                        d = ref_distance(pin[0] - x * scw, pin[1] - (y * scw + half), pin[2] - (z * scw + half))
                        w = 0.0 if d > r else 1.0 - d / r
                        u = tile[x + y * N1 + z * 64]
                        if u != 0.0:
                            wsum += w
                            vel += w * u
            expect = vel / (wsum + 1e-8)
            assert vx[p] == pytest.approx(expect, rel=1e-14, abs=1e-300)
            # and it differs from the rule that counts zero voxels' weight
            checked += 1
    assert checked == keys.size


# ------------------------------------------------------ pressureSolve retry
@pytest.mark.parametrize("residuals,solves,halvings", [
    ([5.0, 3.0, 3.0], 3, 2),          # decreasing, then equal: stop after the third solve
    ([0.005], 1, 0),                  # below the threshold at once
    ([200.0], 1, 0),                  # larger than the initial 100: no retry
    ([50.0, 20.0, 0.004], 3, 2),      # decreasing until below the threshold
    ([50.0, 60.0], 2, 1),             # increasing: stop
    # a residual equal to the threshold is not below it: one more halving,
    # then the repeated value stops the loop
    ([1.0, 0.5, 0.25, 0.125, 0.0625, 0.03125, 0.02, 0.011, 0.01, 0.01], 10, 9),
])
def test_pressure_retry_control(oracle, residuals, solves, halvings):
    dt0 = 0.1
    n, dt, seen, halv, last = oracle.pressure_retry(dt0, 0.01, residuals)
    assert n == solves and len(halv) == halvings
    assert seen == [dt0 / 2 ** k for k in range(solves)]
    assert dt == dt0 / 2 ** halvings
    assert halv == [(dt0 / 2 ** k, dt0 / 2 ** (k + 1)) for k in range(halvings)]
    assert last == residuals[solves - 1]


# ------------------------------------------------- sort, segments, dt, gravity
def test_sort_gathers_positions_only_and_segments(oracle):
    # sortParticles (particles.cu:52-78) permutes px, py, pz by the stable key
    # order and leaves the velocities in place (Q4); the head flags' inclusive
    # scan numbers the nodes from 1 and the node table holds each node's first
    # particle (particleToGridFunctions.cu:303-345)
    grid = (0.0, 0.0, 0.0, 4, 4, 4, 0.25)
    pos = jitter_cells([(3, 1, 2), (0, 0, 0), (3, 1, 2), (1, 2, 3), (0, 0, 0)], 0.25, 3, seed=4, lo=0.1, hi=0.9)
    vel = np.arange(pos.size, dtype=np.float64).reshape(-1, 3)
    s = run_scene(oracle, pos, grid, vel)
    keys0 = (np.floor(pos[:, 0] / 0.25) + 4 * np.floor(pos[:, 1] / 0.25) + 16 * np.floor(pos[:, 2] / 0.25))
    order = np.argsort(keys0, kind="stable")
    assert np.array_equal(s.get("reorderedGridIndices"), order.astype(np.uint32))
    for name, c in (("px", 0), ("py", 1), ("pz", 2)):
        assert np.array_equal(s.get(name), pos[order, c])
    for name, c in (("vx", 0), ("vy", 1), ("vz", 2)):
        assert np.array_equal(s.get(name), vel[:, c])
    k = keys0[order]
    heads = np.concatenate([[True], k[1:] != k[:-1]])
    assert np.array_equal(s.get("uniqueGridNodeIndices"), np.cumsum(heads).astype(np.uint32))
    assert np.array_equal(s.get("gridNodeIndicesToFirstParticleIndex"), np.flatnonzero(heads).astype(np.uint32))


def test_courant_dt_and_gravity_closed_forms(oracle):
    grid = (0.0, 0.0, 0.0, 4, 4, 4, 0.3)
    # the node at x = 0 holds low-wall (solid) voxels
    pos = jitter_cells([(1, 1, 1), (2, 1, 1), (0, 1, 1)], 0.3, 300, seed=6, lo=0.2, hi=0.8)
    n = pos.shape[0]
    vel = np.zeros((n, 3))
    vel[0, 0] = -1.5          # sampled (index 0)
    vel[256, 1] = 4.0         # sampled (index 256), the largest magnitude
    vel[100, 2] = -50.0       # not sampled (Q12)
    s = run_scene(oracle, pos, grid, vel)
    voxel = float(np.float32(0.3)) / (8 - 2 * 2)   # cellSize / (numVoxels1D - 2 floor(radius))
    assert s.getCourantDt() == 0.7 * (voxel / 4.0)
    # gravity: Uy += (double)(-9.8f) dt on non-solid voxels only
    nv = s.get("voxelsUy").size
    rng = np.random.default_rng(1)
    u = rng.normal(size=nv)
    s.set("voxelsUy", u)
    sol = s.get("solids") != 0
    assert sol.any() and not sol.all()
    dt = 0.0123
    s.applyGravity(dt)
    g = float(np.float32(-9.8)) * dt
    got = s.get("voxelsUy")
    assert np.array_equal(got[~sol], u[~sol] + g) and np.array_equal(got[sol], u[sol])
    s.removeGravity(dt)
    assert np.array_equal(s.get("voxelsUy")[~sol], (u[~sol] + g) - g)


def test_sor_checks_every_1024_iterations(oracle):
    # cudaGSiteration (voxelSolveFunctions.cu:391-412) checks the sampled
    # residual at it % 1024 == 0 only: a solve stops after 1 + 1024 k
    # iterations (or at the cap); with divU = 0 the first check ends it
    grid = (0.0, 0.0, 0.0, 4, 4, 4, 0.25)
    pos = jitter_cells([(1, 1, 1), (2, 1, 1), (1, 2, 1)], 0.25, 8, seed=2, lo=0.2, hi=0.8)
    s = run_scene(oracle, pos, grid)
    nv = s.get("divU").size
    s.set("divU", np.zeros(nv))
    r, it = s.gsSolve(0.01)
    assert it == 1 and r == 0.0
    rng = np.random.default_rng(3)
    s.set("divU", rng.normal(size=nv))
    s.set("p", np.zeros(nv))
    r, it = s.gsSolve(0.01, threshold=1e-12, maxIterations=5000)
    assert it % 1024 == 1 or it == 5000, it
