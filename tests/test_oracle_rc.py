"""Closed-form pins of the FP64 oracle's reference-compatible stages.

The reference holds no golden vectors for these stages, so each is checked
against what its definition fixes on a hand-built input:
  - P2G (particles.cu:379-488): a particle on a face point gets weight 1 there
    and the hat weights 1 - d/(2 h_sub) at the neighbouring face points;
  - G2P (particles.cu:562-662): a constant field comes back as the constant
    up to the 1e-8 regulariser;
  - divU (voxelSolveFunctions.cu:188-243): a linear field inside one node has
    the constant forward-difference divergence;
  - SOR (voxelSolveFunctions.cu:293-389): the converged red/black iteration is
    the solution of its (dense-solved) linear system;
  - velocity update (voxelSolveFunctions.cu:415-486) of a linear pressure and
    Euler advection (particles.cu:664-669) have closed forms.
"""
import numpy as np
import pytest

H = 0.25          # cell size (4^3 cells on [0,1)^3)
SCW = H / 4       # sub-voxel width: 8 voxels per node minus a 2-voxel apron
HS = SCW / 2
N1 = 8


def voxel_xyz(ids):
    ids = np.asarray(ids, np.int64)
    return ids % N1, (ids // N1) % N1, ids // (N1 * N1)


def setup(oracle, pos, vel=None):
    pos = np.asarray(pos, np.float64).reshape(-1, 3)
    n = pos.shape[0]
    s = oracle.OracleParticles(n)
    s.setDomain(0.0, 0.0, 0.0, 4, 4, 4, H)
    s.set_positions(pos[:, 0], pos[:, 1], pos[:, 2])
    v = np.zeros((n, 3)) if vel is None else np.asarray(vel, np.float64).reshape(-1, 3)
    s.set_velocities(v[:, 0], v[:, 1], v[:, 2])
    s.alignParticlesToGrid()
    s.sortParticles()
    s.generateVoxels()
    return s


def test_p2g_single_particle_hat_weights(oracle):
    # cell (1,1,1); node-local frame p_in = p - cell origin + 2 h_sub.
    # Put the particle on the x-face point of voxel (3,3,3): (3 h, 3 h + h/2, 3 h + h/2)
    pin = np.array([3 * SCW, 3 * SCW + HS, 3 * SCW + HS])
    pos = np.array([H, H, H]) + pin - 2 * SCW
    v = np.array([2.0, -3.0, 5.0])
    s = setup(oracle, pos, v)
    s.particleVelToVoxels()
    ids = s.get("voxelIDsUsed")
    Ux, Uy, Uz = s.get("voxelsUx"), s.get("voxelsUy"), s.get("voxelsUz")
    at = {int(i): k for k, i in enumerate(ids)}
    vid = lambda x, y, z: x + N1 * y + N1 * N1 * z
    reg = 1e-7

    def expect(w, vc):
        return w * vc / (w + reg)

    # x-face of (3,3,3): d = 0 -> w = 1
    assert Ux[at[vid(3, 3, 3)]] == pytest.approx(expect(1.0, v[0]), rel=1e-15)
    # x-faces of (2,3,3) and (4,3,3): d = h_sub -> w = 1/2
    for x in (2, 4):
        assert Ux[at[vid(x, 3, 3)]] == pytest.approx(expect(0.5, v[0]), rel=1e-15)
    # x-face of (4,4,3): d = sqrt(2) h_sub
    assert Ux[at[vid(4, 4, 3)]] == pytest.approx(expect(1 - np.sqrt(2) / 2, v[0]), rel=1e-14)
    # y-face of (3,3,3) sits at (3h + h/2, 3h, 3h + h/2): d = h/sqrt(2)
    assert Uy[at[vid(3, 3, 3)]] == pytest.approx(expect(1 - 1 / (2 * np.sqrt(2)), v[1]), rel=1e-14)
    # z-face of (3,3,3) at (3h + h/2, 3h + h/2, 3h): same distance
    assert Uz[at[vid(3, 3, 3)]] == pytest.approx(expect(1 - 1 / (2 * np.sqrt(2)), v[2]), rel=1e-14)
    # x-face of (3,3,1) is 2 h_sub away in z: weight 0 (and the voxel is not used)
    assert vid(3, 3, 1) not in at or Ux[at[vid(3, 3, 1)]] == 0.0


def block_positions(seed, cells=((1, 1, 1),), per_cell=8):
    rng = np.random.default_rng(seed)
    out = []
    for c in cells:
        out.append((np.asarray(c) + 0.15 + 0.7 * rng.random((per_cell, 3))) * H)
    return np.concatenate(out)


def test_g2p_constant_field(oracle):
    pos = block_positions(1, cells=((1, 1, 1), (2, 1, 1), (1, 2, 1)))
    s = setup(oracle, pos)
    c = np.array([1.5, -0.25, 4.0])
    nv = s.get("voxelsUx").size
    s.set("voxelsUx", np.full(nv, c[0]))
    s.set("voxelsUy", np.full(nv, c[1]))
    s.set("voxelsUz", np.full(nv, c[2]))
    s.voxelVelsToParticles()
    v = np.stack(s.velocities(), axis=1)
    # v = c W / (W + 1e-8) with W >= 1/2 (the nearest face point is within h_sub)
    assert np.all(np.abs(v / c - 1.0) <= 2e-8)
    assert np.all(np.abs(v) <= np.abs(c))


def test_divu_linear_field_in_one_node(oracle):
    pos = block_positions(2, per_cell=12)
    s = setup(oracle, pos)
    ids = s.get("voxelIDsUsed")
    x, y, z = voxel_xyz(ids)
    a, b, c = 3.0, 5.0, 7.0
    s.set("voxelsUx", a * x)
    s.set("voxelsUy", b * y)
    s.set("voxelsUz", c * z)
    s.calcDivU()
    div = s.get("divU")
    used = set(int(i) for i in ids)
    ok = [k for k, i in enumerate(ids)
          if x[k] < N1 - 1 and y[k] < N1 - 1 and z[k] < N1 - 1
          and int(i) + 1 in used and int(i) + N1 in used and int(i) + N1 * N1 in used]
    assert len(ok) >= 20
    assert np.all(div[ok] == a + b + c)
    # a voxel whose +x neighbour is unused sees u(i+x) = 0
    edge = [k for k, i in enumerate(ids) if x[k] < N1 - 1 and int(i) + 1 not in used
            and int(i) + N1 in used and int(i) + N1 * N1 in used and y[k] < N1 - 1 and z[k] < N1 - 1]
    assert edge
    for k in edge:
        assert div[k] == -a * x[k] + b + c


def sor_system(ids, solids, divu, scale):
    """Dense matrix of the SOR fixed point: updated voxels are non-solid and
    off the apron layer (x or z on the tile border; the reference's y index
    equals x); off-diagonals -scale to non-solid neighbours with divU != 0;
    diagonal = scale x (non-solid neighbours) + 1e-9."""
    tile_div = np.zeros(N1 ** 3)
    tile_sol = np.zeros(N1 ** 3, bool)
    tile_div[ids] = divu
    tile_sol[ids] = solids != 0
    x, _, z = voxel_xyz(ids)
    upd = [int(i) for k, i in enumerate(ids)
           if not tile_sol[i] and 1 <= x[k] < N1 - 1 and 1 <= z[k] < N1 - 1]
    pos = {v: r for r, v in enumerate(upd)}
    M = np.zeros((len(upd), len(upd)))
    for r, v in enumerate(upd):
        d = 0.0
        for o in (-1, 1, -N1, N1, -N1 * N1, N1 * N1):
            nb = v + o
            if not tile_sol[nb]:
                d += scale
                if tile_div[nb] != 0.0 and nb in pos:
                    M[r, pos[nb]] = -scale
        M[r, r] = d + 1e-9
    return upd, M, tile_div


def test_sor_converges_to_dense_solve(oracle):
    pos = block_positions(3, per_cell=16)
    s = setup(oracle, pos)
    ids = s.get("voxelIDsUsed")
    rng = np.random.default_rng(4)
    divu = rng.uniform(0.5, 1.5, ids.size) * rng.choice([-1.0, 1.0], ids.size)
    s.set("divU", divu)
    s.set("p", np.zeros(ids.size))
    dt = 1e-3
    scale = dt / (0.014 * SCW * SCW)
    s.gsIterations(dt, 3000)
    p = s.get("p")
    upd, M, tile_div = sor_system(ids, s.get("solids"), divu, scale)
    assert len(upd) >= 20 and np.count_nonzero(M - np.diag(np.diag(M))) >= 40
    # apron and solid voxels keep p = 0, so they drop out of the system
    want = np.linalg.solve(M, -tile_div[upd])
    at = {int(i): k for k, i in enumerate(ids)}
    got = np.array([p[at[v]] for v in upd])
    assert np.max(np.abs(got - want)) <= 1e-10 * np.max(np.abs(want))
    fixed = [k for k, i in enumerate(ids) if int(i) not in set(upd)]
    assert np.all(p[fixed] == 0.0)


def test_velocity_update_linear_pressure(oracle):
    pos = block_positions(5, per_cell=12)
    s = setup(oracle, pos)
    ids = s.get("voxelIDsUsed")
    x, y, z = voxel_xyz(ids)
    alpha = 0.5
    s.set("p", alpha * x.astype(np.float64))
    s.set("voxelsUx", np.zeros(ids.size))
    s.set("voxelsUy", np.zeros(ids.size))
    s.set("voxelsUz", np.zeros(ids.size))
    dt = 2e-3
    s.setTime(dt, 1.0 / 24)
    s.updateVoxelVelocities()
    Ux, Uy, Uz = s.get("voxelsUx"), s.get("voxelsUy"), s.get("voxelsUz")
    scale = dt / (0.014 * SCW * SCW)
    used = set(int(i) for i in ids)
    solids = s.get("solids")
    inner = [k for k, i in enumerate(ids) if 1 <= x[k] < N1 - 1 and 1 <= z[k] < N1 - 1 and 1 <= y[k] < N1 - 1
             and solids[k] == 0 and int(i) - 1 in used and int(i) - N1 in used and int(i) - N1 * N1 in used]
    assert len(inner) >= 10
    # u -= scale (p(v) - p(v - e)): only the x component sees a gradient
    assert np.allclose(Ux[inner], -scale * alpha, rtol=1e-14)
    assert np.all(Uy[inner] == 0.0) and np.all(Uz[inner] == 0.0)


def test_euler_advection(oracle):
    pos = block_positions(6, per_cell=5)
    rng = np.random.default_rng(7)
    vel = rng.normal(size=pos.shape)
    s = setup(oracle, pos, vel)
    p0 = np.stack(s.positions(), axis=1)
    v0 = np.stack(s.velocities(), axis=1)
    dt = 1.0 / 480
    s.setTime(dt, 1.0 / 24)
    s.advectParticles()
    p1 = np.stack(s.positions(), axis=1)
    assert np.array_equal(p1, p0 + dt * v0)
