"""Independent pins of the FP64 oracle's north-star method (oracle/ns_oracle.c).

The method has no reference counterpart, so each function is checked
against what its definition fixes, on inputs chosen so that a plausible
slip (a dropped half-cell offset, a transposed axis, a wrong sign, a face
of the wrong cell) fails:
  - P2G: a uniform velocity is reproduced exactly on every face with weight;
    one particle's weights are the products of 1D hats on each component's
    own staggered lattice and sum to 1 away from the walls; a particle on a
    face point puts weight 1 there and nowhere else;
  - walls + gravity: closed form (domain-wall and solid faces 0, v += g dt
    elsewhere, the saved field without gravity);
  - assembly, operator and Jacobi-PCG: the operator equals a scipy sparse
    7-point Laplacian built from the cell types alone, the right-hand side
    is -(h/dt) times a numpy MAC divergence, and the PCG solution equals
    scipy's sparse direct solve;
  - pressure gradient: a pressure linear in x changes the u faces between
    fluid cells by the constant -(dt/h) dp and leaves air-air faces;
  - G2P: trilinear interpolation reproduces a linear field exactly; FLIP 0
    is PIC, FLIP 1 adds the interpolated change;
  - RK2: a uniform field moves particles by dt v; the stretching field
    u = a x gives the midpoint-rule closed form x (1 + a dt + (a dt)^2 / 2),
    not Euler's x (1 + a dt);
  - the whole serial step equals its stages run one by one.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from flip2_amd import scenes


def grid_of(n, h=None, origin=(0.0, 0.0, 0.0)):
    return (origin[0], origin[1], origin[2], n, n, n, h if h is not None else 1.0 / n)


def ijk(n):
    """Cell coordinates [i, j, k] arrays (indexing i fastest in memory order)."""
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return i.ravel(), j.ravel(), k.ravel()


def to_dense(oracle, a, grid):
    """Brick-tiled array -> dense [k, j, i]."""
    n = grid[3]
    i, j, k = ijk(n)
    return np.asarray(a)[oracle.ns_tiled_index(i, j, k, grid)].reshape(n, n, n)


def to_tiled(oracle, d, grid):
    n = grid[3]
    i, j, k = ijk(n)
    out = np.zeros(n ** 3, dtype=np.asarray(d).dtype)
    out[oracle.ns_tiled_index(i, j, k, grid)] = np.asarray(d).reshape(-1)
    return out


# --------------------------------------------------------------------------- P2G
def test_p2g_uniform_velocity_reproduced(oracle):
    n = 16
    grid = grid_of(n)
    x, y, z = scenes.uniform_random(3000, seed=11, lo=0.1, hi=0.9)
    v = (np.full_like(x, 1.25), np.full_like(x, -0.5), np.full_like(x, 2.0))
    U, V, W, WU, WV, WW = oracle.ns_p2g(x, y, z, *v, grid, weights=True)
    for F, Wt, c in ((U, WU, 1.25), (V, WV, -0.5), (W, WW, 2.0)):
        assert np.all(Wt >= 0.0)
        assert np.any(Wt > 0.0)
        assert np.allclose(F[Wt > 0.0], c, rtol=0, atol=1e-14)
        assert np.all(F[Wt == 0.0] == 0.0)


def hat1(q, node):
    return np.maximum(0.0, 1.0 - np.abs(q - node))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_p2g_single_particle_weights_are_staggered_hats(oracle, seed):
    n = 16
    grid = grid_of(n)
    h = 1.0 / n
    rng = np.random.default_rng(seed)
    p = rng.uniform(0.3, 0.7, 3).astype(np.float32)
    q = p.astype(np.float64) / h
    _, _, _, WU, WV, WW = oracle.ns_p2g(p[:1], p[1:2], p[2:], *(np.ones(1, np.float32),) * 3, grid, weights=True)
    i, j, k = (a.reshape(n, n, n) for a in ijk(n))  # [k, j, i]
    for comp, Wt in enumerate((WU, WV, WW)):
        # component c sits at (i, j, k) + 1/2 on the two other axes
        node = [i + (0.0 if comp == 0 else 0.5), j + (0.0 if comp == 1 else 0.5), k + (0.0 if comp == 2 else 0.5)]
        expect = hat1(q[0], node[0]) * hat1(q[1], node[1]) * hat1(q[2], node[2])
        got = to_dense(oracle, Wt, grid)
        assert np.allclose(got, expect, rtol=0, atol=1e-14), comp
        assert abs(got.sum() - 1.0) <= 1e-14  # partition of unity away from the walls
        assert np.count_nonzero(got) <= 8


def test_p2g_particle_on_face_point(oracle):
    n = 8
    h = 1.0 / n
    grid = grid_of(n)
    # the u-face point of cell (3, 4, 5): (3 h, 4.5 h, 5.5 h)
    x, y, z = (np.array([v * h], np.float32) for v in (3.0, 4.5, 5.5))
    _, _, _, WU, _, _ = oracle.ns_p2g(x, y, z, *(np.ones(1, np.float32),) * 3, grid, weights=True)
    d = to_dense(oracle, WU, grid)
    assert d[5, 4, 3] == 1.0
    assert np.count_nonzero(d) == 1


def test_p2g_weights_drop_outside_faces(oracle):
    # next to the low x wall: the u lattice (not offset in x) keeps both x
    # nodes, the v lattice (offset by half a cell) loses the node at i = -1,
    # so the particle's v weights on the grid sum to q_x + 1/2
    n = 8
    grid = grid_of(n)
    x, y, z = (np.array([v], np.float32) for v in (0.02, 0.5, 0.5))
    WU, WV, WW = oracle.ns_p2g(x, y, z, *(np.ones(1, np.float32),) * 3, grid, weights=True)[3:]
    qx = float(x[0]) * n
    assert abs(WU.sum() - 1.0) <= 1e-14
    assert abs(WV.sum() - (qx + 0.5)) <= 1e-14
    assert abs(WW.sum() - (qx + 0.5)) <= 1e-14


# ---------------------------------------------------------------- walls, gravity
def test_prepare_faces_closed_form(oracle):
    n = 8
    grid = grid_of(n)
    rng = np.random.default_rng(4)
    U, V, W = (rng.normal(size=n ** 3) for _ in range(3))
    solid = np.zeros((n, n, n), np.uint8)  # [k, j, i]
    solid[2:4, 3, 5] = 1
    st = to_tiled(oracle, solid, grid)
    gdt = -0.25
    F, F0 = oracle.ns_prepare_faces(U, V, W, grid, gdt, st)
    i, j, k = (a.reshape(n, n, n) for a in ijk(n))
    sd = solid.astype(bool)
    for a, (Fa, F0a, orig) in enumerate(zip(F, F0, (U, V, W))):
        prev = np.zeros_like(sd)  # the neighbour on the low side of the face
        coord = (i, j, k)[a]
        if a == 0:
            prev[:, :, 1:] = sd[:, :, :-1]
        elif a == 1:
            prev[:, 1:, :] = sd[:, :-1, :]
        else:
            prev[1:, :, :] = sd[:-1, :, :]
        wall = sd | prev | (coord == 0)
        o = to_dense(oracle, orig, grid)
        exp0 = np.where(wall, 0.0, o)
        exp = exp0 + (gdt if a == 1 else 0.0) * (~wall)
        assert np.array_equal(to_dense(oracle, F0a, grid), exp0)
        assert np.allclose(to_dense(oracle, Fa, grid), exp, rtol=0, atol=1e-15)


# ---------------------------------------------------- assembly, operator, PCG
def mixed_scene(n, seed):
    """Counts and solids on an n^3 grid: fluid below a bumpy surface, a solid
    block, air above (dense [k, j, i])."""
    rng = np.random.default_rng(seed)
    counts = np.zeros((n, n, n), np.uint32)
    height = rng.integers(n // 3, n - 2, size=(n, n))  # per (k, i)
    for kk in range(n):
        for ii in range(n):
            counts[kk, :height[kk, ii], ii] = rng.integers(1, 9, size=height[kk, ii])
    solid = np.zeros((n, n, n), np.uint8)
    solid[n // 4:n // 2, 1:3, n // 3:n // 3 + 3] = 1
    counts[solid.astype(bool)] = 0
    counts[n // 2, n // 5, 2] = 0  # an air bubble
    return counts, solid


def reference_operator(counts, solid):
    """7-point operator over fluid cells from the cell types alone: diagonal =
    number of non-solid neighbours inside the domain, -1 to fluid neighbours."""
    n = counts.shape[0]
    fluid = (counts > 0) & (solid == 0)
    idx = -np.ones((n, n, n), np.int64)
    cells = np.argwhere(fluid)
    idx[tuple(cells.T)] = np.arange(len(cells))
    rows, cols, vals = [], [], []
    for r, (kk, jj, ii) in enumerate(cells):
        diag = 0
        for dk, dj, di in ((0, 0, 1), (0, 0, -1), (0, 1, 0), (0, -1, 0), (1, 0, 0), (-1, 0, 0)):
            a, b, c = kk + dk, jj + dj, ii + di
            if not (0 <= a < n and 0 <= b < n and 0 <= c < n) or solid[a, b, c]:
                continue
            diag += 1
            if fluid[a, b, c]:
                rows.append(r); cols.append(idx[a, b, c]); vals.append(-1.0)
        rows.append(r); cols.append(r); vals.append(float(diag))
    A = sp.csr_matrix((vals, (rows, cols)), shape=(len(cells), len(cells)))
    return A, cells


def mac_divergence(U, V, W):
    """D = u(i+1)-u(i) + v(j+1)-v(j) + w(k+1)-w(k) on dense [k, j, i] faces
    (faces beyond the high walls are 0)."""
    D = -U - V - W
    D[:, :, :-1] += U[:, :, 1:]
    D[:, :-1, :] += V[:, 1:, :]
    D[:-1, :, :] += W[1:, :, :]
    return D


@pytest.mark.parametrize("n,seed", [(8, 1), (16, 2)])
def test_assemble_apply_and_pcg_match_scipy(oracle, n, seed):
    grid = grid_of(n)
    h, dt = grid[6], 1.0 / 240.0
    counts, solid = mixed_scene(n, seed)
    rng = np.random.default_rng(seed + 10)
    Ud, Vd, Wd = (rng.normal(size=(n, n, n)) for _ in range(3))
    ct = np.append(to_tiled(oracle, counts, grid), np.uint32(0))
    st = to_tiled(oracle, solid, grid)
    ty, coef, b = oracle.ns_assemble(ct, grid, h / dt, *(to_tiled(oracle, a, grid) for a in (Ud, Vd, Wd)), solid=st)
    fluid = (counts > 0) & (solid == 0)
    td = to_dense(oracle, ty, grid)
    assert np.array_equal(td, np.where(solid > 0, 2, np.where(counts > 0, 1, 0)))
    A, cells = reference_operator(counts, solid)
    sel = tuple(cells.T)
    bd = to_dense(oracle, b, grid)
    assert np.allclose(bd[sel], -(h / dt) * mac_divergence(Ud, Vd, Wd)[sel], rtol=1e-13, atol=1e-9)
    assert np.all(bd[~fluid] == 0.0)
    # operator: y = A x for a random x
    xr = rng.normal(size=(n, n, n))
    y = to_dense(oracle, oracle.ns_apply_A(coef, grid, to_tiled(oracle, xr, grid)), grid)
    assert np.allclose(y[sel], A @ xr[sel], rtol=0, atol=1e-12)
    assert np.all(y[~fluid] == 0.0)
    # PCG to 1e-12 equals the sparse direct solve
    x, it, rel = oracle.ns_pcg_jacobi(coef, grid, b, tol=1e-12)
    assert rel <= 1e-12 and it > 0
    xs = spla.spsolve(A.tocsc(), bd[sel])
    xd = to_dense(oracle, x, grid)
    assert np.max(np.abs(xd[sel] - xs)) <= 1e-9 * np.max(np.abs(xs))


# ----------------------------------------------------------- pressure gradient
def test_velocity_update_linear_pressure(oracle):
    n = 8
    grid = grid_of(n)
    h, dt = grid[6], 1.0 / 240.0
    counts = np.zeros((n, n, n), np.uint32)
    counts[:, :4, :] = 1  # fluid in the lower half
    ct = np.append(to_tiled(oracle, counts, grid), np.uint32(0))
    zero = np.zeros(n ** 3)
    ty, _, _ = oracle.ns_assemble(ct, grid, h / dt, zero, zero, zero)
    i, j, k = (a.reshape(n, n, n) for a in ijk(n))
    c = 3.0
    p = np.where(counts > 0, c * i + 7.0 * k, 0.0)
    rng = np.random.default_rng(5)
    U, V, W = (rng.normal(size=(n, n, n)) for _ in range(3))
    Fu, Fv, Fw = oracle.ns_velocity_update(ty, grid, to_tiled(oracle, p, grid), dt / h,
                                           *(to_tiled(oracle, a, grid) for a in (U, V, W)))
    Fu, Fv, Fw = (to_dense(oracle, a, grid) for a in (Fu, Fv, Fw))
    fl = counts > 0
    # u faces between two fluid cells: constant change -(dt/h) c
    inner = fl.copy()
    inner[:, :, 0] = False
    assert np.allclose(Fu[inner] - U[inner], -(dt / h) * c, rtol=0, atol=1e-12)
    # w faces between fluid cells: -(dt/h) 7
    wz = fl.copy()
    wz[0] = False
    assert np.allclose(Fw[wz] - W[wz], -(dt / h) * 7.0, rtol=0, atol=1e-12)
    # v faces at the free surface (fluid cell j = 3 below, air j = 4 above):
    # p_air = 0, so the change is -(dt/h)(0 - p[j = 3])
    assert np.allclose((Fv - V)[:, 4, :], (dt / h) * p[:, 3, :], rtol=0, atol=1e-12)
    # v faces inside the fluid: p does not vary with j
    assert np.allclose((Fv - V)[:, 1:4, :], 0.0, rtol=0, atol=1e-12)
    # faces between two air cells and the domain walls are untouched
    air = (j >= 5)
    assert np.array_equal(Fv[air], V[air]) and np.array_equal(Fu[~fl], U[~fl])
    assert np.array_equal(Fu[:, :, 0], U[:, :, 0])


# ---------------------------------------------------------------------- G2P
def linear_field(oracle, grid, coef):
    """MAC faces of the linear velocity field v_c(x) = a_c + b_c . x, sampled
    at each component's own face points."""
    n, h = grid[3], grid[6]
    i, j, k = (a.reshape(n, n, n) for a in ijk(n))
    out = []
    for comp in range(3):
        pos = [(i + (0.0 if comp == 0 else 0.5)) * h, (j + (0.0 if comp == 1 else 0.5)) * h,
               (k + (0.0 if comp == 2 else 0.5)) * h]
        a, bvec = coef[comp]
        out.append(to_tiled(oracle, a + bvec[0] * pos[0] + bvec[1] * pos[1] + bvec[2] * pos[2], grid))
    return out


def test_g2p_reproduces_linear_field_and_flip_blend(oracle):
    n = 16
    grid = grid_of(n)
    coef = [(0.3, (1.0, -2.0, 0.5)), (-1.0, (0.25, 0.75, -1.5)), (2.0, (-0.5, 0.0, 1.0))]
    F = linear_field(oracle, grid, coef)
    F0 = [0.5 * f for f in F]
    x, y, z = scenes.uniform_random(500, seed=3, lo=0.2, hi=0.8)
    v0 = [np.full_like(x, 0.125), np.full_like(x, -0.25), np.full_like(x, 1.0)]
    pos = [a.astype(np.float64) for a in (x, y, z)]
    exact = [a + b[0] * pos[0] + b[1] * pos[1] + b[2] * pos[2] for a, b in coef]
    pic = oracle.ns_g2p_flip(x, y, z, *v0, *F, *F0, grid, 0.0)
    for got, ex in zip(pic, exact):
        assert np.allclose(got, ex, rtol=0, atol=1e-12)
    flip1 = oracle.ns_g2p_flip(x, y, z, *v0, *F, *F0, grid, 1.0)
    for got, ex, v in zip(flip1, exact, v0):
        assert np.allclose(got, v + 0.5 * ex, rtol=0, atol=1e-12)
    blend = oracle.ns_g2p_flip(x, y, z, *v0, *F, *F0, grid, 0.95)
    for got, ex, v in zip(blend, exact, v0):
        assert np.allclose(got, 0.95 * (v + 0.5 * ex) + 0.05 * ex, rtol=0, atol=1e-12)


# ---------------------------------------------------------------------- RK2
def test_rk2_uniform_and_stretching_fields(oracle):
    n = 16
    grid = grid_of(n)
    dt = 0.01
    x, y, z = scenes.uniform_random(400, seed=5, lo=0.3, hi=0.6)
    F = linear_field(oracle, grid, [(0.5, (0, 0, 0)), (-1.0, (0, 0, 0)), (0.25, (0, 0, 0))])
    xo = oracle.ns_advect_rk2(x, y, z, *F, grid, dt, 0.0)
    for got, p0, v in zip(xo, (x, y, z), (0.5, -1.0, 0.25)):
        assert np.allclose(got, p0.astype(np.float64) + dt * v, rtol=0, atol=1e-14)
    a = 2.0
    F = linear_field(oracle, grid, [(0.0, (a, 0, 0)), (0.0, (0, 0, 0)), (0.0, (0, 0, 0))])
    xo = oracle.ns_advect_rk2(x, y, z, *F, grid, dt, 0.0)
    x0 = x.astype(np.float64)
    assert np.allclose(xo[0], x0 * (1.0 + a * dt + 0.5 * (a * dt) ** 2), rtol=0, atol=1e-14)
    assert not np.allclose(xo[0], x0 * (1.0 + a * dt), rtol=0, atol=1e-9)  # not forward Euler
    assert np.array_equal(xo[1], y.astype(np.float64))


def test_rk2_clamps_to_margin_and_keeps_particles_out_of_solids(oracle):
    n = 8
    grid = grid_of(n)
    h = grid[6]
    # u = -1 everywhere: the midpoint (x = 0) is still inside the grid, the
    # end point (x = -0.05) is clamped to the margin
    F = linear_field(oracle, grid, [(-1.0, (0, 0, 0)), (0.0, (0, 0, 0)), (0.0, (0, 0, 0))])
    x, y, z = (np.array([v], np.float32) for v in (0.05, 0.5, 0.5))
    xo = oracle.ns_advect_rk2(x, y, z, *F, grid, 0.1, 1e-3 * h)
    assert xo[0][0] == pytest.approx(1e-3 * h, abs=1e-15)
    F = linear_field(oracle, grid, [(-10.0, (0, 0, 0)), (0.0, (0, 0, 0)), (0.0, (0, 0, 0))])
    solid = np.zeros((n, n, n), np.uint8)
    solid[4, 4, 3] = 1  # the cell the particle would move into
    x, y, z = (np.array([v], np.float32) for v in (0.52, 0.55, 0.55))
    xo = oracle.ns_advect_rk2(x, y, z, *F, grid, 0.01, 0.0, to_tiled(oracle, solid, grid))
    assert xo[0][0] == pytest.approx(float(x[0]), abs=0) and xo[1][0] == float(y[0])


# ------------------------------------------------------- whole serial step
def test_serial_step_equals_its_stages(oracle):
    n = 16
    grid = grid_of(n)
    h, dt = grid[6], 1.0 / 240.0
    (x, y, z), vel = scenes.colliding_spheres(n, seed=4, radius=0.25)
    st = oracle.NsStep(x, y, z, *vel, grid)
    st.run()
    # the same by stages
    keys = oracle.ns_bin(x, y, z, grid)
    perm = np.argsort(keys, kind="stable")
    xs, ys, zs, vxs, vys, vzs = (a[perm] for a in (x, y, z, *vel))
    counts = np.bincount(keys, minlength=n ** 3 + 1).astype(np.uint32)
    P = oracle.ns_p2g(xs, ys, zs, vxs, vys, vzs, grid)
    F, F0 = oracle.ns_prepare_faces(*P, grid, -9.81 * dt)
    ty, coef, b = oracle.ns_assemble(counts, grid, h / dt, *F)
    p, it, rel = oracle.ns_pcg_jacobi(coef, grid, b, tol=1e-6)
    Fp = oracle.ns_velocity_update(ty, grid, p, dt / h, *F)
    v = oracle.ns_g2p_flip(xs, ys, zs, vxs, vys, vzs, *Fp, *F0, grid, 0.95)
    xo = oracle.ns_advect_rk2(xs, ys, zs, *Fp, grid, dt, 1e-3 * h)
    assert st.iterations == it
    got = st.state()
    for g, e in zip(got, list(xo) + list(v)):
        assert np.array_equal(g, e.astype(np.float32))


# ------------------------------------------------------ binning, tiled index
def test_tiled_index_is_the_brick_order(oracle):
    n = 16
    grid = grid_of(n)
    i, j, k = ijk(n)
    t = oracle.ns_tiled_index(i, j, k, grid)
    assert np.array_equal(np.sort(t), np.arange(n ** 3))   # a bijection
    # hand case: cell (9, 2, 13) is in brick (1, 0, 1) of 2 x 2 x 2 bricks ->
    # brick (1 * 2 + 0) * 2 + 1 = 5, local (1, 2, 5)
    assert oracle.ns_tiled_index(9, 2, 13, grid) == 5 * 512 + 5 * 64 + 2 * 8 + 1
    assert oracle.lib().orc_ns_tiled_index(9, 2, 13, oracle.OrcGrid(*grid)) == 5 * 512 + 5 * 64 + 2 * 8 + 1


def test_ns_bin_cells_and_clamping(oracle):
    n = 16
    grid = grid_of(n, origin=(-1.0, 0.5, 0.0), h=0.125)
    x = np.array([-1.0, -0.99, 0.999, -5.0, 9.0], np.float32)
    y = np.array([0.5, 0.5 + 0.125 * 3.5, 2.49, 0.0, 0.6], np.float32)
    z = np.array([0.0, 1.0, 1.99, 0.3, -2.0], np.float32)
    keys = oracle.ns_bin(x, y, z, grid)
    cell = [(0, 0, 0), (0, 3, 8), (15, 15, 15), (0, 0, 2), (15, 0, 0)]  # clamped into [0, n - 1]
    for k, (a, b, c) in zip(keys, cell):
        assert k == oracle.ns_tiled_index(a, b, c, grid)
