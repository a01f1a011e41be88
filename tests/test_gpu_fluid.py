"""GPU checks of the north-star FLIP step (mixed-precision PCG on bricks).

Stages are compared with the FP64 oracle on the same inputs (the GPU's own
inputs of each stage), 1e-5 normwise relative for transfers, faces, G2P and
advection.  The solve is judged by its FP64 true residual (<= 1e-6 relative,
recomputed here by the oracle's operator) and, against a tightly converged
FP64 solve of the same system, by the field error that residual permits
(stated per test).  Closed form: a still full-width slab has hydrostatic
pressure p_j = |g| h (J - j) and zero velocity after projection.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from flip2_amd import scenes
from tests.fluid_checks import dense_from_slots, rel_err, staged_step_vs_oracle


def dam_break_sim(flip2, n=32, precond="jacobi", seed=1):
    x, y, z = scenes.dam_break(n, seed=seed)
    zero = np.zeros_like(x)
    grid = (0.0, 0.0, 0.0, n, n, n, 1.0 / n)
    sim = flip2.FluidSim(x.size, grid, preconditioner=precond)
    sim.set_particles(x, y, z, zero, zero, zero)
    return sim, grid, (x, y, z)


def test_stages_match_oracle(flip2, oracle):
    sim, grid, _ = dam_break_sim(flip2, 32)
    err, _ = staged_step_vs_oracle(sim, grid, oracle)
    assert err["pressure"] <= 2e-3  # residual 1e-6 x condition number of the 32^3 Poisson operator


@pytest.mark.parametrize("precond", ["jacobi", "mic0"])
def test_hydrostatic_slab_closed_form(flip2, precond):
    n = 32
    h = 1.0 / n
    J = n // 2
    x, y, z = scenes.jittered_cells((n, n, n), h, seed=9, keep=lambda a, b, c: b < np.float32(0.5))
    zero = np.zeros_like(x)
    grid = (0.0, 0.0, 0.0, n, n, n, h)
    sim = flip2.FluidSim(x.size, grid, preconditioner=precond)
    sim.set_particles(x, y, z, zero, zero, zero)
    for s in ("transfer", "prepare", "system", "solve"):
        sim.stage(s)
    assert sim.last_solve()["converged"]
    p = dense_from_slots(sim, sim.get("p"), n ** 3)
    from oracle import pyoracle
    jj = np.arange(n)
    i, j, k = np.meshgrid(jj, jj, jj, indexing="ij")
    idx = pyoracle.ns_tiled_index(i, j, k, grid)
    expect = np.where(j < J, 9.81 * h * (J - j), 0.0)
    got = p[idx]
    assert rel_err(got, expect) <= 1e-4
    sim.stage("project")
    V = sim.get("V")
    assert np.max(np.abs(V[idx][j <= J])) <= 1e-4 * 9.81 / 240.0


def test_mic0_fewer_iterations_than_jacobi(flip2):
    sim, grid, _ = dam_break_sim(flip2, 64)
    for s in ("transfer", "prepare", "system"):
        sim.stage(s)
    sj = sim.solve_with("jacobi")
    sm = sim.solve_with("mic0")
    assert sj["converged"] and sm["converged"]
    assert sm["rel_residual"] <= 1e-6 and sj["rel_residual"] <= 1e-6
    assert sm["iterations"] < sj["iterations"], (sm, sj)


def test_dam_break_steps(flip2):
    sim, grid, (x, y, z) = dam_break_sim(flip2, 64)
    n0 = x.size
    y0 = float(np.mean(y))
    for _ in range(10):
        sim.step()
        st = sim.last_solve()
        assert st["converged"] and st["rel_residual"] <= 1e-6, st
    xs, ys, zs, vx, vy, vz = sim.particles()
    assert xs.size == n0
    for a in (xs, ys, zs):
        assert np.all(a > 0.0) and np.all(a < 1.0)
    assert float(np.mean(ys)) < y0          # the column falls
    assert float(np.mean(vy)) < 0.0
    assert np.all(np.isfinite(vx)) and np.all(np.isfinite(vy)) and np.all(np.isfinite(vz))


def test_moving_solid_box_boundary(flip2, oracle):
    """A solid box moving at u_p (a paddle): faces between box cells and fluid
    cells carry u_p before and after the projection, domain walls stay 0,
    and a particle that ends inside the box is pushed out of it."""
    n = 32
    h = 1.0 / n
    x, y, z = scenes.jittered_cells((n, n, n), h, seed=12,
                                    keep=lambda a, b, c: (a >= np.float32(0.25)) & (a < np.float32(0.75))
                                    & (b < np.float32(0.5)))
    # one particle inside the box region, at rest
    x = np.append(x, np.float32(0.2)); y = np.append(y, np.float32(0.1)); z = np.append(z, np.float32(0.5))
    zero = np.zeros_like(x)
    grid = (0.0, 0.0, 0.0, n, n, n, h)
    sim = flip2.FluidSim(x.size, grid, preconditioner="mic0")
    sim.set_particles(x, y, z, zero, zero, zero)
    up = 0.5
    sim.set_solid_box((0.0, 0.0, 0.0), (0.25, 1.0, 1.0), (up, 0.0, 0.0))
    sim.stage("transfer")
    sim.stage("prepare")
    U = sim.get("U")
    jj = np.arange(n)
    i, j, k = np.meshgrid(jj, jj, jj, indexing="ij")
    idx = oracle.ns_tiled_index(i, j, k, grid)
    Ud = U[idx]  # [i, j, k]
    assert np.all(Ud[8, :, :] == np.float32(up))      # box | fluid-or-air faces
    assert np.all(Ud[1:8, :, :] == np.float32(up))    # faces between box cells
    assert np.all(Ud[0, :, :] == 0.0)                 # domain wall
    for s in ("system", "solve"):
        sim.stage(s)
    assert sim.last_solve()["converged"]
    sim.stage("project")
    U2 = sim.get("U")[idx]
    assert np.all(U2[8, :, :] == np.float32(up))
    sim.stage("g2p")
    sim.stage("advect")
    px, py, pz, _, _, _ = sim.particles()
    assert np.all(px >= np.float32(0.25)), float(px.min())
    # without the box the wall faces inside x < 0.25 are ordinary faces again
    sim.set_solid_box(None)
    sim.stage("transfer")
    sim.stage("prepare")
    assert np.all(sim.get("U")[idx][1:8, :, :] == 0.0)  # no particles there: P2G gives 0


@pytest.mark.parametrize("solid", ["none", "box", "static"])
def test_fused_transfer_prepare_bit_identical(flip2, oracle, solid):
    """The step's fused transfer + prepare (scatter only in bricks within one
    brick of a particle, prepared zero faces elsewhere) gives the faces,
    saved faces and sorted records of the transfer stage followed by the
    prepare stage, bit for bit, on a 64^3 grid whose particles leave most
    bricks untouched, with a moving solid box or a static solid mask; over
    two steps, so the second fused transfer (which rewrites only the bricks
    that left the touched set) is covered too."""
    n = 64
    h = 1.0 / n
    rng = np.random.default_rng(8)
    blob = rng.random((20000, 3)) * np.array([0.3, 0.25, 0.3]) + np.array([0.1, 0.0, 0.4])
    edge = rng.random((500, 3)) * np.array([1.0, 0.02, 0.02]) + np.array([0.0, 0.97, 0.97])  # along the top corner
    pos = np.concatenate([blob, edge]).astype(np.float32)
    x, y, z = (np.ascontiguousarray(pos[:, d]) for d in range(3))
    v = [(rng.standard_normal(x.size) * 0.5).astype(np.float32) for _ in range(3)]
    v[1][-500:] = -50.0  # the edge particles leave their bricks within one step (12 cells)
    v[2][-500:] = -50.0
    grid = (0.0, 0.0, 0.0, n, n, n, h)
    names = ("U", "V", "W", "U0", "V0", "W0")
    out = []
    for fused in (False, True):
        sim = flip2.FluidSim(x.size, grid)
        sim.set_particles(x, y, z, *v)
        if solid == "box":
            sim.set_solid_box((0.0, 0.0, 0.0), (0.25, 0.5, 1.0), (0.5, 0.0, 0.25))
        elif solid == "static":
            m = np.zeros((n, n, n), bool)
            m[20:40, 0:6, 10:50] = True
            m[60:, :, :] = True
            from tests.fluid_checks import tiled_mask
            sim.set_solid(tiled_mask(m.transpose(2, 1, 0), grid, oracle))
        # two steps: the second fused transfer rewrites only the bricks that
        # left the touched set (the particles move between the two)
        for step in range(2):
            if step:
                for s in ("system", "solve", "project"):
                    sim.stage(s)
                if fused:
                    sim.stage("g2p_advect")
                else:
                    sim.stage("g2p")
                    sim.stage("advect")
            if fused:
                sim.stage("transfer_prepare")
            else:
                sim.stage("transfer")
                sim.stage("prepare")
        out.append([sim.get(k) for k in names] + list(sim.particles()))
    for name, a, b in zip(names + ("x", "y", "z", "vx", "vy", "vz"), *out):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), name


@pytest.mark.parametrize("box", [False, True])
def test_fused_g2p_advect_bit_identical(flip2, box):
    """The step's fused G2P + RK2 (by brick, from LDS copies of the faces,
    after the fused transfer) gives the records of the g2p stage followed by
    the advect stage, bit for bit (with and without a moving solid box,
    whose push-out and return-to-start branches it shares; the edge
    particles' fast motion takes some RK2 midpoints outside their brick's
    region, to the grid fallback)."""
    n = 32
    h = 1.0 / n
    x, y, z = scenes.dam_break(n, seed=21)
    x = np.append(x, np.float32(0.2)); y = np.append(y, np.float32(0.1)); z = np.append(z, np.float32(0.5))
    rng = np.random.default_rng(5)
    v = [(rng.standard_normal(x.size) * 0.5).astype(np.float32) for _ in range(3)]
    v[0][:2000] = 60.0  # a fast stretch: its grid velocity moves RK2 midpoints past the brick's region
    grid = (0.0, 0.0, 0.0, n, n, n, h)
    out = []
    for fused in (False, True):
        sim = flip2.FluidSim(x.size, grid, preconditioner="mic0")
        sim.set_particles(x, y, z, *v)
        if box:
            sim.set_solid_box((0.0, 0.0, 0.0), (0.25, 1.0, 1.0), (0.5, 0.0, 0.0))
        # the fused path as the step runs it: after the fused transfer the
        # G2P + RK2 kernel works by brick from LDS copies of the faces
        for s in (("transfer_prepare",) if fused else ("transfer", "prepare")) + ("system", "solve", "project"):
            sim.stage(s)
        if fused:
            sim.stage("g2p_advect")
        else:
            sim.stage("g2p")
            sim.stage("advect")
        out.append(sim.particles())
    for name, a, b in zip(("x", "y", "z", "vx", "vy", "vz"), *out):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), name


def test_snapshot_resume_bit_identical(flip2, tmp_path):
    """A run saved mid-way and resumed in a fresh simulation (default
    configuration: the snapshot restores the settings, the static solid and
    the solid box) steps to the same particles, bit for bit, as the run that
    continued without stopping."""
    n = 32
    (x, y, z), solid = scenes.dam_break_sphere(n, seed=3)
    zero = np.zeros_like(x)
    grid = (0.0, 0.0, 0.0, n, n, n, 1.0 / n)
    sim = flip2.FluidSim(x.size, grid, preconditioner="mic0")
    sim.set_particles(x, y, z, zero, zero, zero)
    sim.set_solid(solid)
    sim.set_solid_box((0.0, 0.0, 0.0), (0.05, 1.0, 1.0), (0.2, 0.0, 0.0))
    for _ in range(3):
        sim.step()
    path = tmp_path / "run.snap"
    sim.save_snapshot(path)
    for _ in range(3):
        sim.step()
    a = sim.particles()

    other = flip2.FluidSim(x.size, grid)  # Jacobi-PCG by default; the snapshot restores MIC(0)
    other.load_snapshot(path)
    for _ in range(3):
        other.step()
    b = other.particles()
    for u, v in zip(a, b):
        assert u.dtype == v.dtype and np.array_equal(u.view(np.uint32), v.view(np.uint32))
    assert other.last_solve()["iterations"] == sim.last_solve()["iterations"]

    # a snapshot of another grid size, or a cut file, is refused
    small = flip2.FluidSim(x.size, (0.0, 0.0, 0.0, 16, 16, 16, 1.0 / 16))
    with pytest.raises(flip2.Flip2Error):
        small.load_snapshot(path)
    cut = tmp_path / "cut.snap"
    cut.write_bytes(path.read_bytes()[:1000])
    with pytest.raises(flip2.Flip2Error):
        other.load_snapshot(cut)


def test_kernel_timing_sample(flip2):
    """The per-kernel event timer records every launch, or with a sampling
    stride one in every `sample_every` launches of each kernel (the same
    deterministic step on two simulations: the same launch count, divided)."""
    n = 32
    x, y, z = scenes.dam_break(n, seed=3)
    zero = np.zeros_like(x)
    grid = (0.0, 0.0, 0.0, n, n, n, 1.0 / n)
    counts = []
    for every in (1, 4):
        sim = flip2.FluidSim(x.size, grid, preconditioner="mic0")
        sim.set_particles(x, y, z, zero, zero, zero)
        sim.step()
        sim.enable_timing(["cg_update_mic_kernel"], sample_every=every)
        sim.step()
        t = sim.timing()
        sim.disable_timing()
        assert set(t) == {"cg_update_mic_kernel"}, t
        ms, launches = t["cg_update_mic_kernel"]
        assert ms > 0.0
        counts.append(launches)
    assert counts[0] > 8 and counts[1] == (counts[0] + 3) // 4, counts


def test_step_frame_cfl_substeps(flip2):
    """Adaptive step: with a fast initial velocity the frame is split into
    substeps of dt = min(remaining, cfl h / max|v_c|) (exact maximum over
    all particles); the same substeps taken one by one with step() give the
    same particles bit for bit."""
    n = 32
    h = float(np.float32(1.0 / n))
    x, y, z = scenes.dam_break(n, seed=4)
    vx = np.full_like(x, np.float32(6.0))
    zero = np.zeros_like(x)
    grid = (0.0, 0.0, 0.0, n, n, n, 1.0 / n)
    frame, cfl = 1.0 / 240.0, 0.5

    a = flip2.FluidSim(x.size, grid, preconditioner="mic0")
    a.set_particles(x, y, z, vx, zero, zero)
    k = a.step_frame(frame, cfl=cfl, max_substeps=64)
    assert k >= 2

    b = flip2.FluidSim(x.size, grid, preconditioner="mic0")
    b.set_particles(x, y, z, vx, zero, zero)
    t, steps = 0.0, 0
    while frame - t > 1e-12 * frame:
        _, _, _, u, v, w = b.particles()
        vmax = float(np.max(np.abs(np.concatenate([u, v, w]))))
        dt = min(frame - t, cfl * h / vmax) if vmax > 0.0 else frame - t
        b.set_dt(dt)
        b.step()
        t += dt
        steps += 1
    assert steps == k
    for p, q in zip(a.particles(), b.particles()):
        assert np.array_equal(p.view(np.uint32), q.view(np.uint32))
    assert a.cfg.dt == np.float32(1.0 / 240.0)  # the configured dt is left as it was

    # one allowed substep takes the whole frame
    c = flip2.FluidSim(x.size, grid, preconditioner="mic0")
    c.set_particles(x, y, z, vx, zero, zero)
    assert c.step_frame(frame, cfl=cfl, max_substeps=1) == 1
    with pytest.raises(flip2.Flip2Error):
        c.step_frame(-1.0)
