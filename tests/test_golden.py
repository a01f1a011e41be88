"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the oracle reproduces every fixture bit for bit, and the fixtures obey
the reference's own result invariant (testing.h:34-53: keys ascending, the
permutation a bijection that maps inputs onto the sorted keys).
GPU: the library reproduces the fixtures without the oracle at run time:
integer arrays bit-exact, FP64 reference-compatible stages to 1e-5 (relative
to the array's max magnitude), north-star stages to 1e-5 with the pressure
solved to a 1e-11 FP64 residual (so the solve tolerance does not show in the
downstream fields).
"""
import numpy as np
import pytest

from tests.golden import make_golden as mg

RTOL = 1e-5


def close(a, b, name):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    assert a.shape == b.shape, (name, a.shape, b.shape)
    scale = max(float(np.max(np.abs(b))) if b.size else 0.0, 1e-300)
    err = float(np.max(np.abs(a - b))) if a.size else 0.0
    assert err <= RTOL * scale, f"{name}: {err:.3e} vs {scale:.3e}"


def test_oracle_reproduces_fixtures():
    assert mg.main(["--check"]) == 0


def test_fixture_sort_invariants():
    f = mg.load("primitives")
    k, out, perm = f["sort_keys"], f["sort_keys_out"], f["sort_perm"]
    assert np.all(np.diff(out.astype(np.int64)) >= 0)
    assert np.array_equal(np.sort(perm), np.arange(k.size))
    assert np.array_equal(k[perm], out)
    assert np.array_equal(perm, np.argsort(k, kind="stable"))
    v = f["scan_in"].astype(np.uint64)
    assert np.array_equal(f["scan_incl"], (np.cumsum(v) & 0xFFFFFFFF).astype(np.uint32))
    r = mg.load("rc_block8")
    g = r["sort_gridCell"]
    assert np.all(np.diff(g.astype(np.int64)) >= 0)
    assert np.array_equal(r["align_gridCell"][r["sort_reorderedGridIndices"]], g)


@pytest.mark.gpu
def test_gpu_primitives_match_fixtures(flip2):
    f = mg.load("primitives")
    incl, _ = flip2.scan_u32(f["scan_in"], exclusive=False)
    excl, _ = flip2.scan_u32(f["scan_in"], exclusive=True)
    assert np.array_equal(incl, f["scan_incl"])
    assert np.array_equal(excl, f["scan_excl"])
    k, v = flip2.sort_pairs_u32(f["sort_keys"], None, 0, 12)
    assert np.array_equal(k, f["sort_keys_out"])
    assert np.array_equal(v, f["sort_perm"])


@pytest.mark.gpu
def test_gpu_rc_stages_match_fixtures(flip2):
    f = mg.load("rc_block8")
    n = f["in_px"].size
    g = flip2.Particles(n)
    g.setDomain(*mg.RC_DOMAIN)
    g.set_positions(f["in_px"], f["in_py"], f["in_pz"])
    g.set_velocities(f["in_vx"], f["in_vy"], f["in_vz"])
    g.alignParticlesToGrid()
    assert np.array_equal(g.get("gridCell"), f["align_gridCell"])
    g.sortParticles()
    for k in ("gridCell", "reorderedGridIndices", "px", "py", "pz"):
        assert np.array_equal(g.get(k), f["sort_" + k]), k
    g.generateVoxels()
    for k in ("uniqueGridNodeIndices", "gridNodeIndicesToFirstParticleIndex", "yDimNumUsedGridNodes",
              "nodeIndexUsedVoxels", "voxelIDsUsed", "solids"):
        assert np.array_equal(g.get(k), f["vox_" + k]), k
    g.particleVelToVoxels()
    for k in ("voxelsUx", "voxelsUy", "voxelsUz"):
        close(g.get(k), f["p2g_" + k], k)
    g.setTime(mg.RC_DT, 1.0 / 24)
    g.applyGravity(mg.RC_DT)
    g.calcDivU()
    close(g.get("voxelsUy"), f["div_voxelsUy"], "Uy+g")
    close(g.get("divU"), f["div_divU"], "divU")
    g.gsIterations(mg.RC_DT, 64)
    close(g.get("p"), f["sor_p"], "p")
    close(g.get("residuals"), f["sor_residuals"], "residuals")
    g.updateVoxelVelocities()
    for k in ("voxelsUx", "voxelsUy", "voxelsUz"):
        close(g.get(k), f["upd_" + k], "upd " + k)
    g.voxelVelsToParticles()
    for k in ("vx", "vy", "vz"):
        close(g.get(k), f["g2p_" + k], "g2p " + k)
    g.advectParticles()
    for k in ("px", "py", "pz"):
        close(g.get(k), f["adv_" + k], "adv " + k)


@pytest.mark.gpu
def test_gpu_ns_stages_match_fixtures(flip2):
    f = mg.load("ns_dam16")
    g = mg.NS_GRID
    n = f["in_x"].size
    sim = flip2.FluidSim(n, g, dt=mg.NS_DT, tolerance=1e-11, max_iterations=4000)
    sim.set_particles(f["in_x"], f["in_y"], f["in_z"], f["in_vx"], f["in_vy"], f["in_vz"])
    sim.stage("transfer")
    assert np.array_equal(sim.get("counts")[:mg.NS_N ** 3], f["counts"])
    for c in ("U", "V", "W"):
        close(sim.get(c), f["p2g_" + c], "p2g " + c)
    sim.stage("prepare")
    for c in ("U", "V", "W"):
        close(sim.get(c), f["prep_" + c], "prep " + c)
    sim.stage("system")
    ab = sim.get("activeBricks").astype(np.int64)
    idx = (ab[:, None] * 512 + np.arange(512)[None, :]).reshape(-1)

    def dense(a):
        out = np.zeros(mg.NS_N ** 3, a.dtype)
        out[idx] = a[:idx.size]
        return out

    assert np.array_equal(dense(sim.get("coef")), f["coef"])
    close(dense(sim.get("b")), f["b"], "b")
    sim.stage("solve")
    st = sim.last_solve()
    assert st["converged"] and st["rel_residual"] <= 1e-11, st
    close(dense(sim.get("p")), f["p"], "p")
    sim.stage("project")
    for c in ("U", "V", "W"):
        close(sim.get(c), f["proj_" + c], "proj " + c)
    # particle state is in cell-sorted (stable) order after the transfer
    order = np.argsort(f["keys"], kind="stable")
    sim.stage("g2p")
    p = sim.particles()
    for a, c in zip(p[3:], ("vx", "vy", "vz")):
        close(a, f["g2p_" + c][order], "g2p " + c)
    sim.stage("advect")
    p = sim.particles()
    for a, c in zip(p[:3], ("x", "y", "z")):
        close(a, f["adv_" + c][order], "adv " + c)
