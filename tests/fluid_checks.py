"""Shared checks of the north-star FLIP step against the FP64 oracle.

staged_step_vs_oracle runs one step stage by stage and compares each stage
with the oracle fed the GPU's own inputs of that stage:
  - transfer (P2G faces), walls + gravity, projected faces, particle
    velocities after the FLIP update and positions after RK2: max-norm
    relative error <= 1e-5 (RTOL);
  - operator coefficients and per-cell counts: equal (integer work);
  - right-hand side: <= 1e-5;
  - the solve: FP64 true residual of the GPU pressure, recomputed with the
    oracle's operator, <= 1e-6 relative (the solver's tolerance), and the
    post-projection divergence <= 2e-6 of the divergence before;
  - the pressure against a tight (1e-12) FP64 solve of the same system: the
    error the 1e-6 residual permits (returned, bounded by the caller).
Returns the measured errors per field.
"""
import numpy as np

RTOL = 1e-5


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def dense_from_slots(sim, arr, cells):
    ab = sim.get("activeBricks").astype(np.int64)
    out = np.zeros(cells, dtype=arr.dtype)
    idx = (ab[:, None] * 512 + np.arange(512)[None, :]).reshape(-1)
    out[idx] = arr[:idx.size]
    return out


def tiled_mask(mask, grid, oracle):
    """A solid mask indexed x + y*nx + z*nx*ny in the brick-tiled cell order."""
    nx, ny, nz = grid[3], grid[4], grid[5]
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    out = np.zeros(nx * ny * nz, np.uint8)
    out[oracle.ns_tiled_index(i.ravel(), j.ravel(), k.ravel(), grid)] = np.asarray(mask, np.uint8).reshape(-1)
    return out


def staged_step_vs_oracle(sim, grid, oracle, dt=1.0 / 240.0, gravity=-9.81, flip=0.95, solid=None):
    n = grid[3]
    cells = grid[3] * grid[4] * grid[5]
    h = grid[6]
    err = {}
    xs, ys, zs, vxs, vys, vzs = sim.particles()
    # transfer: counts (integer) and P2G faces
    sim.stage("transfer")
    keys = oracle.ns_bin(xs, ys, zs, grid)
    counts = sim.get("counts")
    assert np.array_equal(counts, np.bincount(keys, minlength=cells + 1).astype(np.uint32))
    P = [sim.get(c).astype(np.float64) for c in ("U", "V", "W")]
    Po = oracle.ns_p2g(xs, ys, zs, vxs, vys, vzs, grid)
    for name, a, b in zip("UVW", P, Po):
        e = rel_err(a, b) if np.max(np.abs(b)) > 0.0 else float(np.max(np.abs(a)))
        err["P2G " + name] = e
        assert e <= RTOL, (name, e)
    # walls + gravity, saved field
    sim.stage("prepare")
    F, F0 = oracle.ns_prepare_faces(*P, grid, gravity * dt, solid)
    for name, ref in zip(("U", "V", "W", "U0", "V0", "W0"), F + F0):
        e = rel_err(sim.get(name), ref) if np.max(np.abs(ref)) > 0.0 else float(np.max(np.abs(sim.get(name))))
        err["faces " + name] = e
        assert e <= RTOL, (name, e)
    # system
    sim.stage("system")
    Fg = [sim.get(c).astype(np.float64) for c in ("U", "V", "W")]
    ty, coef, b = oracle.ns_assemble(counts, grid, h / dt, *Fg, solid=solid)
    assert np.array_equal(dense_from_slots(sim, sim.get("coef"), cells), coef)
    err["rhs b"] = rel_err(dense_from_slots(sim, sim.get("b"), cells), b)
    assert err["rhs b"] <= RTOL
    # solve: FP64 true residual of the GPU pressure
    sim.stage("solve")
    st = sim.last_solve()
    assert st["converged"], st
    p = dense_from_slots(sim, sim.get("p"), cells)
    r = b - oracle.ns_apply_A(coef, grid, p)
    res = float(np.linalg.norm(r) / np.linalg.norm(b))
    err["FP64 true residual"] = res
    assert res <= 1.0001e-6, res
    pref, _, relref = oracle.ns_pcg_jacobi(coef, grid, b, tol=1e-12)
    assert relref <= 1e-12
    err["pressure"] = rel_err(p, pref)
    err["pressure (normwise)"] = float(np.linalg.norm(p - pref) / np.linalg.norm(pref))
    # project
    sim.stage("project")
    Fp = oracle.ns_velocity_update(ty, grid, p, dt / h, *Fg)
    for name, ref in zip("UVW", Fp):
        err["projected " + name] = rel_err(sim.get(name), ref)
        assert err["projected " + name] <= RTOL, name
    _, _, b_after = oracle.ns_assemble(counts, grid, h / dt, *[sim.get(c).astype(np.float64) for c in "UVW"],
                                       solid=solid)
    err["divergence after / before"] = float(np.linalg.norm(b_after) / np.linalg.norm(b))
    assert err["divergence after / before"] <= 2e-6, err
    # g2p (state in sorted order after the transfer)
    xs, ys, zs, vxs, vys, vzs = sim.particles()
    Fn = [sim.get(c) for c in ("U", "V", "W", "U0", "V0", "W0")]
    vo = oracle.ns_g2p_flip(xs, ys, zs, vxs, vys, vzs, *Fn, grid, flip)
    sim.stage("g2p")
    _, _, _, vx2, vy2, vz2 = sim.particles()
    for c, a, ref in zip("xyz", (vx2, vy2, vz2), vo):
        err["particle v" + c] = rel_err(a, ref)
        assert err["particle v" + c] <= RTOL
    # advect
    xo = oracle.ns_advect_rk2(xs, ys, zs, *Fn[:3], grid, dt, 1e-3 * h, solid)
    sim.stage("advect")
    x2, y2, z2, _, _, _ = sim.particles()
    for c, a, ref in zip("xyz", (x2, y2, z2), xo):
        err["particle " + c] = rel_err(a, ref)
        assert err["particle " + c] <= RTOL
    print("max-norm relative error vs the FP64 oracle:", {k: float(f"{v:.2e}") for k, v in err.items()})
    return err, st
