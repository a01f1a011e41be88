"""Writes the golden fixtures of this directory from the FP64 oracle.

  python tests/golden/make_golden.py          # rewrite the .npz files
  python tests/golden/make_golden.py --check  # compare, exit 1 on difference

The reference ships no golden vectors (its only result check is the
sortedness/count invariant of testing.h:34-53), so these fixtures are the
oracle's outputs on small seeded scenes, kept so that the library's results
can be compared with fixed data and the oracle itself cannot drift:

  primitives.npz  inclusive/exclusive u32 scans, stable key/value radix sort
  rc_block8.npz   reference-compatible stages on an 8^3 grid: binned keys,
                  sorted keys and permutation, node/voxel tables, P2G, gravity
                  + divergence, 64 SOR iterations, velocity update, G2P, Euler
                  advection (particles.cu:48-368, 379-669;
                  voxelSolveFunctions.cu:20-486)
  ns_dam16.npz    north-star stages on a 16^3 grid with a fluid block
                  x < 0.3, y < 0.5, z < 0.5: keys, per-cell
                  counts, P2G, wall faces + gravity, operator and right-hand
                  side, FP64 pressure (1e-12), projection, FLIP G2P, RK2
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import pyoracle as orc  # noqa: E402
from flip2_amd import scenes  # noqa: E402


def primitives():
    rng = np.random.default_rng(20)
    vals = rng.integers(0, 1 << 20, 3000, dtype=np.uint32)
    keys = rng.integers(0, 1 << 12, 4097, dtype=np.uint32)  # many duplicates: stability matters
    sk, perm = orc.radix_sort_u32(keys, 12)
    return {
        "scan_in": vals,
        "scan_incl": orc.scan_u32(vals, exclusive=False),
        "scan_excl": orc.scan_u32(vals, exclusive=True),
        "sort_keys": keys,
        "sort_keys_out": sk,
        "sort_perm": perm,
    }


RC_DOMAIN = (0.0, 0.0, 0.0, 8, 8, 8, 1.0 / 8)
RC_DT = 1.0 / 480


def rc_block8():
    pos, vel = scenes.rc_block_scene(8, 1.0 / 8, (2, 3, 2), seed=31, origin_cell=(2, 0, 3), vel_amp=0.5,
                                     vel_seed=32)
    n = pos[0].size
    s = orc.OracleParticles(n)
    s.setDomain(*RC_DOMAIN)
    s.set_positions(*pos)
    s.set_velocities(*vel)
    out = {"in_px": pos[0], "in_py": pos[1], "in_pz": pos[2], "in_vx": vel[0], "in_vy": vel[1], "in_vz": vel[2]}
    s.alignParticlesToGrid()
    out["align_gridCell"] = s.get("gridCell")
    s.sortParticles()
    for k in ("gridCell", "reorderedGridIndices", "px", "py", "pz"):
        out["sort_" + k] = s.get(k)
    s.generateVoxels()
    for k in ("uniqueGridNodeIndices", "gridNodeIndicesToFirstParticleIndex", "yDimNumUsedGridNodes",
              "nodeIndexUsedVoxels", "voxelIDsUsed", "solids"):
        out["vox_" + k] = s.get(k)
    s.particleVelToVoxels()
    for k in ("voxelsUx", "voxelsUy", "voxelsUz"):
        out["p2g_" + k] = s.get(k)
    s.setTime(RC_DT, 1.0 / 24)
    s.applyGravity(RC_DT)
    s.calcDivU()
    out["div_voxelsUy"] = s.get("voxelsUy")
    out["div_divU"] = s.get("divU")
    s.gsIterations(RC_DT, 64)
    out["sor_p"] = s.get("p")
    out["sor_residuals"] = s.get("residuals")
    s.updateVoxelVelocities()
    for k in ("voxelsUx", "voxelsUy", "voxelsUz"):
        out["upd_" + k] = s.get(k)
    s.voxelVelsToParticles()
    for k in ("vx", "vy", "vz"):
        out["g2p_" + k] = s.get(k)
    s.advectParticles()
    for k in ("px", "py", "pz"):
        out["adv_" + k] = s.get(k)
    s.close()
    return out


NS_N = 16
NS_GRID = (0.0, 0.0, 0.0, NS_N, NS_N, NS_N, 1.0 / NS_N)
NS_DT = 1.0 / 240


def ns_dam16():
    h = 1.0 / NS_N
    x, y, z = scenes.jittered_cells((NS_N, NS_N, NS_N), h, 41,
                                    keep=lambda x, y, z: (x < np.float32(0.3)) & (y < np.float32(0.5))
                                    & (z < np.float32(0.5)))
    vx, vy, vz = scenes.uniform_velocities(x.size, 42, 0.25)
    g = NS_GRID
    out = {"in_x": x, "in_y": y, "in_z": z, "in_vx": vx, "in_vy": vy, "in_vz": vz}
    keys = orc.ns_bin(x, y, z, g)
    out["keys"] = keys
    counts = np.bincount(keys, minlength=NS_N ** 3).astype(np.uint32)
    out["counts"] = counts
    U, V, W = orc.ns_p2g(x, y, z, vx, vy, vz, g)
    out["p2g_U"], out["p2g_V"], out["p2g_W"] = U, V, W
    F, F0 = orc.ns_prepare_faces(U, V, W, g, -9.81 * NS_DT)
    out["prep_U"], out["prep_V"], out["prep_W"] = F
    ty, coef, b = orc.ns_assemble(counts, g, h / NS_DT, *F)
    out["types"], out["coef"], out["b"] = ty, coef, b
    p, it, rel = orc.ns_pcg_jacobi(coef, g, b, tol=1e-12)
    assert rel <= 1e-12
    out["p"] = p
    Fp = orc.ns_velocity_update(ty, g, p, NS_DT / h, *F)
    out["proj_U"], out["proj_V"], out["proj_W"] = Fp
    v = orc.ns_g2p_flip(x, y, z, vx, vy, vz, *Fp, *F0, g, 0.95)
    out["g2p_vx"], out["g2p_vy"], out["g2p_vz"] = v
    pos = orc.ns_advect_rk2(x, y, z, *Fp, g, NS_DT, 1e-3 * h)
    out["adv_x"], out["adv_y"], out["adv_z"] = pos
    return out


FIXTURES = {"primitives": primitives, "rc_block8": rc_block8, "ns_dam16": ns_dam16}


def path(name):
    return os.path.join(HERE, name + ".npz")


def load(name):
    with np.load(path(name), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def main(argv):
    check = "--check" in argv
    bad = 0
    for name, fn in FIXTURES.items():
        data = fn()
        if check:
            ref = load(name)
            for k, v in data.items():
                if k not in ref or ref[k].dtype != v.dtype or not np.array_equal(ref[k], v):
                    print(f"{name}:{k} differs")
                    bad += 1
        else:
            np.savez_compressed(path(name), **data)
            print(f"wrote {path(name)} ({os.path.getsize(path(name))} bytes, {len(data)} arrays)")
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
