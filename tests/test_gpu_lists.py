"""Sub-cell lists and per-voxel particle lists on the GPU (lists.hip through
the C ABI: flip2_find_sub_cell, flip2_sum_particles_per_node,
flip2_particle_list_create) against the oracle, bit for bit: the counts'
inclusive scans, the concatenated sub-cell lists of each velocity component,
the per-node scans of distinct voxels and list entries, the used voxels,
list starts and per-voxel particle lists.  Inputs: sorted, clamped
reference-compatible particle state (Particles::initialize) on a small
scene and on the 64^3 dam break (503k particles)."""
import numpy as np
import pytest

from flip2_amd import scenes

pytestmark = pytest.mark.gpu


def rc_state(flip2, n, x, y, z):
    p = flip2.Particles(x.size)
    p.setDomain(0.0, 0.0, 0.0, n, n, n, 1.0 / n)
    p.set_positions(x.astype(np.float64), y.astype(np.float64), z.astype(np.float64))
    zero = np.zeros(x.size)
    p.set_velocities(zero, zero, zero)
    p.initialize()
    return [p.get(a) for a in ("px", "py", "pz", "gridCell", "gridNodeIndicesToFirstParticleIndex")]


@pytest.mark.parametrize("scene", ["block16", "dam64"])
def test_lists_match_oracle(flip2, oracle, scene):
    if scene == "block16":
        n = 16
        mask = np.zeros((n, n, n), dtype=bool)
        mask[2:9, 0:7, 3:12] = True
        x, y, z = scenes.jittered_cells((n, n, n), 1.0 / n, seed=12, cell_mask=mask)
    else:
        n = 64
        x, y, z = scenes.dam_break(n, seed=1)
    grid = (0.0, 0.0, 0.0, n, n, n, 1.0 / n)
    print(f"lists: {scene}, {x.size} particles", flush=True)
    px, py, pz, keys, first = rc_state(flip2, n, x, y, z)
    npart = keys.size
    touched, lists = flip2.find_sub_cell(px, py, pz, grid, keys)
    t_o, l_o = oracle.find_sub_cell(px, py, pz, grid, keys)
    for c in range(3):
        assert np.array_equal(touched[c], t_o[c]), c
        assert np.array_equal(lists[c], l_o[c]), c
        nz, cnt = flip2.sum_particles_per_node(first, touched[c], lists[c], npart)
        nz_o, cnt_o = oracle.sum_particles_per_node(first, t_o[c], l_o[c], npart)
        assert np.array_equal(nz, nz_o) and np.array_equal(cnt, cnt_o), c
        vid, start, plist = flip2.particle_list_create(first, touched[c], lists[c], npart, nz, cnt)
        vid_o, start_o, plist_o = oracle.particle_list_create(first, t_o[c], l_o[c], npart, nz_o, cnt_o)
        assert np.array_equal(vid, vid_o) and np.array_equal(start, start_o), c
        assert np.array_equal(plist, plist_o), c
    print(f"lists: {[int(t[-1]) for t in touched]} sub-cell entries, {int(nz[-1])} used voxels (z)", flush=True)
