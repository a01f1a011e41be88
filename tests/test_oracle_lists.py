"""Pins of the oracle's sub-cell lists and per-voxel particle lists
(oracle/rc_oracle.c: orc_find_sub_cell, orc_sum_particles_per_node,
orc_particle_list_create; the reference's particleToGridFunctions.cu:69-262
and perVoxelParticleListFunctions.cu:6-189).

- Sub-cells: particles on a lattice of 1/8 sub-cell, on a grid whose cell
  size is a power of two, so every quantity of the reference's mixed
  float/double arithmetic is exact; a sub-cell (x, y, z) of the 4^3 window
  around the particle's sub-cell s is touched for component c exactly when
  the integer lattice distance to its staggered point is below 2 sub-cells:
  |L - 8 (x,y,z) + 4 (1 - e_c)|^2 < 16^2 (lattice units, e_c the component's
  axis).  Counted and listed in the reference's loop order (x outer, z inner).
- Per-voxel lists: the structure is fixed by the (particle, voxel) incidences
  of each node -- used voxels ascending, each list the node's particles in
  ascending order, lists concatenated in voxel order -- computed here with
  numpy's lexsort and compared.
"""
import numpy as np

from oracle import pyoracle as O


def lattice_particles(n_cells, count, seed):
    rng = np.random.default_rng(seed)
    h = 1.0 / n_cells
    lat = rng.integers(0, 32, size=(count, 3))  # position in a cell, units of h / 32 = 1/8 sub-cell
    cell = rng.integers(0, n_cells, size=(count, 3))
    pos = (cell + lat / 32.0) * h
    return pos, cell, lat


def expected_lists(cell, lat, n_cells):
    """The touched sub-cells per particle and component in integer lattice
    units (sub-cell = 8 units; the particle sits 16 units + lat in the apron-
    offset frame)."""
    out = [[], [], []]
    counts = np.zeros((3, len(lat)), np.int64)
    for i, L in enumerate(lat):
        o = 16 + L  # apron of 2 sub-cells = 16 units
        s = o // 8
        for x in range(s[0] - 2, s[0] + 2):
            for y in range(s[1] - 2, s[1] + 2):
                for z in range(s[2] - 2, s[2] + 2):
                    d = o - 8 * np.array([x, y, z])
                    for c in range(3):
                        dd = d + 4 * (np.arange(3) != c)
                        if int(np.sum(dd * dd)) < 256:
                            out[c].append(x + 8 * y + 64 * z)
                            counts[c, i] += 1
    return [np.cumsum(v).astype(np.uint32) for v in counts], [np.array(v, np.uint32) for v in out]


def test_sub_cells_on_a_lattice():
    n = 8
    grid = (0.0, 0.0, 0.0, n, n, n, 1.0 / n)
    pos, cell, lat = lattice_particles(n, 400, seed=3)
    keys = (cell[:, 0] + cell[:, 1] * n + cell[:, 2] * n * n).astype(np.uint32)
    touched, lists = O.find_sub_cell(pos[:, 0], pos[:, 1], pos[:, 2], grid, keys)
    t_exp, l_exp = expected_lists(cell, lat, n)
    for c in range(3):
        assert np.array_equal(touched[c], t_exp[c]), c
        assert np.array_equal(lists[c], l_exp[c]), c


def test_sub_cell_closed_forms():
    # a particle at a sub-cell centre: for each component the touched
    # staggered points are those of the 4^3 window within 2 sub-cells
    n = 4
    grid = (0.0, 0.0, 0.0, n, n, n, 1.0 / n)
    h = 1.0 / n
    p = np.array([[1.5 * h / 4 + h, 2.5 * h / 4 + h, 0.5 * h / 4 + 2 * h]])
    keys = np.array([1 + 1 * n + 2 * n * n], np.uint32)
    touched, lists = O.find_sub_cell(p[:, 0], p[:, 1], p[:, 2], grid, keys)
    # the window is x in [s - 2, s + 2) and component c's staggered point of
    # sub-cell x sits at x on its axis and x - 1/2 on the others, so the
    # offsets from a particle at s + 1/2 are {2.5, 1.5, 0.5, -0.5} on the
    # axis and {3, 2, 1, 0} on the other two (sub-cell units); by symmetry
    # every component touches the same number: 11
    offs_axis = np.array([2.5, 1.5, 0.5, -0.5])
    offs_other = np.array([3.0, 2.0, 1.0, 0.0])
    expect = sum(1 for a in offs_axis for b in offs_other for c in offs_other if a * a + b * b + c * c < 4.0)
    assert expect == 11
    assert [int(t[0]) for t in touched] == [expect] * 3


def incidences(first, touched, lists, n):
    B = first.size
    node_of = np.zeros(n, np.int64)
    ends = np.r_[first[1:], n]
    for b in range(B):
        node_of[first[b]:ends[b]] = b
    per = np.diff(np.r_[0, touched.astype(np.int64)])
    particle = np.repeat(np.arange(n), per)
    return node_of[particle], particle, lists.astype(np.int64)


def test_per_voxel_lists_match_a_lexsort():
    n_cells = 6
    grid = (0.0, 0.0, 0.0, n_cells, n_cells, n_cells, 1.0 / n_cells)
    rng = np.random.default_rng(8)
    x = rng.uniform(0.1, 0.9, size=(1500, 3))
    px, py, pz, k = O.find_grid_cell(x[:, 0], x[:, 1], x[:, 2], grid)
    kk, perm = O.radix_sort_u32(k)
    px, py, pz = px[perm], py[perm], pz[perm]
    first = np.flatnonzero(np.r_[True, kk[1:] != kk[:-1]]).astype(np.uint32)
    touched, lists = O.find_sub_cell(px, py, pz, grid, kk)
    npart = kk.size
    for c in range(3):
        nz, cnt = O.sum_particles_per_node(first, touched[c], lists[c], npart)
        node, particle, voxel = incidences(first, touched[c], lists[c], npart)
        B = first.size
        # per-node counts
        exp_cnt = np.bincount(node, minlength=B)
        exp_nz = np.array([np.unique(voxel[node == b]).size for b in range(B)])
        assert np.array_equal(cnt, np.cumsum(exp_cnt).astype(np.uint32))
        assert np.array_equal(nz, np.cumsum(exp_nz).astype(np.uint32))
        vid, start, plist = O.particle_list_create(first, touched[c], lists[c], npart, nz, cnt)
        order = np.lexsort((particle, voxel, node))  # node, then voxel, then particle
        assert np.array_equal(plist, particle[order].astype(np.uint32))
        key = node[order] * 512 + voxel[order]
        heads = np.r_[True, key[1:] != key[:-1]]
        assert np.array_equal(vid, voxel[order][heads].astype(np.uint32))
        assert np.array_equal(start, np.flatnonzero(heads).astype(np.uint32))
