"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The shipped library never uses
it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None

_u32p = ctypes.POINTER(ctypes.c_uint32)
_f64p = ctypes.POINTER(ctypes.c_double)


class OrcGrid(ctypes.Structure):
    _fields_ = [("negX", ctypes.c_float), ("negY", ctypes.c_float), ("negZ", ctypes.c_float),
                ("sizeX", ctypes.c_uint32), ("sizeY", ctypes.c_uint32), ("sizeZ", ctypes.c_uint32),
                ("cellSize", ctypes.c_float)]


class OrcParticles(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_uint32),
        ("px", _f64p), ("py", _f64p), ("pz", _f64p), ("vx", _f64p), ("vy", _f64p), ("vz", _f64p),
        ("gridCell", _u32p), ("reorderedGridIndices", _u32p), ("uniqueGridNodeIndices", _u32p),
        ("numUsedGridNodes", ctypes.c_uint32),
        ("gridNodeIndicesToFirstParticleIndex", _u32p),
        ("yDimRows", ctypes.c_uint32),
        ("yDimNumUsedGridNodes", _u32p),
        ("nodeIndexUsedVoxels", _u32p),
        ("numUsedVoxels", ctypes.c_uint32),
        ("voxelIDsUsed", _u32p),
        ("solids", ctypes.POINTER(ctypes.c_int8)),
        ("voxelsUx", _f64p), ("voxelsUy", _f64p), ("voxelsUz", _f64p),
        ("divU", _f64p), ("p", _f64p), ("residuals", _f64p),
        ("grid", OrcGrid),
        ("radius", ctypes.c_double),
        ("refinementLevel", ctypes.c_uint32), ("numVoxels1D", ctypes.c_uint32),
        ("numVoxelsPerNode", ctypes.c_uint32),
        ("frameDt", ctypes.c_double), ("dt", ctypes.c_double), ("elapsedTime", ctypes.c_double),
        ("elapsedTimeThisFrame", ctypes.c_double),
        ("lastSorIterations", ctypes.c_uint32), ("lastSolveAttempts", ctypes.c_uint32),
        ("lastResidual", ctypes.c_double),
        ("outOfTileEvents", ctypes.c_uint32),
        ("moveSolidsCounter", ctypes.c_uint64),
        ("relocatedParticles", ctypes.c_uint32),
        ("sorThreshold", ctypes.c_double),
        ("sorMaxIterations", ctypes.c_uint32),
    ]


SOLVE_CB = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_void_p, ctypes.c_double)
HALVE_CB = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_double, ctypes.c_double)


def pressure_retry(dt, threshold, residuals):
    """Drive the oracle's pressureSolve retry control with a scripted
    sequence of terminating residuals; returns (solves, final dt, the dt of
    each solve, the (old, new) dt of each halving, final residual)."""
    seq = list(residuals)
    seen, halvings = [], []

    def solve(_ctx, d):
        seen.append(d)
        return seq[len(seen) - 1]

    def halve(_ctx, old, new):
        halvings.append((old, new))

    cs, ch = SOLVE_CB(solve), HALVE_CB(halve)
    d = ctypes.c_double(dt)
    last = ctypes.c_double()
    n = lib().orc_pressure_retry(ctypes.byref(d), threshold, cs, ch, None, ctypes.byref(last))
    return n, d.value, seen, halvings, last.value


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        l = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        l.orc_find_grid_cell.argtypes = [vp, vp, vp, ctypes.c_uint32, OrcGrid, vp]
        l.orc_scan_u32.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_int]
        l.orc_radix_sort_u32.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_uint32]
        l.orc_radix_sort_pairs_u32.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_uint32]
        l.orc_sampled_absmax.argtypes = [vp, ctypes.c_uint32]
        l.orc_sampled_absmax.restype = ctypes.c_double
        l.orc_philox4x32_10.argtypes = [vp, vp, vp]
        P = ctypes.POINTER(OrcParticles)
        l.orc_particles_create.argtypes = [ctypes.c_uint32]
        l.orc_particles_create.restype = P
        l.orc_particles_destroy.argtypes = [P]
        l.orc_set_domain.argtypes = [P] + [ctypes.c_double] * 3 + [ctypes.c_uint32] * 3 + [ctypes.c_double]
        for name in ("align_particles_to_grid", "sort_particles", "generate_voxels", "particle_vel_to_voxels",
                     "calc_div_u", "pressure_solve", "update_voxel_velocities", "voxel_vels_to_particles",
                     "advect_particles", "move_solid_particles_only", "move_solid_particles", "initialize"):
            getattr(l, "orc_" + name).argtypes = [P]
        l.orc_get_courant_dt.argtypes = [P]
        l.orc_get_courant_dt.restype = ctypes.c_double
        l.orc_apply_gravity.argtypes = [P, ctypes.c_double]
        l.orc_remove_gravity.argtypes = [P, ctypes.c_double]
        l.orc_gs_solve.argtypes = [P, ctypes.c_double, ctypes.c_double, ctypes.c_uint32, _u32p]
        l.orc_gs_solve.restype = ctypes.c_double
        l.orc_gs_iterations.argtypes = [P, ctypes.c_double, ctypes.c_uint32]
        l.orc_solve_frame.argtypes = [P, ctypes.c_double]
        l.orc_resize_voxels.argtypes = [P, ctypes.c_uint32]
        l.orc_resize_nodes.argtypes = [P, ctypes.c_uint32]
        l.orc_find_sub_cell.argtypes = [vp] * 3 + [ctypes.c_uint32, OrcGrid, vp, ctypes.c_uint32,
                                                    ctypes.c_double] + [vp] * 6
        l.orc_sum_particles_per_node.argtypes = [ctypes.c_uint32, ctypes.c_uint32, vp, vp, vp, ctypes.c_uint32, vp, vp]
        l.orc_particle_list_create.argtypes = [ctypes.c_uint32] * 3 + [vp] * 8
        l.orc_ns_tiled_index.argtypes = [ctypes.c_uint32] * 3 + [OrcGrid]
        l.orc_ns_tiled_index.restype = ctypes.c_uint32
        l.orc_ns_bin.argtypes = [vp, vp, vp, ctypes.c_uint32, OrcGrid, vp]
        l.orc_ns_p2g.argtypes = [vp] * 6 + [ctypes.c_uint32, OrcGrid] + [vp] * 6
        l.orc_ns_prepare_faces.argtypes = [vp] * 7 + [OrcGrid, ctypes.c_double]
        l.orc_ns_assemble.argtypes = [vp, vp, OrcGrid, ctypes.c_double, vp, vp, vp, vp, vp, vp]
        l.orc_ns_apply_A.argtypes = [vp, OrcGrid, vp, vp]
        l.orc_ns_pcg_jacobi.argtypes = [vp, OrcGrid, vp, vp, ctypes.c_double, ctypes.c_uint32, vp]
        l.orc_ns_pcg_jacobi.restype = ctypes.c_uint32
        l.orc_ns_velocity_update.argtypes = [vp, vp, OrcGrid, vp, ctypes.c_double, vp, vp, vp]
        l.orc_ns_g2p_flip.argtypes = [ctypes.c_uint32] + [vp] * 12 + [OrcGrid, ctypes.c_double]
        l.orc_ns_advect_rk2.argtypes = [ctypes.c_uint32] + [vp] * 7 + [OrcGrid, ctypes.c_double, ctypes.c_double]
        l.orc_ns_transfer_step.argtypes = [vp] * 6 + [ctypes.c_uint32, OrcGrid, ctypes.c_uint32] + [vp] * 8
        l.orc_pressure_retry.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_double, SOLVE_CB, HALVE_CB, vp,
                                         ctypes.POINTER(ctypes.c_double)]
        l.orc_pressure_retry.restype = ctypes.c_uint32
        l.orc_ns_work_create.argtypes = [ctypes.c_uint32, OrcGrid]
        l.orc_ns_work_create.restype = vp
        l.orc_ns_work_destroy.argtypes = [vp]
        l.orc_ns_step.argtypes = [vp] + [vp] * 6 + [OrcGrid, vp] + [ctypes.c_double] * 4 + [ctypes.c_uint32,
                                                                                        ctypes.c_double, vp]
        l.orc_ns_step.restype = ctypes.c_uint32
        _lib = l
    return _lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def find_sub_cell(px, py, pz, grid, keys, refinement_level=1, radius=2.0):
    """cudaFindSubCell: (inclusive scans of the per-particle counts, lists)
    for the x, y and z velocity components."""
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (px, py, pz)]
    k = np.ascontiguousarray(keys, dtype=np.uint32)
    n = k.size
    t = [np.empty(n, np.uint32) for _ in range(3)]
    g = OrcGrid(*grid)
    lib().orc_find_sub_cell(*(_p(v) for v in a), n, g, _p(k), refinement_level, radius, *(_p(v) for v in t),
                            None, None, None)
    lists = [np.empty(int(v[-1]) if n else 0, np.uint32) for v in t]
    if n:
        t = [np.empty(n, np.uint32) for _ in range(3)]
        lib().orc_find_sub_cell(*(_p(v) for v in a), n, g, _p(k), refinement_level, radius, *(_p(v) for v in t),
                                *(_p(v) for v in lists))
    return t, lists


def sum_particles_per_node(first, touched_incl, sub_cells, n, num_voxels_per_node=512):
    """cudaSumParticlesPerNodeAndWriteNumUsedVoxels: inclusive scans over the
    nodes of (distinct voxels, list entries) of each node's particles."""
    f = np.ascontiguousarray(first, np.uint32)
    t = np.ascontiguousarray(touched_incl, np.uint32)
    sc = np.ascontiguousarray(sub_cells, np.uint32)
    nz = np.empty(f.size, np.uint32)
    np_ = np.empty(f.size, np.uint32)
    lib().orc_sum_particles_per_node(f.size, n, _p(f), _p(t), _p(sc) if sc.size else None, num_voxels_per_node,
                                     _p(nz), _p(np_))
    return nz, np_


def particle_list_create(first, touched_incl, sub_cells, n, used_incl, lists_incl, num_voxels_per_node=512):
    """cudaParticleListCreate: (voxelIDs, per-voxel list starts, lists)."""
    f = np.ascontiguousarray(first, np.uint32)
    t = np.ascontiguousarray(touched_incl, np.uint32)
    sc = np.ascontiguousarray(sub_cells, np.uint32)
    u = np.ascontiguousarray(used_incl, np.uint32)
    li = np.ascontiguousarray(lists_incl, np.uint32)
    V = int(u[-1]) if u.size else 0
    L = int(li[-1]) if li.size else 0
    vid = np.empty(max(V, 1), np.uint32)
    start = np.empty(max(V, 1), np.uint32)
    lists = np.empty(max(L, 1), np.uint32)
    lib().orc_particle_list_create(f.size, n, num_voxels_per_node, _p(u), _p(li), _p(t), _p(sc) if sc.size else None,
                                   _p(f), _p(vid), _p(start), _p(lists))
    return vid[:V], start[:V], lists[:L]


def find_grid_cell(px, py, pz, grid):
    a = [np.array(v, dtype=np.float64, copy=True) for v in (px, py, pz)]
    keys = np.empty(a[0].size, dtype=np.uint32)
    lib().orc_find_grid_cell(_p(a[0]), _p(a[1]), _p(a[2]), a[0].size, OrcGrid(*grid), _p(keys))
    return a[0], a[1], a[2], keys


def scan_u32(values, exclusive=False):
    v = np.ascontiguousarray(values, dtype=np.uint32)
    out = np.empty_like(v)
    lib().orc_scan_u32(_p(v), _p(out), v.size, int(exclusive))
    return out


def radix_sort_u32(keys, nbits=32):
    k = np.array(keys, dtype=np.uint32, copy=True)
    perm = np.empty_like(k)
    lib().orc_radix_sort_u32(_p(k), _p(perm), k.size, nbits)
    return k, perm


def radix_sort_pairs_u32(keys, vals, nbits=32):
    k = np.array(keys, dtype=np.uint32, copy=True)
    v = np.array(vals, dtype=np.uint32, copy=True)
    lib().orc_radix_sort_pairs_u32(_p(k), _p(v), k.size, nbits)
    return k, v


def sampled_absmax(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return lib().orc_sampled_absmax(_p(a), a.size)


def philox4x32_10(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.empty(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def ns_bin(x, y, z, grid):
    a = [np.ascontiguousarray(v, dtype=np.float32) for v in (x, y, z)]
    keys = np.empty(a[0].size, dtype=np.uint32)
    lib().orc_ns_bin(_p(a[0]), _p(a[1]), _p(a[2]), a[0].size, OrcGrid(*grid), _p(keys))
    return keys


def ns_p2g(x, y, z, vx, vy, vz, grid, weights=False):
    """FP64 trilinear MAC P2G onto the brick-tiled grid: (U, V, W[, WU, WV, WW])."""
    a = [np.ascontiguousarray(v, dtype=np.float32) for v in (x, y, z, vx, vy, vz)]
    g = OrcGrid(*grid)
    cells = g.sizeX * g.sizeY * g.sizeZ
    out = [np.empty(cells) for _ in range(6 if weights else 3)]
    ptrs = [_p(o) for o in out] + ([None] * 3 if not weights else [])
    lib().orc_ns_p2g(*(_p(v) for v in a), a[0].size, g, *ptrs)
    return tuple(out)


class NsTransferStep:
    """Serial transfer step (bin, sort, gather, counts, scan, P2G) with
    preallocated scratch, for timing the CPU baseline."""

    def __init__(self, x, y, z, vx, vy, vz, grid):
        self.a = [np.ascontiguousarray(v, dtype=np.float32) for v in (x, y, z, vx, vy, vz)]
        self.g = OrcGrid(*grid)
        self.n = self.a[0].size
        cells = self.g.sizeX * self.g.sizeY * self.g.sizeZ
        self.key_bits = max(1, int(cells - 1).bit_length())
        self.keys = np.empty(self.n, np.uint32)
        self.perm = np.empty(self.n, np.uint32)
        self.sorted = np.empty(6 * self.n, np.float32)
        self.counts = np.empty(cells + 1, np.uint32)
        self.start = np.empty(cells + 1, np.uint32)
        self.out = [np.empty(cells) for _ in range(3)]

    def run(self):
        lib().orc_ns_transfer_step(*(_p(v) for v in self.a), self.n, self.g, self.key_bits, _p(self.keys),
                                   _p(self.perm), _p(self.sorted), _p(self.counts), _p(self.start),
                                   *(_p(o) for o in self.out))


class NsStep:
    """Whole serial FP64 step of the north-star method (orc_ns_step): the
    particle state (FP32 arrays held here) is replaced by the sorted, updated
    state after each run()."""

    def __init__(self, x, y, z, vx, vy, vz, grid, solid=None, dt=1.0 / 240.0, gravity=-9.81, flip=0.95, tol=1e-6,
                 max_it=100000, margin=1e-3):
        self.a = [np.array(v, dtype=np.float32, copy=True) for v in (x, y, z, vx, vy, vz)]
        self.g = OrcGrid(*grid)
        self.solid = None if solid is None else np.ascontiguousarray(solid, np.uint8)
        self.dt, self.gravity, self.flip, self.tol, self.max_it = dt, gravity, flip, tol, max_it
        self.margin = margin * self.g.cellSize
        self.w = lib().orc_ns_work_create(self.a[0].size, self.g)
        if not self.w:
            raise MemoryError("orc_ns_work_create failed")
        self.iterations, self.rel = 0, 0.0

    def run(self):
        rel = ctypes.c_double()
        self.iterations = lib().orc_ns_step(self.w, *(_p(v) for v in self.a), self.g,
                                            None if self.solid is None else _p(self.solid), self.dt, self.gravity,
                                            self.flip, self.tol, self.max_it, self.margin, ctypes.byref(rel))
        self.rel = rel.value

    def state(self):
        return [v.copy() for v in self.a]

    def close(self):
        if getattr(self, "w", None):
            lib().orc_ns_work_destroy(self.w)
            self.w = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _f64(a):
    return np.array(a, dtype=np.float64, copy=True)


def ns_prepare_faces(U, V, W, grid, gdt, solid=None):
    F = [_f64(a) for a in (U, V, W)]
    F0 = [np.empty_like(a) for a in F]
    s = None if solid is None else np.ascontiguousarray(solid, np.uint8)
    lib().orc_ns_prepare_faces(*(_p(a) for a in F), *(_p(a) for a in F0), None if s is None else _p(s),
                               OrcGrid(*grid), gdt)
    return F, F0


def ns_assemble(counts, grid, h_over_dt, U, V, W, solid=None):
    g = OrcGrid(*grid)
    cells = g.sizeX * g.sizeY * g.sizeZ
    c = np.ascontiguousarray(counts[:cells], np.uint32)
    F = [np.ascontiguousarray(a, np.float64) for a in (U, V, W)]
    ty = np.empty(cells, np.uint8)
    coef = np.empty(cells, np.uint16)
    b = np.empty(cells)
    s = None if solid is None else np.ascontiguousarray(solid, np.uint8)
    lib().orc_ns_assemble(_p(c), None if s is None else _p(s), g, h_over_dt, *(_p(a) for a in F), _p(ty), _p(coef),
                          _p(b))
    return ty, coef, b


def ns_apply_A(coef, grid, x):
    coef = np.ascontiguousarray(coef, np.uint16)
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    lib().orc_ns_apply_A(_p(coef), OrcGrid(*grid), _p(x), _p(y))
    return y


def ns_pcg_jacobi(coef, grid, b, tol=1e-12, max_it=100000):
    coef = np.ascontiguousarray(coef, np.uint16)
    b = np.ascontiguousarray(b, np.float64)
    x = np.empty_like(b)
    rel = ctypes.c_double()
    it = lib().orc_ns_pcg_jacobi(_p(coef), OrcGrid(*grid), _p(b), _p(x), tol, max_it, ctypes.byref(rel))
    return x, it, rel.value


def ns_velocity_update(types, grid, p, dt_over_h, U, V, W):
    F = [_f64(a) for a in (U, V, W)]
    lib().orc_ns_velocity_update(_p(np.ascontiguousarray(types, np.uint8)), None, OrcGrid(*grid),
                                 _p(np.ascontiguousarray(p, np.float64)), dt_over_h, *(_p(a) for a in F))
    return F


def ns_g2p_flip(x, y, z, vx, vy, vz, U, V, W, U0, V0, W0, grid, flip=0.95):
    pos = [np.ascontiguousarray(a, np.float32) for a in (x, y, z)]
    vel = [_f64(a) for a in (vx, vy, vz)]
    F = [np.ascontiguousarray(a, np.float64) for a in (U, V, W, U0, V0, W0)]
    lib().orc_ns_g2p_flip(pos[0].size, *(_p(a) for a in pos), *(_p(a) for a in vel), *(_p(a) for a in F),
                          OrcGrid(*grid), flip)
    return vel


def ns_advect_rk2(x, y, z, U, V, W, grid, dt, margin, solid=None):
    pos = [_f64(a) for a in (x, y, z)]
    F = [np.ascontiguousarray(a, np.float64) for a in (U, V, W)]
    s = None if solid is None else np.ascontiguousarray(solid, np.uint8)
    lib().orc_ns_advect_rk2(pos[0].size, *(_p(a) for a in pos), *(_p(a) for a in F), None if s is None else _p(s),
                            OrcGrid(*grid), dt, margin)
    return pos


def ns_tiled_index(i, j, k, grid):
    """Brick-tiled index of cell (i, j, k) (numpy arrays accepted)."""
    g = OrcGrid(*grid)
    nbx, nby = g.sizeX >> 3, g.sizeY >> 3
    i, j, k = (np.asarray(v, dtype=np.int64) for v in (i, j, k))
    return ((((k >> 3) * nby + (j >> 3)) * nbx + (i >> 3)) * 512 + (k & 7) * 64 + (j & 7) * 8 + (i & 7))


_RC_ARRAYS = {
    # name: (dtype, length attribute)
    "px": (np.float64, "n"), "py": (np.float64, "n"), "pz": (np.float64, "n"),
    "vx": (np.float64, "n"), "vy": (np.float64, "n"), "vz": (np.float64, "n"),
    "gridCell": (np.uint32, "n"), "reorderedGridIndices": (np.uint32, "n"),
    "uniqueGridNodeIndices": (np.uint32, "n"),
    "gridNodeIndicesToFirstParticleIndex": (np.uint32, "numUsedGridNodes"),
    "yDimNumUsedGridNodes": (np.uint32, "yDimRows"),
    "nodeIndexUsedVoxels": (np.uint32, "numUsedGridNodes"),
    "voxelIDsUsed": (np.uint32, "numUsedVoxels"), "solids": (np.int8, "numUsedVoxels"),
    "voxelsUx": (np.float64, "numUsedVoxels"), "voxelsUy": (np.float64, "numUsedVoxels"),
    "voxelsUz": (np.float64, "numUsedVoxels"), "divU": (np.float64, "numUsedVoxels"),
    "p": (np.float64, "numUsedVoxels"), "residuals": (np.float64, "numUsedVoxels"),
}


class OracleParticles:
    """The oracle's restatement of class Particles, same method names."""

    def __init__(self, size):
        self._s = lib().orc_particles_create(int(size))
        self.size = int(size)

    def close(self):
        if getattr(self, "_s", None):
            lib().orc_particles_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def state(self) -> OrcParticles:
        return self._s.contents

    def setDomain(self, nx, ny, nz, x, y, z, cellSize):
        lib().orc_set_domain(self._s, nx, ny, nz, x, y, z, cellSize)

    def __getattr__(self, name):
        table = {
            "alignParticlesToGrid": "orc_align_particles_to_grid", "sortParticles": "orc_sort_particles",
            "generateVoxels": "orc_generate_voxels", "particleVelToVoxels": "orc_particle_vel_to_voxels",
            "calcDivU": "orc_calc_div_u", "pressureSolve": "orc_pressure_solve",
            "updateVoxelVelocities": "orc_update_voxel_velocities",
            "voxelVelsToParticles": "orc_voxel_vels_to_particles", "advectParticles": "orc_advect_particles",
            "moveSolids": "orc_move_solid_particles_only", "moveSolidParticles": "orc_move_solid_particles",
            "initialize": "orc_initialize",
        }
        if name in table:
            fn = getattr(lib(), table[name])
            return lambda: fn(self._s)
        raise AttributeError(name)

    def getCourantDt(self):
        return lib().orc_get_courant_dt(self._s)

    def applyGravity(self, dt):
        lib().orc_apply_gravity(self._s, dt)

    def removeGravity(self, dt):
        lib().orc_remove_gravity(self._s, dt)

    def gsSolve(self, dt, threshold=0.01, maxIterations=16384):
        it = ctypes.c_uint32()
        r = lib().orc_gs_solve(self._s, dt, threshold, maxIterations, ctypes.byref(it))
        return r, it.value

    def gsIterations(self, dt, count):
        lib().orc_gs_iterations(self._s, dt, count)

    def solveFrame(self, fps):
        lib().orc_solve_frame(self._s, fps)

    def setSolverLimits(self, threshold=0.01, maxIterations=16384):
        s = self.state
        s.sorThreshold, s.sorMaxIterations = threshold, maxIterations

    def setTime(self, dt, frameDt, elapsedTime=0.0, elapsedTimeThisFrame=0.0):
        s = self.state
        s.dt, s.frameDt, s.elapsedTime, s.elapsedTimeThisFrame = dt, frameDt, elapsedTime, elapsedTimeThisFrame

    def _view(self, name):
        dtype, attr = _RC_ARRAYS[name]
        s = self.state
        count = getattr(s, attr)
        ptr = getattr(s, name)
        if count == 0 or not ptr:
            return np.empty(0, dtype=dtype)
        return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(np.ctypeslib.as_ctypes_type(dtype))),
                                     shape=(count,))

    def get(self, name):
        return self._view(name).copy()

    def set(self, name, values):
        v = self._view(name)
        values = np.asarray(values, dtype=v.dtype)
        assert values.size == v.size, (name, values.size, v.size)
        v[:] = values

    def set_positions(self, px, py, pz):
        for n, a in zip(("px", "py", "pz"), (px, py, pz)):
            self.set(n, a)

    def set_velocities(self, vx, vy, vz):
        for n, a in zip(("vx", "vy", "vz"), (vx, vy, vz)):
            self.set(n, a)

    def positions(self):
        return [self.get(n) for n in ("px", "py", "pz")]

    def velocities(self):
        return [self.get(n) for n in ("vx", "vy", "vz")]

    def stats(self):
        s = self.state
        return {"numUsedGridNodes": s.numUsedGridNodes, "numUsedVoxels": s.numUsedVoxels,
                "lastSorIterations": s.lastSorIterations, "lastSolveAttempts": s.lastSolveAttempts,
                "lastResidual": s.lastResidual, "dt": s.dt, "frameDt": s.frameDt,
                "elapsedTime": s.elapsedTime, "elapsedTimeThisFrame": s.elapsedTimeThisFrame,
                "outOfTileEvents": s.outOfTileEvents, "relocatedParticles": s.relocatedParticles}
