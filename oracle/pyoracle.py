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
        _lib = l
    return _lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


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
