"""Python bindings of libparticles.so, the MI355X build of flip2's per-step
particle pipeline (binning, radix sort, scans, sparse-voxel P2G, pressure
solve, G2P/advection).

The library is C++/HIP behind the C ABI in include/flip2.h; this module only
marshals numpy arrays through ctypes.  There is no fallback: if the built
library is missing or the GPU is absent, calls raise.

Reference interfaces mirrored here (paths relative to the reference repo):
  - class Particles (engine/particles.hu:11-45)
  - cudaParallelPrefixSum (engine/algorithms/parallelPrefixSumKernels.hu:53-54)
  - cudaRadixSortUint (engine/algorithms/radixSortKernels.hu:16-17)
  - the ctypes load of libparticles.so in plugin/plugin.py:4-5
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REPO_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(REPO_ROOT, "build", "engine", "libparticles.so")
HEADER_PATH = os.path.join(REPO_ROOT, "include", "flip2.h")

_lib = None


class Flip2Error(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"flip2 error {code}: {message}")
        self.code = code


class GridC(ctypes.Structure):
    """flip2_grid: same layout as the reference's Grid (engine/grid.hpp:35-37)."""

    _fields_ = [
        ("negX", ctypes.c_float),
        ("negY", ctypes.c_float),
        ("negZ", ctypes.c_float),
        ("sizeX", ctypes.c_uint32),
        ("sizeY", ctypes.c_uint32),
        ("sizeZ", ctypes.c_uint32),
        ("cellSize", ctypes.c_float),
    ]


class FluidConfigC(ctypes.Structure):
    _fields_ = [
        ("grid", GridC),
        ("dt", ctypes.c_float),
        ("gravity", ctypes.c_float),
        ("flip", ctypes.c_float),
        ("preconditioner", ctypes.c_int),
        ("tolerance", ctypes.c_double),
        ("max_iterations", ctypes.c_uint32),
        ("check_every", ctypes.c_uint32),
        ("margin", ctypes.c_float),
    ]


class SolveStatsC(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_uint32),
        ("refinements", ctypes.c_uint32),
        ("rel_residual", ctypes.c_double),
        ("b_norm", ctypes.c_double),
        ("active_bricks", ctypes.c_uint32),
        ("fluid_cells", ctypes.c_uint32),
        ("solve_ms", ctypes.c_double),
        ("converged", ctypes.c_int),
        ("used_fallback", ctypes.c_int),
    ]


PREC = {"jacobi": 0, "mic0": 1, "learned": 2}


class RcStatsC(ctypes.Structure):
    _fields_ = [
        ("numUsedGridNodes", ctypes.c_uint32),
        ("numUsedVoxels", ctypes.c_uint32),
        ("lastSorIterations", ctypes.c_uint32),
        ("lastSolveAttempts", ctypes.c_uint32),
        ("lastResidual", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("frameDt", ctypes.c_double),
        ("elapsedTime", ctypes.c_double),
        ("elapsedTimeThisFrame", ctypes.c_double),
        ("outOfTileEvents", ctypes.c_uint32),
        ("relocatedParticles", ctypes.c_uint32),
        ("substeps", ctypes.c_uint32),
    ]


def lib() -> ctypes.CDLL:
    """Load build/engine/libparticles.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing; build it with `make -C flip2_amd/csrc` "
                "or __graft_entry__.build()")
        l = ctypes.CDLL(LIB_PATH)
        l.flip2_last_error.restype = ctypes.c_char_p
        l.flip2_scan_workspace_bytes.restype = ctypes.c_size_t
        l.flip2_sort_workspace_bytes.restype = ctypes.c_size_t
        l.flip2_scan_workspace_bytes.argtypes = [ctypes.c_uint32]
        l.flip2_sort_workspace_bytes.argtypes = [ctypes.c_uint32] * 3
        _declare(l)
        _lib = l
    return _lib


_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_f64 = ctypes.c_double
_sz = ctypes.c_size_t


def _declare(l: ctypes.CDLL) -> None:
    sig = {
        "flip2_malloc": [ctypes.POINTER(_vp), _sz],
        "flip2_free": [_vp],
        "flip2_memcpy_htod": [_vp, _vp, _sz],
        "flip2_memcpy_dtoh": [_vp, _vp, _sz],
        "flip2_memcpy_dtod": [_vp, _vp, _sz],
        "flip2_memset": [_vp, ctypes.c_int, _sz],
        "flip2_copy": [_vp, _vp, _sz, _vp],
        "flip2_stream_create": [ctypes.POINTER(_vp)],
        "flip2_stream_destroy": [_vp],
        "flip2_stream_synchronize": [_vp],
        "flip2_timer_create": [ctypes.POINTER(_vp)],
        "flip2_timer_destroy": [_vp],
        "flip2_timer_start": [_vp, _vp],
        "flip2_timer_stop": [_vp, _vp],
        "flip2_timer_elapsed_ms": [_vp, ctypes.POINTER(ctypes.c_float)],
        "flip2_scan_u32": [_vp, _vp, _u32, ctypes.c_int, _vp, _vp, _sz, _vp],
        "flip2_sort_pairs_u32": [_vp, _vp, _vp, _vp, _vp, _vp, _u32, _u32, _u32, _vp, _sz, _vp],
        "flip2_find_grid_cell": [_vp, _vp, _vp, _u32, GridC, _vp, _vp],
        "flip2_find_sub_cell": [_vp, _vp, _vp, _u32, GridC, _vp, _u32, _f64] + [_vp] * 9,
        "flip2_sum_particles_per_node": [_u32, _u32] + [_vp] * 6,
        "flip2_particle_list_create": [_u32, _u32] + [_vp] * 9,
        "flip2_philox4x32_10": [_vp, _u32, _vp],
        "flip2_particles_create": [_u32, ctypes.POINTER(_vp)],
        "flip2_particles_destroy": [_vp],
        "flip2_set_domain": [_vp, _f64, _f64, _f64, _u32, _u32, _u32, _f64],
        "flip2_set_positions": [_vp, _vp, _vp, _vp],
        "flip2_set_velocities": [_vp, _vp, _vp, _vp],
        "flip2_get_positions": [_vp, _vp, _vp, _vp],
        "flip2_get_velocities": [_vp, _vp, _vp, _vp],
        "flip2_get_courant_dt": [_vp, ctypes.POINTER(_f64)],
        "flip2_apply_gravity": [_vp, _f64],
        "flip2_remove_gravity": [_vp, _f64],
        "flip2_gs_solve": [_vp, _f64, _f64, _u32, ctypes.POINTER(_f64), ctypes.POINTER(_u32)],
        "flip2_gs_iterations": [_vp, _f64, _u32],
        "flip2_solve_frame": [_vp, _f64],
        "flip2_write_positions_to_file": [_vp, ctypes.c_char_p],
        "flip2_get_stats": [_vp, ctypes.POINTER(RcStatsC)],
        "flip2_get_array": [_vp, ctypes.c_char_p, _vp, _sz, ctypes.POINTER(_sz)],
        "flip2_set_array": [_vp, ctypes.c_char_p, _vp, _sz],
        "flip2_set_time": [_vp, _f64, _f64, _f64, _f64],
        "flip2_set_solver_limits": [_vp, _f64, _u32],
        "flip2_ns_bin": [_vp, _vp, _vp, _u32, GridC, _vp, _vp],
        "flip2_ns_p2g": [_vp, _vp, GridC] + [_vp] * 7,
        "flip2_p2g_pipeline_create": [_u32, GridC, ctypes.POINTER(_vp)],
        "flip2_p2g_pipeline_destroy": [_vp],
        "flip2_p2g_pipeline_run": [_vp, _u32] + [_vp] * 10,
        "flip2_p2g_pipeline_get": [_vp, ctypes.c_char_p, _vp, _sz, ctypes.POINTER(_sz)],
        "flip2_p2g_pipeline_enable_timing": [_vp, ctypes.c_int],
        "flip2_p2g_pipeline_run_records": [_vp, _u32, _vp, _vp, _vp, _vp, _vp],
        "flip2_fluid_create": [_u32, ctypes.POINTER(FluidConfigC), ctypes.POINTER(_vp)],
        "flip2_fluid_destroy": [_vp],
        "flip2_fluid_set_config": [_vp, ctypes.POINTER(FluidConfigC)],
        "flip2_fluid_set_particles": [_vp, _u32] + [_vp] * 6,
        "flip2_fluid_get_particles": [_vp] * 7,
        "flip2_fluid_set_solid": [_vp, _vp],
        "flip2_fluid_set_solid_box": [_vp, _vp, _vp, _vp],
        "flip2_fluid_num_particles": [_vp, _vp],
        "flip2_fluid_get_config": [_vp, ctypes.POINTER(FluidConfigC)],
        "flip2_fluid_step_frame": [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_uint32, _vp],
        "flip2_fluid_save_snapshot": [_vp, ctypes.c_char_p],
        "flip2_fluid_load_snapshot": [_vp, ctypes.c_char_p],
        "flip2_fluid_step": [_vp],
        "flip2_fluid_stage": [_vp, ctypes.c_char_p],
        "flip2_fluid_solve_with": [_vp, ctypes.c_int, ctypes.POINTER(SolveStatsC)],
        "flip2_fluid_last_solve": [_vp, ctypes.POINTER(SolveStatsC)],
        "flip2_fluid_get_array": [_vp, ctypes.c_char_p, _vp, _sz, ctypes.POINTER(_sz)],
        "flip2_p2g_pipeline_timing": [_vp, ctypes.c_int, ctypes.c_char_p, _sz, ctypes.POINTER(_f64),
                                      ctypes.POINTER(ctypes.c_int)],
        "flip2_unet_create": [ctypes.c_int, ctypes.POINTER(_vp)],
        "flip2_unet_destroy": [_vp],
        "flip2_unet_num_params": [_vp, ctypes.POINTER(_sz)],
        "flip2_unet_init": [_vp, ctypes.c_uint64],
        "flip2_unet_set_params": [_vp, _vp, _sz],
        "flip2_unet_get_params": [_vp, _vp, _sz],
        "flip2_unet_apply_dense": [_vp, _vp, _vp, _vp],
        "flip2_unet_loss_grad": [_vp, _vp, _vp, ctypes.POINTER(_f64), _vp],
        "flip2_unet_add_frame": [_vp, _vp],
        "flip2_unet_set_smoothing": [_vp, ctypes.c_float],
        "flip2_unet_num_frames": [_vp, ctypes.POINTER(ctypes.c_int)],
        "flip2_unet_capture_residuals": [_vp, _vp, _vp, ctypes.c_int],
        "flip2_unet_clear_frames": [_vp],
        "flip2_unet_train": [_vp, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                             ctypes.POINTER(_f64)],
        "flip2_unet_eval": [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_f64)],
        "flip2_fluid_set_unet": [_vp, _vp, _u32],
        "flip2_fluid_set_flex_directions": [_vp, _u32],
        "flip2_fluid_set_flex_check_every": [_vp, _u32],
        "flip2_fluid_set_learned_graph": [_vp, ctypes.c_int],
        "flip2_fluid_set_learned_smoother": [_vp, ctypes.c_int],
        "flip2_fluid_set_learned_smooth_steps": [_vp, ctypes.c_uint32],
        "flip2_fluid_enable_timing": [_vp, ctypes.c_char_p],
        "flip2_fluid_set_timing_sample": [_vp, ctypes.c_uint32],
        "flip2_fluid_disable_timing": [_vp],
        "flip2_fluid_timing": [_vp, ctypes.c_int, ctypes.c_char_p, _sz, ctypes.POINTER(_f64),
                               ctypes.POINTER(ctypes.c_uint64)],
    }
    for name in ("align_particles_to_grid", "sort_particles", "generate_voxels", "particle_vel_to_voxels",
                 "calc_div_u", "pressure_solve", "update_voxel_velocities", "voxel_vels_to_particles",
                 "advect_particles", "move_solids", "move_solid_particles", "initialize"):
        sig["flip2_" + name] = [_vp]
    for name, args in sig.items():
        fn = getattr(l, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int


def check(rc: int) -> None:
    if rc != 0:
        raise Flip2Error(rc, lib().flip2_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


# --------------------------------------------------------------------------
# Device memory
# --------------------------------------------------------------------------
class DeviceBuffer:
    """Device allocation owned from Python (flip2_malloc / flip2_free)."""

    def __init__(self, nbytes: int):
        p = _vp()
        check(lib().flip2_malloc(ctypes.byref(p), max(int(nbytes), 1)))
        self.ptr = p.value
        self.nbytes = int(nbytes)

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            check(lib().flip2_memcpy_htod(b.ptr, _ptr(a), a.nbytes))
        return b

    def to_array(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            check(lib().flip2_memcpy_dtoh(_ptr(out), self.ptr, out.nbytes))
        return out

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if a.nbytes:
            check(lib().flip2_memcpy_htod(self.ptr, _ptr(a), a.nbytes))

    def free(self) -> None:
        if self.ptr:
            lib().flip2_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize(stream: Optional[int] = None) -> None:
    check(lib().flip2_stream_synchronize(stream))


# --------------------------------------------------------------------------
# Primitives (host-array conveniences over the device entry points)
# --------------------------------------------------------------------------
def scan_u32(values: np.ndarray, exclusive: bool = False) -> Tuple[np.ndarray, int]:
    """Prefix scan mod 2^32 on the GPU; returns (scan, total)."""
    values = np.ascontiguousarray(values, dtype=np.uint32)
    n = values.size
    d_in = DeviceBuffer.from_array(values)
    d_out = DeviceBuffer(4 * n)
    d_tot = DeviceBuffer(4)
    ws_bytes = lib().flip2_scan_workspace_bytes(n)
    d_ws = DeviceBuffer(ws_bytes)
    check(lib().flip2_scan_u32(d_in.ptr, d_out.ptr, n, int(exclusive), d_tot.ptr, d_ws.ptr, ws_bytes, None))
    synchronize()
    return d_out.to_array(np.uint32, n), int(d_tot.to_array(np.uint32, 1)[0]) if n else 0


def sort_pairs_u32(keys: np.ndarray, values: Optional[np.ndarray] = None, begin_bit: int = 0,
                   end_bit: int = 32) -> Tuple[np.ndarray, np.ndarray]:
    """Stable radix sort on the GPU; values=None returns the permutation."""
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    n = keys.size
    d_k = DeviceBuffer.from_array(keys)
    d_v = DeviceBuffer.from_array(np.ascontiguousarray(values, dtype=np.uint32)) if values is not None else None
    d_ko, d_vo, d_ka, d_va = (DeviceBuffer(4 * n) for _ in range(4))
    ws_bytes = lib().flip2_sort_workspace_bytes(n, begin_bit, end_bit)
    d_ws = DeviceBuffer(ws_bytes)
    check(lib().flip2_sort_pairs_u32(d_k.ptr, d_v.ptr if d_v else None, d_ko.ptr, d_vo.ptr, d_ka.ptr, d_va.ptr, n,
                                     begin_bit, end_bit, d_ws.ptr, ws_bytes, None))
    synchronize()
    return d_ko.to_array(np.uint32, n), d_vo.to_array(np.uint32, n)


def find_grid_cell(px, py, pz, grid: Sequence) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """rootCell on the GPU: returns clamped (px, py, pz) and keys."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (px, py, pz)]
    n = arrs[0].size
    d = [DeviceBuffer.from_array(a) for a in arrs]
    dk = DeviceBuffer(4 * n)
    g = GridC(*grid)
    check(lib().flip2_find_grid_cell(d[0].ptr, d[1].ptr, d[2].ptr, n, g, dk.ptr, None))
    synchronize()
    return tuple(b.to_array(np.float64, n) for b in d) + (dk.to_array(np.uint32, n),)


def find_sub_cell(px, py, pz, grid: Sequence, keys, refinement_level: int = 1, radius: float = 2.0):
    """cudaFindSubCell on the GPU: ([inclusive scans of the per-particle
    counts] * 3, [sub-cell lists] * 3) for the x, y, z components."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (px, py, pz)]
    k = np.ascontiguousarray(keys, dtype=np.uint32)
    n = k.size
    d = [DeviceBuffer.from_array(a) for a in arrs]
    dk = DeviceBuffer.from_array(k)
    dt = [DeviceBuffer(4 * n) for _ in range(3)]
    totals = np.zeros(3, np.uint32)
    g = GridC(*grid)
    check(lib().flip2_find_sub_cell(*(b.ptr for b in d), n, g, dk.ptr, refinement_level, radius,
                                    *(b.ptr for b in dt), None, None, None, None, _ptr(totals), None))
    caps = totals.copy()
    dl = [DeviceBuffer(4 * max(int(t), 1)) for t in totals]
    check(lib().flip2_find_sub_cell(*(b.ptr for b in d), n, g, dk.ptr, refinement_level, radius,
                                    *(b.ptr for b in dt), *(b.ptr for b in dl), _ptr(caps), _ptr(totals), None))
    return ([b.to_array(np.uint32, n) for b in dt], [b.to_array(np.uint32, int(t)) for b, t in zip(dl, totals)])


def sum_particles_per_node(first, touched_incl, sub_cells, n: int):
    """cudaSumParticlesPerNodeAndWriteNumUsedVoxels on the GPU: inclusive scans
    over the nodes of (distinct voxels, list entries)."""
    f = np.ascontiguousarray(first, np.uint32)
    B = f.size
    df, dt = DeviceBuffer.from_array(f), DeviceBuffer.from_array(np.ascontiguousarray(touched_incl, np.uint32))
    dl = DeviceBuffer.from_array(np.ascontiguousarray(sub_cells, np.uint32))
    dnz, dne = DeviceBuffer(4 * B), DeviceBuffer(4 * B)
    check(lib().flip2_sum_particles_per_node(B, n, df.ptr, dt.ptr, dl.ptr, dnz.ptr, dne.ptr, None))
    return dnz.to_array(np.uint32, B), dne.to_array(np.uint32, B)


def particle_list_create(first, touched_incl, sub_cells, n: int, nonzero_incl, entries_incl):
    """cudaParticleListCreate on the GPU: (voxel ids, list starts, lists)."""
    f = np.ascontiguousarray(first, np.uint32)
    B = f.size
    nz = np.ascontiguousarray(nonzero_incl, np.uint32)
    ne = np.ascontiguousarray(entries_incl, np.uint32)
    V, L = (int(nz[-1]), int(ne[-1])) if B else (0, 0)
    bufs = [DeviceBuffer.from_array(a) for a in (nz, ne, np.ascontiguousarray(touched_incl, np.uint32),
                                                 np.ascontiguousarray(sub_cells, np.uint32), f)]
    dv, ds, dp = DeviceBuffer(4 * max(V, 1)), DeviceBuffer(4 * max(V, 1)), DeviceBuffer(4 * max(L, 1))
    check(lib().flip2_particle_list_create(B, n, *(b.ptr for b in bufs), dv.ptr, ds.ptr, dp.ptr, None))
    return dv.to_array(np.uint32, V), ds.to_array(np.uint32, V), dp.to_array(np.uint32, L)


def philox4x32_10(ctr_key: np.ndarray) -> np.ndarray:
    """Philox4x32-10 on the GPU for rows (c0, c1, c2, c3, k0, k1)."""
    a = np.ascontiguousarray(ctr_key, dtype=np.uint32).reshape(-1, 6)
    out = np.empty((a.shape[0], 4), dtype=np.uint32)
    check(lib().flip2_philox4x32_10(_ptr(a), a.shape[0], _ptr(out)))
    return out


# --------------------------------------------------------------------------
# Particle -> grid transfer on 8^3 bricks (north-star method)
# --------------------------------------------------------------------------
_PIPE_DTYPES = {"keys": np.uint32, "sortedKeys": np.uint32, "perm": np.uint32, "counts": np.uint32,
                "cellStart": np.uint32, "x": np.float32, "y": np.float32, "z": np.float32,
                "vx": np.float32, "vy": np.float32, "vz": np.float32}


class P2GPipeline:
    """bin -> key/value radix sort -> gather -> per-cell counts -> exclusive
    scan -> trilinear MAC P2G, on a brick-tiled grid (sizes multiples of 8)."""

    def __init__(self, max_particles: int, grid: Sequence):
        h = _vp()
        self.grid = GridC(*grid)
        check(lib().flip2_p2g_pipeline_create(int(max_particles), self.grid, ctypes.byref(h)))
        self._h = h.value
        self.num_cells = self.grid.sizeX * self.grid.sizeY * self.grid.sizeZ

    def close(self):
        if getattr(self, "_h", None):
            lib().flip2_p2g_pipeline_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, n: int, state: Sequence[DeviceBuffer], out: Sequence[DeviceBuffer], stream=None):
        """state: device x, y, z, vx, vy, vz (FP32); out: device U, V, W."""
        check(lib().flip2_p2g_pipeline_run(self._h, int(n), *(b.ptr for b in state), *(b.ptr for b in out), stream))

    def run_records(self, n: int, records: DeviceBuffer, out: Sequence[DeviceBuffer], stream=None):
        """records: device particle records (8 floats each); out: device U, V, W."""
        check(lib().flip2_p2g_pipeline_run_records(self._h, int(n), records.ptr, *(b.ptr for b in out), stream))

    def get(self, name: str) -> np.ndarray:
        cnt = _sz()
        check(lib().flip2_p2g_pipeline_get(self._h, name.encode(), None, 0, ctypes.byref(cnt)))
        out = np.empty(cnt.value, dtype=_PIPE_DTYPES[name])
        check(lib().flip2_p2g_pipeline_get(self._h, name.encode(), _ptr(out) if out.size else None, out.size,
                                           ctypes.byref(cnt)))
        return out

    def enable_timing(self, capacity: int) -> None:
        check(lib().flip2_p2g_pipeline_enable_timing(self._h, int(capacity)))

    def timing(self) -> Tuple[dict, int]:
        """({kernel: total ms}, timed runs)."""
        res = {}
        runs = ctypes.c_int()
        k = 0
        while True:
            name = ctypes.create_string_buffer(64)
            ms = _f64()
            rc = lib().flip2_p2g_pipeline_timing(self._h, k, name, 64, ctypes.byref(ms), ctypes.byref(runs))
            if rc != 0:
                break
            res[name.value.decode()] = ms.value
            k += 1
        return res, runs.value


def transfer(x, y, z, vx, vy, vz, grid: Sequence):
    """Host convenience: run the transfer pipeline once; returns (U, V, W, pipeline)."""
    arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (x, y, z, vx, vy, vz)]
    n = arrs[0].size
    pl = P2GPipeline(max(n, 1), grid)
    state = [DeviceBuffer.from_array(a) for a in arrs]
    out = [DeviceBuffer(4 * pl.num_cells) for _ in range(3)]
    pl.run(n, state, out)
    synchronize()
    return tuple(b.to_array(np.float32, pl.num_cells) for b in out) + (pl,)


# --------------------------------------------------------------------------
# North-star FLIP step
# --------------------------------------------------------------------------
_FLUID_DTYPES = {"U": np.float32, "V": np.float32, "W": np.float32, "U0": np.float32, "V0": np.float32,
                 "W0": np.float32, "counts": np.uint32, "cellStart": np.uint32, "activeBricks": np.uint32,
                 "brickSlot": np.int32, "cellType": np.uint8, "coef": np.uint16, "b": np.float64, "p": np.float64}


def _stats_dict(s: SolveStatsC) -> dict:
    return {f: getattr(s, f) for f, _ in SolveStatsC._fields_}


class FluidSim:
    """Mixed-precision FLIP step on 8^3 bricks (see include/flip2.h)."""

    def __init__(self, max_particles: int, grid: Sequence, dt=1.0 / 240.0, gravity=-9.81, flip=0.95,
                 preconditioner="jacobi", tolerance=1e-6, max_iterations=4000, check_every=8, margin=1e-3):
        cfg = FluidConfigC()
        cfg.grid = GridC(*grid)
        cfg.dt, cfg.gravity, cfg.flip = dt, gravity, flip
        cfg.preconditioner = PREC[preconditioner] if isinstance(preconditioner, str) else int(preconditioner)
        cfg.tolerance, cfg.max_iterations, cfg.check_every, cfg.margin = tolerance, max_iterations, check_every, margin
        self.cfg = cfg
        h = _vp()
        check(lib().flip2_fluid_create(int(max_particles), ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h.value
        self.grid = tuple(grid)
        self.n = 0

    def close(self):
        if getattr(self, "_h", None):
            lib().flip2_fluid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_preconditioner(self, name):
        self.cfg.preconditioner = PREC[name] if isinstance(name, str) else int(name)
        check(lib().flip2_fluid_set_config(self._h, ctypes.byref(self.cfg)))

    def set_particles(self, x, y, z, vx, vy, vz):
        a = [np.ascontiguousarray(v, dtype=np.float32) for v in (x, y, z, vx, vy, vz)]
        self.n = a[0].size
        check(lib().flip2_fluid_set_particles(self._h, self.n, *(_ptr(v) for v in a)))

    def num_particles(self) -> int:
        n = ctypes.c_uint32()
        check(lib().flip2_fluid_num_particles(self._h, ctypes.byref(n)))
        return int(n.value)

    def particles(self):
        # sized by the library's own count (a snapshot load changes it)
        self.n = self.num_particles()
        out = [np.empty(self.n, np.float32) for _ in range(6)]
        check(lib().flip2_fluid_get_particles(self._h, *(_ptr(v) for v in out)))
        return out

    def set_solid(self, mask):
        if mask is None:
            check(lib().flip2_fluid_set_solid(self._h, None))
        else:
            m = np.ascontiguousarray(mask, dtype=np.uint8)
            check(lib().flip2_fluid_set_solid(self._h, _ptr(m)))

    def set_solid_box(self, lo=None, hi=None, vel=(0.0, 0.0, 0.0)):
        """Moving solid box [lo, hi) with face velocity vel; lo=None removes it."""
        if lo is None:
            check(lib().flip2_fluid_set_solid_box(self._h, None, None, None))
            return
        a = [np.ascontiguousarray(v, dtype=np.float32) for v in (lo, hi, vel)]
        check(lib().flip2_fluid_set_solid_box(self._h, *(_ptr(v) for v in a)))

    def step_frame(self, frame_dt: float, cfl: float = 0.5, max_substeps: int = 64) -> int:
        """Advance by frame_dt in CFL-limited substeps; returns their number."""
        k = ctypes.c_uint32()
        check(lib().flip2_fluid_step_frame(self._h, float(frame_dt), float(cfl), int(max_substeps),
                                           ctypes.byref(k)))
        return int(k.value)

    def set_dt(self, dt: float) -> None:
        """Fixed step of step() (the configuration's dt)."""
        self.cfg.dt = dt
        check(lib().flip2_fluid_set_config(self._h, ctypes.byref(self.cfg)))

    def save_snapshot(self, path: str) -> None:
        """Write the state needed to resume the run (see include/flip2.h)."""
        check(lib().flip2_fluid_save_snapshot(self._h, os.fspath(path).encode()))

    def load_snapshot(self, path: str) -> None:
        """Restore a state written by save_snapshot (same grid dimensions)."""
        check(lib().flip2_fluid_load_snapshot(self._h, os.fspath(path).encode()))
        self.n = self.num_particles()
        # the restored settings replace this wrapper's copy (set_dt and
        # set_preconditioner write the whole configuration back)
        check(lib().flip2_fluid_get_config(self._h, ctypes.byref(self.cfg)))

    def step(self):
        check(lib().flip2_fluid_step(self._h))

    def stage(self, name: str):
        check(lib().flip2_fluid_stage(self._h, name.encode()))

    def solve_with(self, preconditioner) -> dict:
        s = SolveStatsC()
        p = PREC[preconditioner] if isinstance(preconditioner, str) else int(preconditioner)
        check(lib().flip2_fluid_solve_with(self._h, p, ctypes.byref(s)))
        return _stats_dict(s)

    def last_solve(self) -> dict:
        s = SolveStatsC()
        check(lib().flip2_fluid_last_solve(self._h, ctypes.byref(s)))
        return _stats_dict(s)

    def set_unet(self, net: Optional[UNet], fallback_iterations: int = 0) -> None:
        """Attach `net` as the "learned" preconditioner (None detaches)."""
        check(lib().flip2_fluid_set_unet(self._h, net._h if net is not None else None, int(fallback_iterations)))
        self._unet = net

    def set_learned_smoother(self, smoother: str) -> None:
        """Smoothing around the learned correction: "jacobi" (damped, the
        network's weight) or "ic0" (brick-local incomplete Cholesky)."""
        kinds = {"jacobi": 0, "ic0": 1}
        if smoother not in kinds:
            raise ValueError(f"unknown smoother {smoother!r} (jacobi or ic0)")
        check(lib().flip2_fluid_set_learned_smoother(self._h, kinds[smoother]))

    def set_learned_smooth_steps(self, steps: int) -> None:
        """IC(0) smoothing steps (1 or 2) before and after the network."""
        check(lib().flip2_fluid_set_learned_smooth_steps(self._h, int(steps)))

    def set_learned_graph(self, enable: bool) -> None:
        """Replay the learned preconditioner from a HIP graph (default) or not."""
        check(lib().flip2_fluid_set_learned_graph(self._h, 1 if enable else 0))

    def set_flex_directions(self, directions: int) -> None:
        """Earlier directions (1..16) flexible CG orthogonalises against."""
        check(lib().flip2_fluid_set_flex_directions(self._h, int(directions)))

    def set_flex_check_every(self, iterations: int) -> None:
        """Iterations (1..1024) between residual checks in flexible CG."""
        check(lib().flip2_fluid_set_flex_check_every(self._h, int(iterations)))

    def enable_timing(self, kernels: Optional[Sequence[str]] = None, sample_every: int = 1) -> None:
        """Time every launch (or those of the named kernels) of the following
        steps, stages and solves with HIP events on the simulation's stream;
        sample_every > 1 records one launch in that many of each kernel."""
        arg = ",".join(kernels).encode() if kernels else None
        check(lib().flip2_fluid_enable_timing(self._h, arg))
        if sample_every != 1:
            check(lib().flip2_fluid_set_timing_sample(self._h, int(sample_every)))

    def disable_timing(self) -> None:
        check(lib().flip2_fluid_disable_timing(self._h))

    def timing(self) -> dict:
        """{kernel: (total ms, launches)} since enable_timing; the instances of
        a kernel template (name<...>) are counted under the kernel's name."""
        out, k = {}, 0
        while True:
            name = ctypes.create_string_buffer(128)
            ms, cnt = _f64(), ctypes.c_uint64()
            rc = lib().flip2_fluid_timing(self._h, k, name, 128, ctypes.byref(ms), ctypes.byref(cnt))
            if rc != 0:
                break
            key = name.value.decode().split("<", 1)[0]
            prev = out.get(key, (0.0, 0))
            out[key] = (prev[0] + ms.value, prev[1] + int(cnt.value))
            k += 1
        return out

    def get(self, name: str) -> np.ndarray:
        cnt = _sz()
        check(lib().flip2_fluid_get_array(self._h, name.encode(), None, 0, ctypes.byref(cnt)))
        out = np.empty(cnt.value, dtype=_FLUID_DTYPES[name])
        check(lib().flip2_fluid_get_array(self._h, name.encode(), _ptr(out) if out.size else None, out.size,
                                          ctypes.byref(cnt)))
        return out


class UNet:
    """Learned preconditioner network on a dense n^3 grid (see include/flip2.h)."""

    def __init__(self, n: int, seed: Optional[int] = None):
        h = _vp()
        check(lib().flip2_unet_create(int(n), ctypes.byref(h)))
        self._h = h.value
        self.n = int(n)
        cnt = _sz()
        check(lib().flip2_unet_num_params(self._h, ctypes.byref(cnt)))
        self.num_params = cnt.value
        if seed is not None:
            check(lib().flip2_unet_init(self._h, int(seed)))

    def close(self):
        if getattr(self, "_h", None):
            lib().flip2_unet_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self, seed: int) -> None:
        check(lib().flip2_unet_init(self._h, int(seed)))

    def params(self) -> np.ndarray:
        out = np.empty(self.num_params, np.float32)
        check(lib().flip2_unet_get_params(self._h, _ptr(out), out.size))
        return out

    def set_params(self, p) -> None:
        a = np.ascontiguousarray(p, dtype=np.float32)
        check(lib().flip2_unet_set_params(self._h, _ptr(a), a.size))

    def _dense(self, b, coef):
        nv = self.n ** 3
        bb = np.ascontiguousarray(b, dtype=np.float32).reshape(-1)
        cc = np.ascontiguousarray(coef, dtype=np.uint16).reshape(-1)
        if bb.size != nv or cc.size != nv:
            raise ValueError(f"expected {nv} values")
        return bb, cc

    def apply_dense(self, b, coef) -> np.ndarray:
        bb, cc = self._dense(b, coef)
        z = np.empty(bb.size, np.float32)
        check(lib().flip2_unet_apply_dense(self._h, _ptr(bb), _ptr(cc), _ptr(z)))
        return z

    def loss_grad(self, b, coef, grad: bool = True):
        bb, cc = self._dense(b, coef)
        loss = _f64()
        g = np.empty(self.num_params, np.float32) if grad else None
        check(lib().flip2_unet_loss_grad(self._h, _ptr(bb), _ptr(cc), ctypes.byref(loss),
                                         _ptr(g) if grad else None))
        return loss.value, g

    def set_smoothing(self, omega: float) -> None:
        check(lib().flip2_unet_set_smoothing(self._h, float(omega)))

    def add_frame(self, sim: "FluidSim") -> None:
        check(lib().flip2_unet_add_frame(self._h, sim._h))

    def capture_residuals(self, sim: Optional["FluidSim"], iterations=()) -> None:
        """Add the residuals after the given inner iteration counts of every
        following solve of `sim` as training frames (empty stops)."""
        its = np.ascontiguousarray(iterations, dtype=np.uint32)
        check(lib().flip2_unet_capture_residuals(self._h, sim._h, _ptr(its) if its.size else None, int(its.size)))
        sim._capture_net = self if its.size else None

    def num_frames(self) -> int:
        c = ctypes.c_int()
        check(lib().flip2_unet_num_frames(self._h, ctypes.byref(c)))
        return c.value

    def clear_frames(self) -> None:
        check(lib().flip2_unet_clear_frames(self._h))

    def train(self, epochs: int, lr: float, first: int = 0, last: Optional[int] = None, seed: int = 0) -> float:
        last = self.num_frames() if last is None else last
        loss = _f64()
        check(lib().flip2_unet_train(self._h, int(epochs), float(lr), int(first), int(last), int(seed),
                                     ctypes.byref(loss)))
        return loss.value

    def eval(self, first: int = 0, last: Optional[int] = None) -> float:
        last = self.num_frames() if last is None else last
        loss = _f64()
        check(lib().flip2_unet_eval(self._h, int(first), int(last), ctypes.byref(loss)))
        return loss.value


# --------------------------------------------------------------------------
# Reference-compatible particle system
# --------------------------------------------------------------------------
_ARRAY_DTYPES = {
    "px": np.float64, "py": np.float64, "pz": np.float64,
    "vx": np.float64, "vy": np.float64, "vz": np.float64,
    "gridCell": np.uint32, "reorderedGridIndices": np.uint32, "uniqueGridNodeIndices": np.uint32,
    "gridNodeIndicesToFirstParticleIndex": np.uint32, "yDimNumUsedGridNodes": np.uint32,
    "nodeIndexUsedVoxels": np.uint32, "voxelIDsUsed": np.uint32, "solids": np.int8,
    "voxelsUx": np.float64, "voxelsUy": np.float64, "voxelsUz": np.float64,
    "divU": np.float64, "p": np.float64, "residuals": np.float64,
}


class Particles:
    """class Particles (engine/particles.hu:11-45) backed by the HIP build."""

    def __init__(self, size: int):
        h = _vp()
        check(lib().flip2_particles_create(int(size), ctypes.byref(h)))
        self._h = h.value
        self.size = int(size)

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().flip2_particles_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # particles.hu:11-45 -------------------------------------------------
    def setDomain(self, nx, ny, nz, x, y, z, cellSize):
        check(lib().flip2_set_domain(self._h, nx, ny, nz, x, y, z, cellSize))

    def alignParticlesToGrid(self):
        check(lib().flip2_align_particles_to_grid(self._h))

    def sortParticles(self):
        check(lib().flip2_sort_particles(self._h))

    def generateVoxels(self):
        check(lib().flip2_generate_voxels(self._h))

    def particleVelToVoxels(self):
        check(lib().flip2_particle_vel_to_voxels(self._h))

    def pressureSolve(self):
        check(lib().flip2_pressure_solve(self._h))

    def updateVoxelVelocities(self):
        check(lib().flip2_update_voxel_velocities(self._h))

    def advectParticles(self):
        check(lib().flip2_advect_particles(self._h))

    def solveFrame(self, fps):
        check(lib().flip2_solve_frame(self._h, float(fps)))

    def initialize(self):
        check(lib().flip2_initialize(self._h))

    def voxelVelsToParticles(self):
        check(lib().flip2_voxel_vels_to_particles(self._h))

    def moveSolidParticles(self):
        check(lib().flip2_move_solid_particles(self._h))

    def writePositionsToFile(self, fileName: str):
        check(lib().flip2_write_positions_to_file(self._h, fileName.encode()))

    def getCourantDt(self) -> float:
        d = _f64()
        check(lib().flip2_get_courant_dt(self._h, ctypes.byref(d)))
        return d.value

    # stage access --------------------------------------------------------
    def applyGravity(self, dt):
        check(lib().flip2_apply_gravity(self._h, dt))

    def removeGravity(self, dt):
        check(lib().flip2_remove_gravity(self._h, dt))

    def calcDivU(self):
        check(lib().flip2_calc_div_u(self._h))

    def gsSolve(self, dt, threshold=0.01, maxIterations=16384):
        r = _f64()
        it = _u32()
        check(lib().flip2_gs_solve(self._h, dt, threshold, maxIterations, ctypes.byref(r), ctypes.byref(it)))
        return r.value, it.value

    def gsIterations(self, dt, count):
        check(lib().flip2_gs_iterations(self._h, dt, count))

    def moveSolids(self):
        check(lib().flip2_move_solids(self._h))

    def setSolverLimits(self, threshold=0.01, maxIterations=16384):
        check(lib().flip2_set_solver_limits(self._h, threshold, maxIterations))

    def setTime(self, dt, frameDt, elapsedTime=0.0, elapsedTimeThisFrame=0.0):
        check(lib().flip2_set_time(self._h, dt, frameDt, elapsedTime, elapsedTimeThisFrame))

    def set_positions(self, px, py, pz):
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (px, py, pz)]
        assert all(v.size == self.size for v in a)
        check(lib().flip2_set_positions(self._h, *(_ptr(v) for v in a)))

    def set_velocities(self, vx, vy, vz):
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (vx, vy, vz)]
        assert all(v.size == self.size for v in a)
        check(lib().flip2_set_velocities(self._h, *(_ptr(v) for v in a)))

    def positions(self):
        out = [np.empty(self.size) for _ in range(3)]
        check(lib().flip2_get_positions(self._h, *(_ptr(v) for v in out)))
        return out

    def velocities(self):
        out = [np.empty(self.size) for _ in range(3)]
        check(lib().flip2_get_velocities(self._h, *(_ptr(v) for v in out)))
        return out

    def get(self, name: str) -> np.ndarray:
        cnt = _sz()
        check(lib().flip2_get_array(self._h, name.encode(), None, 0, ctypes.byref(cnt)))
        out = np.empty(cnt.value, dtype=_ARRAY_DTYPES[name])
        check(lib().flip2_get_array(self._h, name.encode(), _ptr(out) if out.size else None, out.size,
                                    ctypes.byref(cnt)))
        return out

    def set(self, name: str, values: np.ndarray) -> None:
        a = np.ascontiguousarray(values, dtype=_ARRAY_DTYPES[name])
        check(lib().flip2_set_array(self._h, name.encode(), _ptr(a) if a.size else None, a.size))

    def stats(self) -> dict:
        s = RcStatsC()
        check(lib().flip2_get_stats(self._h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in RcStatsC._fields_}
