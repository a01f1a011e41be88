"""Blender operator on libparticles.so: the reference's plugin/plugin.py with
an execute() that simulates.

The reference's add-on opens ../build/engine/libparticles.so relative to the
working directory (plugin/plugin.py:4-5) and registers an operator
"fluid.solver" whose execute() does nothing (plugin.py:22-27).  Here the
same operator (same bl_idname, label, options and menu entry) reads the
vertices of the selected mesh objects as particles, runs the library's
reference-compatible step (Particles::initialize, then solveFrame(fps) per
frame, particles.cu:784-812) through the C ABI of include/flip2.h, and
writes each frame's positions back as one shape key per frame, keyed on
the timeline.

simulate_frames() is the simulation loop itself: plain ctypes on the C ABI,
importable and testable without Blender (bpy is imported only when present).
"""
import ctypes
import os

import numpy as np

bl_info = {
    "name": "FLUID SOLVER",
    "blender": (2, 80, 0),
    "category": "Object",
}

_HERE = os.path.dirname(os.path.abspath(__file__))


def library_path():
    """FLIP2_LIBRARY if set, else the library built next to this package
    (build/engine/libparticles.so at the repository root)."""
    env = os.environ.get("FLIP2_LIBRARY")
    if env:
        return env
    return os.path.join(os.path.dirname(_HERE), "build", "engine", "libparticles.so")


_lib = None


def load_library(path=None):
    """Opens libparticles.so and declares the entry points the loop uses."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or library_path()
    if not os.path.exists(p):
        raise FileNotFoundError(f"{p} is not built (make -C flip2_amd/csrc)")
    lib = ctypes.CDLL(p)
    vp, u32, f64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_double
    lib.flip2_last_error.restype = ctypes.c_char_p
    lib.flip2_last_error.argtypes = []
    sigs = {
        "flip2_particles_create": [u32, ctypes.POINTER(vp)],
        "flip2_particles_destroy": [vp],
        "flip2_set_domain": [vp, f64, f64, f64, u32, u32, u32, f64],
        "flip2_set_positions": [vp, vp, vp, vp],
        "flip2_set_velocities": [vp, vp, vp, vp],
        "flip2_get_positions": [vp, vp, vp, vp],
        "flip2_initialize": [vp],
        "flip2_solve_frame": [vp, f64],
    }
    for name, args in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    if path is None:
        _lib = lib
    return lib


class SimulationError(RuntimeError):
    pass


def _check(lib, rc):
    if rc != 0:
        raise SimulationError(lib.flip2_last_error().decode(errors="replace"))


def domain_around(points, cells=64, margin=0.25):
    """A cubic domain (Particles::setDomain arguments: corner, cells per
    axis, cell size) holding the points with `margin` of their extent
    around them."""
    pts = np.asarray(points, np.float64).reshape(-1, 3)
    lo, hi = pts.min(axis=0), pts.max(axis=0)
    ext = float(max(np.max(hi - lo), 1e-6))
    size = ext * (1.0 + 2.0 * margin)
    corner = lo - margin * ext
    return (float(corner[0]), float(corner[1]), float(corner[2]), cells, cells, cells, size / cells)


def simulate_frames(px, py, pz, frames, fps=24.0, domain=None, vx=None, vy=None, vz=None, lib=None):
    """Yields the particle positions (three float64 arrays) after each of
    `frames` frames of the reference-compatible step.  domain: the
    Particles::setDomain arguments (negX, negY, negZ, sizeX, sizeY, sizeZ,
    cellSize); by default a 64^3 cube around the particles.  Velocities start
    at zero unless given."""
    lib = lib or load_library()
    pos = [np.ascontiguousarray(a, np.float64) for a in (px, py, pz)]
    n = pos[0].size
    if n == 0 or any(a.size != n for a in pos):
        raise ValueError("positions must be three arrays of the same non-zero length")
    if domain is None:
        domain = domain_around(np.stack(pos, axis=1))
    h = ctypes.c_void_p()
    _check(lib, lib.flip2_particles_create(n, ctypes.byref(h)))
    try:
        _check(lib, lib.flip2_set_domain(h, *domain))
        _check(lib, lib.flip2_set_positions(h, *(a.ctypes.data for a in pos)))
        if vx is not None:
            vel = [np.ascontiguousarray(a, np.float64) for a in (vx, vy, vz)]
            _check(lib, lib.flip2_set_velocities(h, *(a.ctypes.data for a in vel)))
        _check(lib, lib.flip2_initialize(h))
        for _ in range(frames):
            _check(lib, lib.flip2_solve_frame(h, float(fps)))
            out = [np.empty(n) for _ in range(3)]
            _check(lib, lib.flip2_get_positions(h, *(a.ctypes.data for a in out)))
            yield out
    finally:
        lib.flip2_particles_destroy(h)


try:  # Blender's Python only
    import bpy
except ImportError:  # pragma: no cover - outside Blender
    bpy = None


def _world_vertices(obj):
    m = obj.matrix_world
    return np.array([tuple(m @ v.co) for v in obj.data.vertices], np.float64)


def simulate_objects(objects, frames, fps, start_frame=1):
    """Runs the simulation on the vertices of `objects` (all of them as one
    particle set) and stores frame k's positions as shape key "flip2_k" of
    each object, keyed to 1 on its frame and 0 on the neighbouring ones."""
    meshes = [o for o in objects if getattr(o, "type", None) == "MESH" and len(o.data.vertices)]
    if not meshes:
        return 0
    verts = [_world_vertices(o) for o in meshes]
    counts = [len(v) for v in verts]
    allv = np.concatenate(verts, axis=0)
    for o in meshes:
        if o.data.shape_keys is None:
            o.shape_key_add(name="Basis")
    for k, (x, y, z) in enumerate(simulate_frames(allv[:, 0], allv[:, 1], allv[:, 2], frames, fps), start=1):
        off = 0
        for o, c in zip(meshes, counts):
            inv = o.matrix_world.inverted()
            key = o.shape_key_add(name=f"flip2_{k}")
            for i in range(c):
                key.data[i].co = inv @ type(o.matrix_world.translation)((x[off + i], y[off + i], z[off + i]))
            f = start_frame + k
            for frame, val in ((f - 1, 0.0), (f, 1.0), (f + 1, 0.0)):
                key.value = val
                key.keyframe_insert("value", frame=frame)
            off += c
    return frames


if bpy is not None:  # pragma: no cover - needs Blender

    class FluidSolver(bpy.types.Operator):
        """Fluid Solver"""
        bl_idname = "fluid.solver"
        bl_label = "Simulate fluid motion of particles"
        bl_options = {'REGISTER', 'UNDO'}

        def execute(self, context):
            scene = context.scene
            frames = max(0, scene.frame_end - scene.frame_start)
            fps = scene.render.fps / scene.render.fps_base
            try:
                simulate_objects(context.selected_objects, frames, fps, scene.frame_start)
            except (SimulationError, FileNotFoundError, ValueError) as e:
                self.report({'ERROR'}, str(e))
                return {'CANCELLED'}
            return {'FINISHED'}

    def menu_func(self, context):
        self.layout.operator(FluidSolver.bl_idname)

    def register():
        bpy.utils.register_class(FluidSolver)
        bpy.types.VIEW3D_MT_object.append(menu_func)

    def unregister():
        bpy.utils.unregister_class(FluidSolver)
else:
    FluidSolver = None


if __name__ == "__main__" and bpy is not None:  # pragma: no cover
    register()
