"""The Blender operator module (flip2_amd/plugin.py): importable without
Blender, the operator wired as the reference's (plugin/plugin.py: same
bl_idname, label, options, menu entry), execute() driving the simulation
loop; on the GPU the loop over the C ABI gives the same positions, bit for
bit, as the same frames through flip2_amd.Particles."""
import importlib
import sys
import types

import numpy as np
import pytest


def test_imports_without_blender():
    from flip2_amd import plugin
    assert plugin.bl_info["name"] == "FLUID SOLVER"
    assert plugin.FluidSolver is None  # no bpy here
    d = plugin.domain_around(np.array([[0.0, 0.0, 0.0], [2.0, 1.0, 1.0]]), cells=32, margin=0.25)
    assert d[:3] == (-0.5, -0.5, -0.5) and d[3:6] == (32, 32, 32) and d[6] == pytest.approx(3.0 / 32)
    lib = plugin.load_library()
    for name in ("flip2_particles_create", "flip2_set_domain", "flip2_solve_frame", "flip2_get_positions"):
        assert hasattr(lib, name)
    with pytest.raises(ValueError):
        next(plugin.simulate_frames(np.zeros(0), np.zeros(0), np.zeros(0), 1))


def test_operator_with_a_stand_in_bpy(monkeypatch):
    registered, menu = [], []

    class Operator:
        def report(self, kind, msg):
            self.reported = (kind, msg)

    bpy = types.ModuleType("bpy")
    bpy.types = types.SimpleNamespace(Operator=Operator,
                                      VIEW3D_MT_object=types.SimpleNamespace(append=menu.append))
    bpy.utils = types.SimpleNamespace(register_class=registered.append, unregister_class=registered.remove)
    monkeypatch.setitem(sys.modules, "bpy", bpy)
    import flip2_amd.plugin as plugin
    plugin = importlib.reload(plugin)
    try:
        op_cls = plugin.FluidSolver
        assert op_cls.bl_idname == "fluid.solver"
        assert op_cls.bl_label == "Simulate fluid motion of particles"
        assert op_cls.bl_options == {'REGISTER', 'UNDO'}
        plugin.register()
        assert registered == [op_cls] and menu == [plugin.menu_func]
        calls = []
        monkeypatch.setattr(plugin, "simulate_objects", lambda objs, frames, fps, start: calls.append(
            (objs, frames, fps, start)))
        scene = types.SimpleNamespace(frame_start=1, frame_end=25,
                                      render=types.SimpleNamespace(fps=48, fps_base=2.0))
        ctx = types.SimpleNamespace(scene=scene, selected_objects=["a"])
        assert op_cls().execute(ctx) == {'FINISHED'}
        assert calls == [(["a"], 24, 24.0, 1)]

        def fail(*a):
            raise plugin.SimulationError("boom")
        monkeypatch.setattr(plugin, "simulate_objects", fail)
        op = op_cls()
        assert op.execute(ctx) == {'CANCELLED'} and op.reported == ({'ERROR'}, "boom")
        plugin.unregister()
        assert registered == []
    finally:
        monkeypatch.delitem(sys.modules, "bpy")
        importlib.reload(plugin)


@pytest.mark.gpu
def test_simulation_loop_matches_particles(flip2):
    from flip2_amd import plugin, scenes
    n = 16
    mask = np.zeros((n, n, n), dtype=bool)
    mask[2:8, 0:8, 2:14] = True
    x, y, z = scenes.jittered_cells((n, n, n), 1.0 / n, seed=9, cell_mask=mask)
    pos = [a.astype(np.float64) for a in (x, y, z)]
    domain = (0.0, 0.0, 0.0, n, n, n, 1.0 / n)
    frames = list(plugin.simulate_frames(*pos, 2, fps=48.0, domain=domain))
    assert len(frames) == 2
    p = flip2.Particles(x.size)
    p.setDomain(*domain)
    p.set_positions(*pos)
    zero = np.zeros(x.size)
    p.set_velocities(zero, zero, zero)
    p.initialize()
    for k in range(2):
        p.solveFrame(48.0)
        for a, b in zip(frames[k], (p.get("px"), p.get("py"), p.get("pz"))):
            assert np.array_equal(a, b)
    assert not np.array_equal(frames[1][1], pos[1])  # the fluid moved
