"""The C-ABI library loads and exports every entry point include/flip2.h declares."""
import ctypes
import re

import pytest

import flip2_amd


def declared_functions():
    text = open(flip2_amd.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(flip2_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert len(names) > 40
    assert "flip2_sort_pairs_u32" in names and "flip2_solve_frame" in names


def test_library_exports_every_declared_symbol():
    lib = flip2_amd.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_grid_struct_matches_reference_layout():
    # engine/grid.hpp:35-37: three floats, three uints, one float
    assert ctypes.sizeof(flip2_amd.GridC) == 28
    assert flip2_amd.GridC.cellSize.offset == 24


def test_errors_are_reported_not_fatal():
    lib = flip2_amd.lib()
    rc = lib.flip2_particles_create(10, None)
    assert rc == -1
    assert b"NULL" in lib.flip2_last_error()


def _config(**kw):
    c = flip2_amd.FluidConfigC()
    flip2_amd.lib().flip2_fluid_default_config(ctypes.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


@pytest.mark.parametrize("field,value", [
    ("max_iterations", 0), ("max_iterations", 0xFFFFFFFF), ("max_iterations", (1 << 24) + 1),
    ("check_every", 1 << 20), ("dt", 0.0), ("dt", float("nan")), ("tolerance", -1.0), ("flip", 1.5),
    ("preconditioner", 7), ("margin", 0.75),
])
def test_fluid_config_refused_before_any_device_call(field, value):
    # the configuration is validated before the simulation touches the device,
    # so a bad value is refused the same way with or without a GPU
    lib = flip2_amd.lib()
    h = ctypes.c_void_p()
    rc = lib.flip2_fluid_create(16, ctypes.byref(_config(**{field: value})), ctypes.byref(h))
    assert rc == -1  # FLIP2_ERR_ARGUMENT
    assert b"config" in lib.flip2_last_error()
    assert not h.value


@pytest.mark.parametrize("n", [448, 512])
def test_unet_size_beyond_32bit_counts_refused(n):
    lib = flip2_amd.lib()
    h = ctypes.c_void_p()
    rc = lib.flip2_unet_create(n, ctypes.byref(h))
    assert rc == -5  # FLIP2_ERR_RANGE
    assert b"440" in lib.flip2_last_error()
    assert not h.value
