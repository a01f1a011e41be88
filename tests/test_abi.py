"""The C-ABI library loads and exports every entry point include/flip2.h declares."""
import ctypes
import re

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
