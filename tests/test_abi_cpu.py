"""CPU-side checks of the C ABI: the library builds for gfx950, loads, and
exports every entry point include/potg.h declares (no device calls)."""
import ctypes
import os
import re

from pot_amd import potg as G

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    hdr = open(os.path.join(REPO, "include", "potg.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(potg_\w+)\(", hdr, re.M)))


def test_header_and_binding_agree():
    assert declared_entry_points() == sorted(G.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = G.lib()
    for name in declared_entry_points():
        assert hasattr(L, name), name


def test_struct_layouts_match_header():
    assert ctypes.sizeof(G.Config) == 8 + 4 + 4 + 8 * 15 + 4 + 4 + 8 + 4 * 4 + 8
    assert G.PARTICLE_DT.itemsize == 56
    assert ctypes.sizeof(G.Grain) == 4 + 4 + 8 + 24 + 8 + 24
    assert ctypes.sizeof(G.StepOut) == 32
