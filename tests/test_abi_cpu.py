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


def _sphere_config(N=9):
    return G.config(N=N, r=1000, mpp=1e-9, dt=1e-16, soft=1e-11, a=1e-8, Te=11600.0,
                    Ti=11600.0)


def _particles(N=9):
    import numpy as np
    p = np.zeros(N, G.PARTICLE_DT)
    R = 1000 * 1e-9
    p["x"] = p["y"] = p["z"] = R
    p["species"][0] = G.DUST
    p["species"][1:] = G.ION
    p["species"][N // 2:] = G.ELECTRON
    p["x"][1:] = R * (0.5 + 0.1 * np.arange(1, N))
    return p


def test_create_rejects_non_positive_temperatures():
    """Te and Ti divide: the ion sub-cycling factor (pot.c:4157-4159), the
    reinjection's shielding length (pot.c:578-580) and chi_b's kT_i/e
    (pot.c:2278).  potg_create validates before any device call."""
    L = G.lib()
    for te, ti in ((11600.0, 0.0), (0.0, 11600.0), (11600.0, -1.0), (float("nan"), 1.0),
                   (11600.0, float("inf"))):
        cfg = _sphere_config()
        cfg.Te, cfg.Ti = te, ti
        h = ctypes.c_void_p()
        rc = L.potg_create(ctypes.byref(cfg), 1, G.POTG_FP64, ctypes.byref(h))
        assert rc == -1 and not h.value, (te, ti, rc)
        assert b"temperatures" in L.potg_last_error()


def test_check_particles_accepts_a_valid_load():
    assert G.check_particles(_sphere_config(), _particles())[0] == 0


def test_check_particles_rejects_out_of_box_and_nan():
    """build_tree refuses a particle outside the root cell [0, 2 r mpp]^3
    (pot.c:1503-1514) and a NaN position (pot.c:1434-1446), with the
    reference's message; the upper face itself is inside."""
    import numpy as np
    cfg = _sphere_config()
    x2 = 2 * 1000 * 1e-9
    for field, value, ok in (("x", -1e-30, False), ("y", x2 * (1 + 1e-15), False),
                             ("z", np.nan, False), ("x", x2, True), ("z", 0.0, True)):
        p = _particles()
        p[field][5] = value
        rc, bad, msg = G.check_particles(cfg, p)
        if ok:
            assert rc == 0, (field, value)
        else:
            assert rc == -3 and bad == 5, (field, value, rc, bad)
            assert msg.startswith("particle 5 at (") and msg.endswith("not in simulation area.")


def test_check_particles_rejects_bad_species():
    p = _particles()
    p["species"][3] = G.DUST
    rc, bad, _ = G.check_particles(_sphere_config(), p)
    assert rc == -1 and bad == 3
    p = _particles()
    p["species"][0] = G.ION
    assert G.check_particles(_sphere_config(), p)[0] == -1
