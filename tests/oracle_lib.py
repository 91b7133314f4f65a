"""ctypes binding of the CPU oracle (oracle/pot_oracle.c) for the tests.

The oracle is test infrastructure: only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg load it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "_build", "liboracle.so")

PARTICLE_DT = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("vx", "<f8"),
                        ("vy", "<f8"), ("vz", "<f8"), ("species", "<i4"), ("pad", "<i4")])
assert PARTICLE_DT.itemsize == 56

DUST, ION, ELECTRON = 0, 1, 2
SPHERE, PROLATE, OBLATE = 0, 1, 2
RNG_WELL512, RNG_PHILOX = 0, 1
BORIS, VERLET, EULER_RICHARDSON = 0, 1, 2


class Config(C.Structure):
    _fields_ = [("N", C.c_uint64), ("r", C.c_uint32), ("geometry", C.c_uint32),
                ("dt", C.c_longdouble), ("mpp", C.c_double), ("soft", C.c_double),
                ("x_drift", C.c_double), ("a", C.c_double), ("settle", C.c_double),
                ("Te", C.c_double), ("Ti", C.c_double), ("A", C.c_double),
                ("theta", C.c_double), ("B", C.c_double * 3), ("max_lin_dev", C.c_double),
                ("mime", C.c_double), ("grain_enabled", C.c_int32), ("sceptic", C.c_int32),
                ("inplace", C.c_int32), ("rng_kind", C.c_int32), ("philox_seed", C.c_uint64),
                ("record_macro", C.c_int32), ("dipoles", C.c_int32),
                ("integrator", C.c_int32), ("out_bounce", C.c_int32), ("bounce_coe", C.c_double)]


class WalkStats(C.Structure):
    _fields_ = [("n_cell", C.c_uint64), ("n_leaf", C.c_uint64), ("n_open", C.c_uint64),
                ("n_mac", C.c_uint64)]


class StepOut(C.Structure):
    _fields_ = [("abs_q", C.c_int32), ("abs_e", C.c_uint32), ("abs_i", C.c_uint32),
                ("n_pushed", C.c_uint32), ("n_reinj_abs", C.c_uint32),
                ("n_reinj_out", C.c_uint32), ("start", C.c_uint32), ("n_wall_inward", C.c_uint32)]


class Prng(C.Structure):
    _fields_ = [("a", C.c_uint32), ("b", C.c_uint32), ("c", C.c_uint32), ("d", C.c_uint32),
                ("index", C.c_uint32), ("state", C.c_uint32 * 16)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    sig = {
        "or_config_set_dt": (None, [C.POINTER(Config), C.c_char_p]),
        "or_strtold_to_double": (C.c_double, [C.c_char_p]),
        "or_create": (P, [C.POINTER(Config)]),
        "or_destroy": (None, [P]),
        "or_last_error": (C.c_char_p, []),
        "or_get_particles": (None, [P, P]),
        "or_set_particles": (None, [P, P]),
        "or_set_state": (None, [P, C.c_int32, C.c_longdouble, C.c_uint64, C.c_double, C.c_uint32]),
        "or_get_state": (None, [P, C.POINTER(C.c_int32), P, C.POINTER(C.c_uint64), dp,
                                C.POINTER(C.c_uint32), P]),
        "or_get_prng": (None, [P, C.POINTER(Prng)]),
        "or_set_prng": (None, [P, C.POINTER(Prng)]),
        "or_get_tss": (None, [P, C.POINTER(C.c_uint32)]),
        "or_init_particles": (None, [P, C.POINTER(C.c_uint32)]),
        "or_build_tree": (C.c_int, [P]),
        "or_tree_canon": (C.c_int64, [P, P, P, P, P, P, P, P, P, C.c_int64]),
        "or_tree_used": (C.c_uint32, [P]),
        "or_tree_canon_dipoles": (C.c_int64, [P, dp, C.c_int64]),
        "or_field_at": (None, [P, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_int32,
                               dp, C.POINTER(WalkStats)]),
        "or_phi_at": (C.c_double, [P, C.c_double, C.c_double, C.c_double]),
        "or_phi_particle_feels": (C.c_double, [P, C.c_uint32]),
        "or_field_from_grain": (None, [P, C.c_uint32, C.c_double, C.c_double, C.c_double, dp]),
        "or_phi_from_grain": (C.c_double, [P, C.c_uint32]),
        "or_direct_sum": (None, [P, C.c_uint32, C.c_double, C.c_double, C.c_double, dp]),
        "or_measure_chi_b": (C.c_double, [P]),
        "or_run_step": (C.c_int, [P, C.POINTER(StepOut)]),
        "or_move_particles": (C.c_int, [P, C.c_int, C.POINTER(StepOut)]),
        "or_write_state": (C.c_int, [P, C.c_char_p]),
        "or_read_state": (C.c_int, [P, C.c_char_p]),
        "or_u_no_flow_icdf": (C.c_double, [C.c_double, C.c_double]),
        "or_u_icdf": (C.c_double, [C.c_double, C.c_double, C.c_double]),
        "or_integrate_for_alpha": (C.c_double, [C.c_double] * 5),
        "or_philox4x32_10": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_uint32)]),
        "or_sceptic_from_uniforms": (C.c_int, [P, P, C.c_double, dp, C.c_int]),
        "or_reinject": (None, [P, P, C.c_uint32]),
        "or_crosses_grain": (C.c_int, [P, dp, dp, dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_config(args, geometry=SPHERE, theta=0.5, B=(0.0, 0.0, 0.0), inplace=1,
                rng_kind=0, philox_seed=0, record_macro=1, grain_enabled=1, sceptic=1,
                dipoles=0, integrator=0, bounce=None):
    """Config from reference-style CLI values (strings), parsed as pot.c does."""
    L = lib()
    cfg = Config()
    cfg.N = int(args["-N"])
    cfg.r = int(args["-r"])
    cfg.geometry = geometry
    L.or_config_set_dt(C.byref(cfg), args["-d"].encode())
    cfg.mpp = float(args["-m"])
    cfg.soft = float(args["-s"])
    cfg.x_drift = float(args.get("-x", "0"))
    cfg.a = float(args["-a"])
    cfg.settle = L.or_strtold_to_double(args.get("-e", "0").encode())
    cfg.Te = L.or_strtold_to_double(args["-E"].encode())
    cfg.Ti = L.or_strtold_to_double(args["-I"].encode())
    cfg.A = float(args.get("-A", "1"))
    cfg.theta = theta
    cfg.B[0], cfg.B[1], cfg.B[2] = B
    sph = geometry != SPHERE
    cfg.max_lin_dev = 0.1 if sph else 0.01
    cfg.mime = 1836.153 if sph else 1836.15
    cfg.grain_enabled = grain_enabled
    cfg.sceptic = sceptic
    cfg.inplace = inplace
    cfg.rng_kind = rng_kind
    cfg.philox_seed = philox_seed
    cfg.record_macro = record_macro
    cfg.dipoles = dipoles
    cfg.integrator = integrator
    cfg.out_bounce = 0 if bounce is None else 1
    cfg.bounce_coe = 0.0 if bounce is None else bounce
    return cfg


class Sim:
    """Owning handle on an oracle simulation."""

    def __init__(self, cfg):
        self.L = lib()
        self.cfg = cfg
        self.N = cfg.N
        self.h = self.L.or_create(C.byref(cfg))
        if not self.h:
            raise RuntimeError(self.L.or_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            self.L.or_destroy(self.h)
            self.h = None

    def init_particles(self, seed_words):
        w = (C.c_uint32 * 16)(*seed_words)
        self.L.or_init_particles(self.h, w)

    def particles(self):
        out = np.zeros(self.N, PARTICLE_DT)
        self.L.or_get_particles(self.h, out.ctypes.data)
        return out

    def set_particles(self, arr):
        arr = np.ascontiguousarray(arr, PARTICLE_DT)
        self.L.or_set_particles(self.h, arr.ctypes.data)

    def state(self):
        q = C.c_int32()
        t = C.c_uint64()
        chi = C.c_double()
        tot = C.c_uint32()
        m = C.c_longdouble()
        self.L.or_get_state(self.h, C.byref(q), C.byref(m), C.byref(t), C.byref(chi), C.byref(tot),
                            None)
        return {"q": q.value, "t": t.value, "chi_b": chi.value, "tot_e": tot.value, "m": m.value}

    def set_state(self, q, m, t, chi_b, tot_e):
        self.L.or_set_state(self.h, q, m, t, chi_b, tot_e)

    def tss(self):
        a = (C.c_uint32 * 3)()
        self.L.or_get_tss(self.h, a)
        return list(a)

    def build_tree(self):
        rc = self.L.or_build_tree(self.h)
        if rc:
            raise RuntimeError("build_tree %d: %s" % (rc, self.L.or_last_error().decode()))

    def tree_canon(self):
        cap = 4 * self.N + 16
        depth = np.zeros(cap, np.int32)
        leaf = np.zeros(cap, np.int32)
        part = np.zeros(cap, np.uint32)
        q = np.zeros(cap, np.int32)
        aq = np.zeros(cap, np.uint32)
        c = np.zeros((cap, 3), np.float64)
        x12 = np.zeros((cap, 2), np.float64)
        nch = np.zeros(cap, np.int32)
        n = self.L.or_tree_canon(self.h, depth.ctypes.data, leaf.ctypes.data, part.ctypes.data,
                                 q.ctypes.data, aq.ctypes.data, c.ctypes.data, x12.ctypes.data,
                                 nch.ctypes.data, cap)
        if n < 0:
            raise RuntimeError("tree_canon failed")
        return {"depth": depth[:n], "is_leaf": leaf[:n], "part": part[:n], "q": q[:n],
                "aq": aq[:n], "c": c[:n], "x1x2": x12[:n], "nchild": nch[:n]}

    def tree_dipoles(self):
        cap = 4 * self.N + 16
        p = np.zeros((cap, 3))
        n = self.L.or_tree_canon_dipoles(self.h, dptr(p), cap)
        if n < 0:
            raise RuntimeError("tree_canon_dipoles failed")
        return p[:n]

    def field_at(self, part, x, y, z, accum_double=0, stats=None):
        E = np.zeros(3)
        self.L.or_field_at(self.h, part, x, y, z, accum_double, dptr(E),
                           C.byref(stats) if stats is not None else None)
        return E

    def direct_sum(self, part, x, y, z):
        E = np.zeros(3)
        self.L.or_direct_sum(self.h, part, x, y, z, dptr(E))
        return E

    def phi_at(self, x, y, z):
        return self.L.or_phi_at(self.h, x, y, z)

    def run_step(self):
        out = StepOut()
        rc = self.L.or_run_step(self.h, C.byref(out))
        if rc:
            raise RuntimeError("run_step %d: %s" % (rc, self.L.or_last_error().decode()))
        return out

    def move(self, compute_epe=0):
        out = StepOut()
        self.L.or_move_particles(self.h, compute_epe, C.byref(out))
        return out

    def write_state(self, path):
        if self.L.or_write_state(self.h, path.encode()):
            raise RuntimeError("write_state")

    def read_state(self, path):
        rc = self.L.or_read_state(self.h, path.encode())
        if rc:
            raise RuntimeError("read_state %d: %s" % (rc, self.L.or_last_error().decode()))


def read_state_file(path, N):
    """Parse a format-C state file (pot.c:3870-3890): header, PRNG, particles."""
    raw = open(path, "rb").read()
    nl = raw.index(b"\n")
    header = raw[:nl].decode()
    prng = np.frombuffer(raw[nl + 1:nl + 1 + 84], np.uint32)
    parts = np.frombuffer(raw[nl + 1 + 84:nl + 1 + 84 + 56 * N], PARTICLE_DT)
    assert len(parts) == N
    return header, prng, parts
