"""ctypes binding of the C ABI in include/potg.h (libpotg.so).

This is the host-side mirror the tests and bench.py drive; the C host
program (pot_amd/host) links the same library directly.  It fails loudly
when the library is missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libpotg.so")

POTG_SPHERE, POTG_PROLATE, POTG_OBLATE = 0, 1, 2
POTG_FP64, POTG_MIXED, POTG_MIXED_GROUP = 0, 1, 2
POTG_BORIS, POTG_VERLET, POTG_EULER_RICHARDSON = 0, 1, 2
POTG_ERR_ARG = -1
POTG_NOALLOC_MAX_POINTS = 65536  # potg_phi_at_noalloc's limit (include/potg.h)
DUST, ION, ELECTRON = 0, 1, 2

PARTICLE_DT = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("vx", "<f8"),
                        ("vy", "<f8"), ("vz", "<f8"), ("species", "<i4"), ("pad", "<i4")])

ERRORS = {-1: "POTG_ERR_ARG", -2: "POTG_ERR_HIP", -3: "POTG_ERR_OUT_OF_BOX",
          -4: "POTG_ERR_COINCIDENT", -5: "POTG_ERR_DEPTH", -6: "POTG_ERR_NOMEM",
          -7: "POTG_ERR_REINJECT"}

# every entry point include/potg.h declares
EXPORTS = ["potg_create", "potg_destroy", "potg_last_error", "potg_upload", "potg_download",
           "potg_step", "potg_run", "potg_build_tree", "potg_tree_canon", "potg_field_at",
           "potg_phi_at", "potg_direct_sum", "potg_stats", "potg_stats_reset",
           "potg_stage_times", "potg_sync", "potg_tree_canon_dipoles", "potg_tune",
           "potg_set_tuning", "potg_set_counting", "potg_check_particles",
           "potg_phi_at_noalloc", "potg_set_retire", "potg_instance_retired"]


class Config(C.Structure):
    _fields_ = [("N", C.c_uint64), ("r", C.c_uint32), ("geometry", C.c_uint32),
                ("dt", C.c_double), ("mpp", C.c_double), ("soft", C.c_double),
                ("x_drift", C.c_double), ("a", C.c_double), ("settle", C.c_double),
                ("Te", C.c_double), ("Ti", C.c_double), ("A", C.c_double), ("theta", C.c_double),
                ("B", C.c_double * 3), ("max_lin_dev", C.c_double), ("mime", C.c_double),
                ("grain_enabled", C.c_int32), ("sceptic", C.c_int32), ("seed", C.c_uint64),
                ("dipoles", C.c_int32), ("integrator", C.c_int32), ("out_bounce", C.c_int32),
                ("pad", C.c_int32), ("bounce_coe", C.c_double)]


class Grain(C.Structure):
    _fields_ = [("q", C.c_int32), ("tot_absor_e", C.c_uint32), ("m", C.c_double),
                ("v", C.c_double * 3), ("chi_b", C.c_double), ("epe", C.c_double * 3)]


class StepOut(C.Structure):
    _fields_ = [("abs_q", C.c_int32), ("abs_e", C.c_uint32), ("abs_i", C.c_uint32),
                ("n_pushed", C.c_uint32), ("n_reinj_abs", C.c_uint32),
                ("n_reinj_out", C.c_uint32), ("n_wall_inward", C.c_uint32), ("pad", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("n_targets", C.c_uint64), ("n_cell", C.c_uint64),
                ("n_leaf", C.c_uint64), ("n_open", C.c_uint64), ("n_nodes", C.c_uint64),
                ("n_particles", C.c_uint64), ("max_depth", C.c_uint64),
                ("n_range_loads", C.c_uint64), ("n_child_visits", C.c_uint64)]


class Tuning(C.Structure):
    _fields_ = [("walk_blocks", C.c_uint32), ("group", C.c_uint32),
                ("variant", C.c_uint32), ("sort_tile", C.c_uint32), ("walk_ms", C.c_double)]


class PotgError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libpotg.so (built by pot_amd.build); raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PotgError("libpotg.so not built: run python -m pot_amd.build")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    up = C.POINTER(C.c_uint32)
    L.potg_last_error.restype = C.c_char_p
    L.potg_create.argtypes = [C.POINTER(Config), C.c_uint32, C.c_int32, C.POINTER(P)]
    L.potg_destroy.argtypes = [P]
    L.potg_upload.argtypes = [P, C.c_uint32, P, C.POINTER(Grain), C.c_uint64]
    L.potg_download.argtypes = [P, C.c_uint32, P, C.POINTER(Grain)]
    L.potg_step.argtypes = [P, C.c_int32, C.POINTER(StepOut)]
    L.potg_run.argtypes = [P, C.c_uint64]
    L.potg_build_tree.argtypes = [P]
    L.potg_tree_canon.argtypes = [P, C.c_uint32, P, P, P, P, P, P, P, P, C.c_int64,
                                  C.POINTER(C.c_int64)]
    L.potg_field_at.argtypes = [P, C.c_uint32, dp, up, C.c_size_t, dp, up]
    L.potg_phi_at.argtypes = [P, C.c_uint32, dp, C.c_size_t, dp]
    L.potg_direct_sum.argtypes = [P, C.c_uint32, up, C.c_size_t, dp]
    L.potg_stats.argtypes = [P, C.POINTER(Stats)]
    L.potg_stats_reset.argtypes = [P]
    L.potg_set_counting.argtypes = [P, C.c_int32]
    L.potg_stage_times.argtypes = [P, C.c_int32, dp]
    L.potg_sync.argtypes = [P]
    L.potg_tree_canon_dipoles.argtypes = [P, C.c_uint32, P, C.c_int64, C.POINTER(C.c_int64)]
    L.potg_tune.argtypes = [P, C.POINTER(Tuning)]
    L.potg_set_tuning.argtypes = [P, C.POINTER(Tuning), C.POINTER(Tuning)]
    L.potg_check_particles.argtypes = [C.POINTER(Config), P, C.POINTER(C.c_uint64)]
    L.potg_phi_at_noalloc.argtypes = [P, C.c_uint32, P, C.c_size_t, P, P]
    L.potg_set_retire.argtypes = [P, C.c_int32]
    L.potg_instance_retired.argtypes = [P, C.c_uint32, C.POINTER(C.c_int32),
                                        C.POINTER(C.c_uint64)]
    for n in EXPORTS:
        if n != "potg_last_error":
            getattr(L, n).restype = C.c_int
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise PotgError("%s: %s" % (ERRORS.get(rc, rc), lib().potg_last_error().decode()))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _up(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def config(N, r, mpp, dt, soft, a, Te, Ti, x_drift=0.0, settle=0.0, geometry=POTG_SPHERE,
           A=1.0, theta=0.5, B=(0.0, 0.0, 0.0), seed=1, grain_enabled=1, sceptic=1,
           mime=None, max_lin_dev=None, dipoles=0, integrator=0, bounce=None):
    c = Config()
    c.N, c.r, c.geometry = int(N), int(r), int(geometry)
    c.dt, c.mpp, c.soft, c.x_drift, c.a, c.settle = dt, mpp, soft, x_drift, a, settle
    c.Te, c.Ti, c.A, c.theta = Te, Ti, A, theta
    c.B[0], c.B[1], c.B[2] = B
    sph = geometry != POTG_SPHERE
    c.max_lin_dev = max_lin_dev if max_lin_dev is not None else (0.1 if sph else 0.01)
    c.mime = mime if mime is not None else (1836.153 if sph else 1836.15)
    c.grain_enabled, c.sceptic, c.seed = grain_enabled, sceptic, seed
    c.dipoles = dipoles
    c.integrator = integrator
    c.out_bounce = 0 if bounce is None else 1
    c.bounce_coe = 0.0 if bounce is None else bounce
    return c


def check_particles(cfg, parts):
    """potg_check_particles: (code, first bad index, message); host only."""
    parts = np.ascontiguousarray(parts, PARTICLE_DT)
    bad = C.c_uint64()
    rc = lib().potg_check_particles(C.byref(cfg), parts.ctypes.data, C.byref(bad))
    return rc, bad.value, lib().potg_last_error().decode() if rc else ""


class Context:
    """A device context holding one or more instances."""

    def __init__(self, configs, precision=POTG_FP64):
        if isinstance(configs, Config):
            configs = [configs]
        self.L = lib()
        self.configs = list(configs)
        arr = (Config * len(configs))(*configs)
        h = C.c_void_p()
        _check(self.L.potg_create(arr, len(configs), precision, C.byref(h)))
        self.h = h
        self.n_inst = len(configs)
        self.precision = precision

    def close(self):
        if getattr(self, "h", None):
            self.L.potg_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def upload(self, inst, parts, q=0, m=1.67e-12, t=0, chi_b=0.0, tot_e=0):
        parts = np.ascontiguousarray(parts, PARTICLE_DT)
        g = Grain()
        g.q, g.m, g.chi_b, g.tot_absor_e = q, m, chi_b, tot_e
        _check(self.L.potg_upload(self.h, inst, parts.ctypes.data, C.byref(g), t))

    def download(self, inst):
        N = self.configs[inst].N
        out = np.zeros(N, PARTICLE_DT)
        g = Grain()
        _check(self.L.potg_download(self.h, inst, out.ctypes.data, C.byref(g)))
        return out, {"q": g.q, "m": g.m, "v": tuple(g.v), "chi_b": g.chi_b,
                     "tot_e": g.tot_absor_e, "epe": tuple(g.epe)}

    def step_async(self, compute_epe=0):
        _check(self.L.potg_step(self.h, compute_epe, None))

    def sync(self):
        _check(self.L.potg_sync(self.h))

    def step(self, compute_epe=0):
        out = (StepOut * self.n_inst)()
        _check(self.L.potg_step(self.h, compute_epe, out))
        return [{"abs_q": o.abs_q, "abs_e": o.abs_e, "abs_i": o.abs_i, "n_pushed": o.n_pushed,
                 "n_reinj_abs": o.n_reinj_abs, "n_reinj_out": o.n_reinj_out,
                 "n_wall_inward": o.n_wall_inward} for o in out]

    def run(self, steps):
        _check(self.L.potg_run(self.h, steps))

    def build_tree(self):
        _check(self.L.potg_build_tree(self.h))

    def tree_canon(self, inst=0):
        cap = 4 * int(self.configs[inst].N) + 16
        depth = np.zeros(cap, np.int32)
        leaf = np.zeros(cap, np.int32)
        part = np.zeros(cap, np.uint32)
        q = np.zeros(cap, np.int32)
        aq = np.zeros(cap, np.uint32)
        c = np.zeros((cap, 3))
        s2 = np.zeros(cap)
        nch = np.zeros(cap, np.int32)
        n = C.c_int64()
        _check(self.L.potg_tree_canon(self.h, inst, depth.ctypes.data, leaf.ctypes.data,
                                      part.ctypes.data, q.ctypes.data, aq.ctypes.data,
                                      c.ctypes.data, s2.ctypes.data, nch.ctypes.data, cap,
                                      C.byref(n)))
        n = n.value
        return {"depth": depth[:n], "is_leaf": leaf[:n], "part": part[:n], "q": q[:n],
                "aq": aq[:n], "c": c[:n], "size2": s2[:n], "nchild": nch[:n]}

    def tree_dipoles(self, inst=0):
        cap = 4 * int(self.configs[inst].N) + 16
        p = np.zeros((cap, 3))
        n = C.c_int64()
        _check(self.L.potg_tree_canon_dipoles(self.h, inst, p.ctypes.data, cap, C.byref(n)))
        return p[:n.value]

    def field_at(self, inst, xyz, exclude=None, counts=False):
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        n = len(xyz)
        E = np.zeros((n, 3))
        ex = None if exclude is None else np.ascontiguousarray(exclude, np.uint32)
        cnt = np.zeros((n, 3), np.uint32) if counts else None
        _check(self.L.potg_field_at(self.h, inst, _dp(xyz), None if ex is None else _up(ex), n,
                                    _dp(E), None if cnt is None else _up(cnt)))
        return (E, cnt) if counts else E

    def phi_at(self, inst, xyz):
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        out = np.zeros(len(xyz))
        _check(self.L.potg_phi_at(self.h, inst, _dp(xyz), len(xyz), _dp(out)))
        return out

    def direct_sum(self, inst, idx):
        idx = np.ascontiguousarray(idx, np.uint32)
        E = np.zeros((len(idx), 3))
        _check(self.L.potg_direct_sum(self.h, inst, _up(idx), len(idx), _dp(E)))
        return E

    def tune(self):
        t = Tuning()
        _check(self.L.potg_tune(self.h, C.byref(t)))
        return {"walk_blocks": t.walk_blocks, "group": t.group, "variant": t.variant,
                "sort_tile": t.sort_tile, "walk_ms": t.walk_ms}

    def tuning(self):
        t = Tuning()
        _check(self.L.potg_set_tuning(self.h, None, C.byref(t)))
        return {"walk_blocks": t.walk_blocks, "group": t.group, "variant": t.variant,
                "sort_tile": t.sort_tile, "walk_ms": t.walk_ms}

    def set_tuning(self, walk_blocks, group=64, variant=0, sort_tile=0):
        t = Tuning()
        t.walk_blocks, t.group, t.variant, t.sort_tile = walk_blocks, group, variant, sort_tile
        _check(self.L.potg_set_tuning(self.h, C.byref(t), None))

    def stats(self):
        s = Stats()
        _check(self.L.potg_stats(self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in Stats._fields_}

    def stats_reset(self):
        _check(self.L.potg_stats_reset(self.h))

    def set_counting(self, on):
        _check(self.L.potg_set_counting(self.h, 1 if on else 0))

    def set_retire(self, on):
        """A failed reinjection retires its instance instead of failing the step."""
        _check(self.L.potg_set_retire(self.h, 1 if on else 0))

    def retired(self, inst):
        """(retired, step of the failure) of instance inst"""
        r, t = C.c_int32(), C.c_uint64()
        _check(self.L.potg_instance_retired(self.h, inst, C.byref(r), C.byref(t)))
        return bool(r.value), int(t.value)

    def stage_times(self, enable=1):
        ms = np.zeros(5)
        _check(self.L.potg_stage_times(self.h, enable, _dp(ms)))
        return ms
