"""Build recipe of the gfx950 library (libpotg.so) and the host driver.

    python -m pot_amd.build          # or pot_amd.build.build()

Everything is compiled in-tree into pot_amd/_build/ with hipcc for gfx950
only.  Flags:
  -ffp-contract=off   no fused multiply-add unless written: the float64
                      tree build and walk repeat the reference's roundings
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed float32 instructions
  -munsafe-fp-atomics hardware float64 atomic add for the momentum tally
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
LIB = os.path.join(BUILD, "libpotg.so")

HIP_SOURCES = ["potg_prim.hip", "potg_tree.hip", "potg_walk.hip", "potg_push.hip", "potg_api.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics", "-Wall", "-Wno-unused-result"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r.stdout


def build_lib(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    procs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                            text=True)))
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed = True
    if failed:
        raise RuntimeError("hipcc failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


HOST_LIB_SOURCES = ["pot_init.c", "pot_macro.c", "pot_state.c", "pot_info.c"]
HOST_MAIN = "pot_main.c"
HOST_LIB = os.path.join(BUILD, "libpothost.so")


def build_host():
    """The host side in C: libpothost.so (initial load, state files, outputs)
    and, when its main exists, the drop-in programs pot and pot-spheroid."""
    os.makedirs(BUILD, exist_ok=True)
    inc = ["-I" + os.path.join(REPO, "include"), "-I" + HOST]
    cflags = ["gcc", "-std=c99", "-O2", "-Wall", "-Wextra", "-D_GNU_SOURCE", "-fPIC"]
    srcs = [os.path.join(HOST, f) for f in HOST_LIB_SOURCES]
    _run(cflags + ["-shared"] + inc + srcs + ["-L" + BUILD, "-lpotg", "-Wl,-rpath,$ORIGIN",
                                             "-lm", "-o", HOST_LIB])
    outs = [HOST_LIB]
    main = os.path.join(HOST, HOST_MAIN)
    if os.path.exists(main):
        for name, geom in (("pot", "SPHERE"), ("pot-spheroid", "SPHEROID")):
            out = os.path.join(BUILD, name)
            _run(cflags + ["-DPOT_GEOMETRY_" + geom] + inc + [main] + srcs +
                 ["-L" + BUILD, "-lpotg", "-Wl,-rpath,$ORIGIN", "-lm", "-o", out])
            outs.append(out)
    return outs


def build():
    lib = build_lib()
    host = build_host()
    return [lib] + host


if __name__ == "__main__":
    for p in build():
        print(p)
