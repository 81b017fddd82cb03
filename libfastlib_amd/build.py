"""Build libfastbit_hip.so (gfx950) and the host-only self-test library.

Invoked by ``__graft_entry__.build()``; run as ``python -m libfastlib_amd.build``.
Everything is compiled in-tree so that the built libraries travel with the
repository snapshot.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
ROOT = os.path.dirname(HERE)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument", "-I" + os.path.join(ROOT, "include")]

SOURCES = ["wah_kernels.hip", "canon_kernels.hip", "index_kernels.hip", "aux_kernels.hip", "runtime.cpp",
           "bitvector_api.cpp",
           "index_api.cpp", "bin_bounds.cpp", "mlp_kernels.hip", "planner_api.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd[:3]) + " ...")
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(LIBDIR, src + ".o")
        lang = ["-x", "hip"]
        cmd = [HIPCC] + HIP_FLAGS + lang + ["-c", path, "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                            text=True)))
        objs.append(obj)
    for cmd, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out + err)
            raise RuntimeError("compile failed: " + cmd[-3])
    so = os.path.join(LIBDIR, "libfastbit_hip.so")
    tmp = so + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, so)
    for o in objs:
        os.remove(o)
    # host-only self-test of the shared host/device step core
    sim = os.path.join(LIBDIR, "libfbsim.so")
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-o", sim,
          os.path.join(CSRC, "selftest_sim.cpp")])
    gen = os.path.join(LIBDIR, "libfbgen.so")
    _run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
          "-o", gen, os.path.join(CSRC, "datagen.cpp")])
    if verbose:
        print("built", so)
    return so


if __name__ == "__main__":
    build(verbose=True)
