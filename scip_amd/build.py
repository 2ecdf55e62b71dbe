"""Build scip_amd/libscipamd.so (HIP engine + C generators) for gfx950.

Usage: python scip_amd/build.py
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(ROOT, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off: no FMA contraction, every FP64 operation is the written IEEE operation
# (bit-identical to a host build); packed FP32 instructions are disabled for all HIP code.
HIPFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{os.path.join(HERE, 'csrc')}",
]
CFLAGS = ["-O2", "-fPIC", "-ffp-contract=off", "-Wall", f"-I{os.path.join(ROOT, 'include')}"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    eng_o = os.path.join(BUILD, "engine.o")
    gen_o = os.path.join(BUILD, "gen.o")
    srv_o = os.path.join(BUILD, "serve.o")
    mps_o = os.path.join(BUILD, "mps.o")
    lib = os.path.join(HERE, "libscipamd.so")
    steps = [
        [HIPCC, *HIPFLAGS, "-c", os.path.join(HERE, "csrc", "engine.hip"), "-o", eng_o],
        ["gcc", *CFLAGS, "-c", os.path.join(HERE, "csrc", "gen.c"), "-o", gen_o],
        ["g++", "-std=c++17", *CFLAGS, "-c", os.path.join(HERE, "csrc", "serve.cpp"), "-o", srv_o],
        ["gcc", *CFLAGS, "-c", os.path.join(HERE, "csrc", "mps.c"), "-o", mps_o],
        [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", eng_o, gen_o, srv_o, mps_o, "-o", lib + ".tmp"],
    ]
    for cmd in steps:
        out = _run(cmd)
        if verbose and out:
            print(out)
    os.replace(lib + ".tmp", lib)
    if verbose:
        print(f"built {lib}")
    build_plugin_testhost(verbose)
    return lib


def build_plugin_testhost(verbose: bool = True) -> str:
    """The SCIP propagator plugin (plugin/prop_gpulinear.c) linked against the test double of
    the SCIP API it calls (plugin/testhost), giving plugin/libgpulinear_testhost.so for tests.
    In a SCIP build the plugin source is compiled against SCIP's own headers instead."""
    plug = os.path.join(HERE, "plugin")
    out = os.path.join(plug, "libgpulinear_testhost.so")
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-Wall", "-Wextra", "-Wno-unused-parameter",
           f"-I{os.path.join(plug, 'testhost')}", f"-I{plug}", f"-I{os.path.join(ROOT, 'include')}",
           os.path.join(plug, "prop_gpulinear.c"), os.path.join(plug, "testhost", "testhost.c"),
           f"-L{HERE}", "-lscipamd", "-Wl,-rpath,$ORIGIN/..", "-lm", "-o", out + ".tmp"]
    msg = _run(cmd)
    if verbose and msg:
        print(msg)
    os.replace(out + ".tmp", out)
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
