"""Build the oracle (test infrastructure only).

  oracle/liboracle.so         C restatement of SCIP's propagation (always built)
  oracle/_ref/libdbldbl_ref.so the reference's own src/scip/dbldblarith.h macros, compiled
                               from the SCIP source tree when it is present (this container);
                               the rest of the reference needs its cmake-generated headers
                               and is not built.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
REFROOT = "/root/reference"
CFLAGS = ["-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wno-unused-parameter"]


def build(verbose: bool = True):
    out = os.path.join(HERE, "liboracle.so")
    subprocess.run(["gcc", *CFLAGS, "-o", out + ".tmp", os.path.join(HERE, "oracle.c"), "-lm"], check=True)
    os.replace(out + ".tmp", out)
    if verbose:
        print(f"built {out}")
    hdr = os.path.join(REFROOT, "src", "scip", "dbldblarith.h")
    if os.path.exists(hdr):
        os.makedirs(os.path.join(HERE, "_ref"), exist_ok=True)
        ref = os.path.join(HERE, "_ref", "libdbldbl_ref.so")
        subprocess.run(["gcc", *CFLAGS, f"-I{os.path.join(REFROOT, 'src')}", "-o", ref + ".tmp",
                        os.path.join(HERE, "ref_dbldbl.c"), "-lm"], check=True)
        os.replace(ref + ".tmp", ref)
        if verbose:
            print(f"built {ref}")


if __name__ == "__main__":
    build()
