"""Builds libxba.so for gfx950 in-tree (hipcc; no JIT, no cache outside the package directory).

    python -m xclean_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "xba_api.hip")
OUT = os.path.join(HERE, "libxba.so")
DEPS = [os.path.join(HERE, "csrc", f) for f in ("xba_api.hip", "xba_lin.hip", "xba_dense.hip", "xba_internal.h",
                                               "xba_np.hip", "xba_bigchol.hip", "xba_fast.hip", "xba_adj.hip",
                                               "xba_preint.hip", "xba_marg.hip")]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         # packed FP32 instructions stay off (v_pk_* FP32 forms are not used anywhere in this library)
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc] + FLAGS + ["-o", OUT + ".tmp", SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libxba.so")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
