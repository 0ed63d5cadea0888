"""Build libfvhoe_hip.so in-tree with hipcc for gfx950.

    python -m fvhoe_amd.csrc.build          (or: python fvhoe_amd/csrc/build.py)

FP64 parity with the reference needs -ffp-contract=off (NumPy never fuses a
multiply into an add); packed FP32 instructions are disabled for the whole
library.
"""

import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
REPO = os.path.dirname(PKG)
OUT = os.path.join(PKG, "libfvhoe_hip.so")
SOURCES = ["elementary.hip", "stage.hip", "predictor.hip", "ensemble.hip"]

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(REPO, "include"), "-I", HERE,
]
# launch-shape overrides for a sweep, e.g. FVHOE_DEFINES="-DFVHOE_DETECT_ROWS=4" (stage.hip documents them)
FLAGS += [d for d in os.environ.get("FVHOE_DEFINES", "").split() if d.startswith("-DFVHOE_")]


def build(verbose: bool = True) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objs = []
    t0 = time.time()
    procs = []
    for src in SOURCES:
        obj = os.path.join(HERE, "_build", src.replace(".hip", ".o"))
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        cmd = [hipcc, *[f for f in FLAGS if f != "-shared"], "-c", os.path.join(HERE, src), "-o", obj]
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            raise RuntimeError(f"hipcc failed on {src}")
        if verbose and out.strip():
            lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
            if lines:
                sys.stderr.write("\n".join(lines[:40]) + "\n")
    tmp = OUT + ".tmp"
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT} in {time.time() - t0:.1f}s", flush=True)
    return OUT


if __name__ == "__main__":
    build()
