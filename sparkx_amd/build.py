"""Builds libsparkx_hb.so in-tree with hipcc for gfx950.

Run: python -m sparkx_amd.build   (also called by __graft_entry__.build()).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsparkx_hb.so")

SOURCES = ["hb_kernels.hip", "hb_api.hip", "hb_batch.hip", "hb_mlp.hip", "hb_exact.hip", "hb_eval.hip", "hb_layout.hip", "hb_ingest.hip", "hb_walk.hip", "hb_graphgen.cpp"]
HEADERS = ["hb_device.h", "hb_kernels.h", "hb_union.h", "hll_bias_table.inc"]

# -ffp-contract=off: the f64 accumulators and f32 estimates must not fuse a*b+c
# (the reference never does); packed-fp32 instruction selection is disabled.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-pthread"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    import tempfile
    compile_flags = [f for f in FLAGS if f != "-shared"]
    with tempfile.TemporaryDirectory(prefix="sparkx_hb_build_") as tmp:
        objs, procs = [], []
        for src in SOURCES:  # one hipcc per source, in parallel
            obj = os.path.join(tmp, src + ".o")
            cmd = [hipcc(), *compile_flags, "-c", os.path.join(CSRC, src), "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
            objs.append(obj)
        errors = []
        for src, pr in procs:
            out, err = pr.communicate()
            if pr.returncode != 0:
                errors.append(f"{src}:\n{out}\n{err}")
        if errors:
            raise RuntimeError("hipcc failed:\n" + "\n".join(errors))
        link = [hipcc(), *FLAGS, *objs, "-o", LIB + ".tmp"]
        res = subprocess.run(link, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"link failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
