"""Builds libsparkx_hb.so in-tree with hipcc for gfx950.

Run: python -m sparkx_amd.build   (also called by __graft_entry__.build()).
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsparkx_hb.so")

SOURCES = ["hb_kernels.hip", "hb_api.hip", "hb_batch.hip", "hb_mlp.hip", "hb_exact.hip", "hb_eval.hip", "hb_layout.hip", "hb_ingest.hip", "hb_walk.hip", "hb_graphgen.cpp"]
HEADERS = ["hb_device.h", "hb_kernels.h", "hb_union.h", "hll_bias_table.inc"]
ABI_HEADERS = ["sparkx_hb.h", "sparkx_hb_graphgen.h"]  # under include/ at the repo root

# -ffp-contract=off: the f64 accumulators and f32 estimates must not fuse a*b+c
# (the reference never does); packed-fp32 instruction selection is disabled.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-pthread"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def source_digest() -> str:
    """SHA-256 over the compile flags and every file compiled into the library (sources,
    kernel headers, the C ABI headers), in a fixed order.  The build passes it to the
    compiler and the library reports it (hb_build_digest), so a test can show that the
    binary it loaded was built from the files beside it."""
    h = hashlib.sha256(" ".join(FLAGS).encode())
    inc = os.path.join(os.path.dirname(HERE), "include")
    files = [os.path.join(CSRC, f) for f in (*SOURCES, *HEADERS)] + [os.path.join(inc, f) for f in ABI_HEADERS]
    for path in files:
        with open(path, "rb") as f:
            data = f.read()
        h.update(os.path.basename(path).encode() + b"\0" + len(data).to_bytes(8, "little") + data)
    return h.hexdigest()


def sanitizer_runtime(kind: str) -> str:
    """Path of the ROCm clang's shared sanitizer runtime for `kind` (address, thread)."""
    clang = os.path.join(os.path.dirname(os.path.realpath(hipcc())), "..", "llvm", "bin", "clang")
    if not os.path.exists(clang):
        clang = "/opt/rocm/llvm/bin/clang"
    name = f"libclang_rt.{'asan' if kind == 'address' else 'tsan'}-x86_64.so"
    res = subprocess.run([clang, "-print-resource-dir"], capture_output=True, text=True, check=True).stdout.strip()
    rt = subprocess.run([clang, "--print-runtime-dir"], capture_output=True, text=True, check=True).stdout.strip()
    for d in (rt, os.path.join(res, "lib", "linux")):
        if os.path.exists(os.path.join(d, name)):
            return os.path.join(d, name)
    raise FileNotFoundError(f"{name} not found under {rt} or {res}/lib/linux")


def build(verbose: bool = False, sanitize: str | None = None) -> str:
    """Builds the library; with `sanitize` ("address" or "thread") the host code is
    instrumented (device code is not; "address" adds UndefinedBehaviorSanitizer) and the
    result is libsparkx_hb_<sanitize>.so, linked against the ROCm clang's sanitizer runtime
    (libclang_rt.asan / .tsan) -- for CPU test runs under SPARKX_HB_LIB with that runtime
    in LD_PRELOAD."""
    import tempfile
    compile_flags = [f for f in FLAGS if f != "-shared"] + [f'-DHB_BUILD_DIGEST="{source_digest()}"']
    target = LIB
    if sanitize:
        if sanitize not in ("address", "thread"):
            raise ValueError(f"sanitize={sanitize!r}: address or thread")
        checks = "address,undefined" if sanitize == "address" else sanitize
        compile_flags += ["-O1", "-g", "-fno-omit-frame-pointer", "-Xarch_host", "-fsanitize=" + checks]
        target = os.path.join(HERE, f"libsparkx_hb_{sanitize}.so")
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
        if sanitize:  # host objects call clang's sanitizer runtime: link against it
            rocm = os.path.dirname(os.path.dirname(os.path.realpath(hipcc())))
            rt = sanitizer_runtime(sanitize)
            link = ["g++", "-shared", "-pthread", *objs, "-L" + os.path.join(rocm, "lib"), "-lamdhip64",
                    "-Wl,-rpath," + os.path.join(rocm, "lib"), rt, "-Wl,-rpath," + os.path.dirname(rt),
                    "-o", target + ".tmp"]
        else:
            link = [hipcc(), *FLAGS, *objs, "-o", target + ".tmp"]
        res = subprocess.run(link, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"link failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    os.replace(target + ".tmp", target)
    return target


if __name__ == "__main__":
    san = next((a.split("=", 1)[1] for a in sys.argv[1:] if a.startswith("--sanitize=")), None)
    print(build(verbose="-v" in sys.argv, sanitize=san))
