"""Builds librtstereo.so (HIP kernels + C ABI, gfx950) and librtstereo_seq.so
(the reference-signature drop-in) in-tree, next to this file.

    python -m rtstereo_amd.build            # build both libraries
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OBJ = os.path.join(ROOT, "_build")
ARCH = os.environ.get("RTS_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["rts_runtime.hip", "k_cost.hip", "k_aggr.hip", "k_aw.hip", "k_wta.hip", "pipeline.hip",
               "server.hip", "batch.hip"]
HOST_SOURCES = ["png_io.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
             # packed FP32 instructions are kept out of every kernel
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             f"-I{INCLUDE}", f"-I{CSRC}"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"build step failed ({p.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build(verbose: bool = False) -> list[str]:
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(OBJ, s.replace(".hip", ".o")) for s in srcs]

    def compile_one(args):
        src, obj = args
        return _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])

    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        for out in ex.map(compile_one, zip(srcs, objs)):
            if verbose and out.strip():
                print(out)
    # host-only sources (image files on zlib)
    for src in HOST_SOURCES:
        obj = os.path.join(OBJ, src.replace(".cpp", ".o"))
        _run(["g++", "-O2", "-fPIC", "-std=c++17", "-Wall", f"-I{INCLUDE}", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    lib = os.path.join(HERE, "librtstereo.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", lib + ".tmp", *objs, "-lz"])
    os.replace(lib + ".tmp", lib)
    seq = os.path.join(HERE, "librtstereo_seq.so")
    _run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", f"-I{INCLUDE}",
          os.path.join(CSRC, "seq_dropin.cpp"), "-o", seq + ".tmp",
          f"-L{HERE}", "-lrtstereo", "-Wl,-rpath,$ORIGIN", "-Wl,--no-undefined",
          "-L/opt/rocm/lib", "-lamdhip64"])
    os.replace(seq + ".tmp", seq)
    return [lib, seq]


if __name__ == "__main__":
    for p in build(verbose="-v" in sys.argv):
        print(p)
