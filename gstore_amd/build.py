"""Build recipe for the engine: libgstore_amd.so (HIP, gfx950) and the
gstoreu / gstored command-line tools (C++ host over the C ABI).

    python -m gstore_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIP_SOURCES = ["gs_graph.hip", "gs_pagerank.hip", "gs_bfs.hip", "gs_bfs_batch.hip", "gs_wcc.hip", "gs_kcore.hip",
               "gs_serve.hip", "gs_stream.hip", "gs_tune.hip", "gs_adj.hip", "gs_pb.hip", "gs_jobs.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -packed-fp32-ops: keep the compiler from emitting packed FP32 VALU ops.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, cwd=cwd, check=True)


def build_library() -> str:
    out = os.path.join(HERE, "libgstore_amd.so")
    tmp = out + ".tmp"
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    _run([HIPCC, *HIP_FLAGS, "-shared", *srcs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_cli() -> list:
    bindir = os.path.join(HERE, "bin")
    os.makedirs(bindir, exist_ok=True)
    outs = []
    for name, directed in (("gstoreu", 0), ("gstored", 1)):
        out = os.path.join(bindir, name)
        _run(["g++", "-O2", "-std=c++17", f"-DGS_CLI_DIRECTED={directed}", f"-I{ROOT}/include",
              os.path.join(CSRC, "gstore_cli.cpp"), "-o", out + ".tmp", f"-L{HERE}", "-lgstore_amd",
              "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,/opt/rocm/lib"])
        os.replace(out + ".tmp", out)
        outs.append(out)
    return outs


def build_all() -> None:
    build_library()
    build_cli()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
