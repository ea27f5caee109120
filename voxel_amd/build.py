"""Builds voxel_amd/libvoxel_c.so (HIP kernels for gfx950 + the C ABI) in-tree.

    python -m voxel_amd.build        # or:  python voxel_amd/build.py

Every source in voxel_amd/csrc is compiled with hipcc (in parallel), objects go
to voxel_amd/_build/, and the shared library is linked next to this file.  The
device code of the vector words (vx_expr_abi.hpp, vx_expr_ops.hpp) is also
written into _build/vx_expr_src.inc as a string literal, the source the
library hands to hipRTC when it compiles a loop body at run time.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libvoxel_c.so")
ARCH = os.environ.get("VOXEL_OFFLOAD_ARCH", "gfx950")

# Packed-FP32 instructions are kept out of the generated code.
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found (ROCm is required to build voxel_amd)")


JIT_HEADERS = ("vx_expr_abi.hpp", "vx_expr_ops.hpp")


def _write_jit_source() -> None:
    """_build/vx_expr_src.inc: the JIT headers' text (without #include and
    #pragma once lines) as one raw string literal."""
    text = []
    for h in JIT_HEADERS:
        with open(os.path.join(SRC, h)) as f:
            for line in f:
                if line.lstrip().startswith(("#include", "#pragma once")):
                    continue
                text.append(line)
    body = "".join(text)
    assert ')VXJIT"' not in body
    out = 'R"VXJIT(' + body + ')VXJIT"\n'
    path = os.path.join(OUT, "vx_expr_src.inc")
    old = open(path).read() if os.path.exists(path) else None
    if old != out:
        with open(path, "w") as f:
            f.write(out)


def _compile(src: str) -> str:
    obj = os.path.join(OUT, os.path.basename(src) + ".o")
    cmd = [hipcc()] + COMMON + ["-I", os.path.join(HERE, "..", "include"), "-I", SRC, "-I", OUT]
    if src.endswith(".hip"):
        cmd += DEVICE + ["-x", "hip"]
    cmd += ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    _write_jit_source()
    srcs = sorted(os.path.join(SRC, f) for f in os.listdir(SRC) if f.endswith((".hip", ".cpp")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lhiprtc", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
