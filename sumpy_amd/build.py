"""Build the native libraries of the package in-tree.

* ``libsumpyhip.so`` -- HIP kernels for gfx950 and their C-ABI
  (``include/sumpyhip.h``), built with hipcc;
* ``libsumpytree.so`` -- the host tree/traversal builder, built with g++.
"""

from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")

HIP_SOURCES = ["misc.hip", "p2p.hip", "expansions.hip", "m2l.hip", "octant.hip", "helmholtz.hip",
               "tree_gpu.hip", "h2d.hip", "qbx.hip",
               "taylor2d.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    return res


def build_hip(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    out = os.path.join(HERE, "libsumpyhip.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


def build_tree() -> str:
    out = os.path.join(HERE, "libsumpytree.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", tmp,
          os.path.join(CSRC, "tree.cpp")])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_tree()
    build_hip()


if __name__ == "__main__":
    build_all()
