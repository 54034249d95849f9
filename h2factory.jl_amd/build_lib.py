"""Builds libh2g.so for gfx950 with hipcc (in tree: h2factory.jl_amd/lib/libh2g.so).

Each source compiles to its own object in parallel (no device code crosses translation units),
then hipcc links the shared library.
"""

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = ["api.hip", "kernels.hip", "leaf_mfma.hip", "method.hip", "m2l_lattice.hip", "m2l_mfma.hip",
           "dense_m2l.hip", "noalloc.hip", "plan.hip", "prims.hip", "operators.hip"]
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _filter(stderr):
    return "\n".join(l for l in stderr.splitlines() if "packed-fp32-ops" not in l)


def build(verbose=False):
    out = os.path.join(HERE, "lib", "libh2g.so")
    objdir = os.path.join(HERE, "lib", "obj")
    os.makedirs(objdir, exist_ok=True)
    srcs = [os.path.join(HERE, "csrc", s) for s in SOURCES if os.path.exists(os.path.join(HERE, "csrc", s))]
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(HERE, "csrc")]

    def compile_one(src):
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        r = subprocess.run(["hipcc", *FLAGS, *inc, "-c", src, "-o", obj], capture_output=True, text=True)
        return src, obj, r

    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(compile_one, srcs))
    failed = False
    for src, _, r in results:
        err = _filter(r.stderr)
        if r.returncode != 0:
            sys.stderr.write(f"{os.path.basename(src)}:\n{err}\n")
            failed = True
        elif verbose and err.strip():
            sys.stderr.write(err + "\n")
    if failed:
        raise RuntimeError("hipcc failed building libh2g.so")
    r = subprocess.run(["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", *[o for _, o, _ in results],
                        "-o", out + ".tmp"], capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(_filter(r.stderr) + "\n")
        raise RuntimeError("hipcc failed linking libh2g.so")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
