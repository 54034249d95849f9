"""Builds the host replay of libh2g (TEST INFRASTRUCTURE ONLY): the library's api.hip, kernels.hip
and method.hip compiled by g++ as C++ against the host emulation in hip/hip_runtime.h, with
AddressSanitizer, into tests/cpu_replay/_build/libh2g_replay.so.  Nothing in the product loads it."""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "h2factory.jl_amd", "csrc")
OUT = os.path.join(HERE, "_build", "libh2g_replay.so")
SOURCES = [os.path.join(CSRC, s) for s in ("api.hip", "kernels.hip", "method.hip", "plan.hip", "operators.hip")] + [os.path.join(HERE, "stubs.cpp")]


def asan_runtime():
    return subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()


def build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = ["g++", "-std=c++20", "-O1", "-g", "-fPIC", "-shared", "-fsanitize=address", "-fno-omit-frame-pointer",
           "-pthread", "-I" + HERE, "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-iquote", CSRC]
    for s in SOURCES:
        if s.endswith(".hip"):
            # the dynamic-LDS declarations (extern __shared__) refer to the emulation's arrays;
            # every other __shared__ array becomes a function-local static.  Line numbers are kept.
            src = open(s).read().replace("extern __shared__", "extern")
            s = os.path.join(os.path.dirname(OUT), os.path.basename(s) + ".cpp")
            with open(s, "w") as fh:
                fh.write(src)
        cmd += ["-x", "c++", s]
    cmd += ["-o", OUT]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr[-20000:])
        raise RuntimeError("g++ failed building the host replay")
    return OUT


if __name__ == "__main__":
    print(build())
