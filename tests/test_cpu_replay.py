"""Host replay of libh2g under AddressSanitizer (CPU; no GPU).

The library's own host code (api.hip) and generic kernels (kernels.hip, method.hip) are compiled
by g++ against the host emulation of HIP in tests/cpu_replay/hip/hip_runtime.h, so every global
and LDS access of those kernels is checked against its allocation while the products are compared
with the oracle.  This is a checker of the kernels' indexing, not an execution path of the product:
the product loads only the gfx950 library.
"""

import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPLAY = os.path.join(HERE, "cpu_replay")


@pytest.fixture(scope="module")
def replay_lib():
    sys.path.insert(0, REPLAY)
    import build as B  # tests/cpu_replay/build.py

    sys.path.remove(REPLAY)
    return B.build(), B.asan_runtime()


@pytest.mark.parametrize("case", ["helmholtz_sphere", "helmholtz_petrov_alpha", "laplace_device_fill",
                                  "random_petrov_symdense",
                                  "random_petrov_diagonal", "noalloc", "diagnostics",
                                  "device_plan", "device_operators"])
def test_replay_in_bounds_and_equal_to_oracle(replay_lib, case):
    """helmholtz_sphere is the map of the skipped GPU Helmholtz test (ComplexF64, p = 3 Chebyshev,
    dense couplings); laplace_device_fill compares the device-filled near values and collections
    (h2g_method_fill) with the oracle's host assembly; random_petrov_symdense covers the permuted
    box GEMM with separate trees.  Any out-of-bounds access aborts the child with an ASan report."""
    _, asan = replay_lib
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    r = subprocess.run([sys.executable, os.path.join(REPLAY, "run_case.py"), case], env=env, capture_output=True,
                       text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-6000:]
    assert "vs oracle" in r.stdout
