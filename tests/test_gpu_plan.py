"""Trees and traversal plans built on the GPU (h2g_plan_build) against the oracle's and the
host builder's, bit for bit, and maps built from them against the oracle's products."""

import time

import numpy as np
import pytest

import plan_check as PC

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", [c[0] for c in PC.point_sets()])
def test_device_plan_equals_oracle(case):
    from h2g.plan import device_plan

    _, a, b, leaf, origin, size, sym = next(c for c in PC.point_sets() if c[0] == case)
    dp = device_plan(a, b, leaf, origin, size, symmetric=sym)
    PC.check_against_oracle(dp, a, b, leaf, origin, size, sym)


@pytest.mark.parametrize("n,geom,symmetric", [(2_000_000, "cube", False), (1_000_000, "sphere", True)])
def test_device_plan_equals_host_builder_at_full_size(n, geom, symmetric):
    """The benchmark map (2M points in the unit cube, leaves of 64) and a sphere map."""
    from h2g.plan import device_plan

    rng = np.random.default_rng(0)
    if geom == "cube":
        pts, origin, size = rng.random((n, 3)), (0.0, 0.0, 0.0), 1.0
    else:
        pts = rng.standard_normal((n, 3))
        pts /= np.linalg.norm(pts, axis=1, keepdims=True)
        origin, size = (-1.0, -1.0, -1.0), 2.0
    t0 = time.perf_counter()
    dp = device_plan(pts, None, 64, origin, size, symmetric=symmetric)
    print(f"device plan of {n} points: {time.perf_counter() - t0:.2f} s, {len(dp.pairs)} M2L pairs, "
          f"{len(dp.near_pairs)} near pairs", flush=True)
    PC.check_against_host_builder(dp, pts, None, 64, origin, size, symmetric)


@pytest.mark.parametrize("kind,petrov", [("toeplitz", False), ("symdense", True)])
def test_map_from_device_plan_equals_oracle(kind, petrov):
    """build_map(plan="device") with device-filled values, every orientation, against the oracle."""
    import h2g
    from oracle import build as OB
    from oracle import h2 as O

    rng = np.random.default_rng(21)
    a = rng.random((2500, 3))
    b = rng.random((1800, 3)) if petrov else None
    p, nodes = (4, "equi") if kind == "toeplitz" else (3, "cheb")
    core = OB.build_core(a, b, p=p, nodekind=nodes, m2l_kind=kind, leaf_size=40)
    arrays = h2g.build_map(a, b, p=p, nodekind=nodes, m2l_kind=kind, leaf_size=40, plan="device")
    A = h2g.H2Workspace(h2g.GpuH2Core(arrays), max_rhs=2)
    for op in "NTC":
        M = {"N": A, "T": h2g.transpose(A), "C": h2g.adjoint(A)}[op]
        x = rng.standard_normal((M.shape[1], 2))
        y = np.zeros((M.shape[0], 2))
        h2g.mul_(y, M, x)
        ref = O.mul(core, x, op).real
        err = np.linalg.norm(y - ref) / np.linalg.norm(ref)
        print(f"{kind} op={op}: rel err vs oracle {err:.2e}", flush=True)
        assert err < 1e-12
