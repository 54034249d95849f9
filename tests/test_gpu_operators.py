"""Operators built on the GPU (h2g_method_operators) against the oracle's builders, and maps
built entirely on the device (plan, operators, near values, collections) against the oracle's
products."""

import numpy as np
import pytest

import ops_check as OC

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", OC.CASES + [(6, "equi", "laplace", 0.0, "toeplitz"),
                                             (6, "cheb", "helmholtz", 10.0, "dense"),
                                             (6, "equi", "laplace", 0.0, "symtoeplitz")])
def test_device_operators_equal_oracle(case):
    e = OC.check(*case)
    print(f"{case}: largest deviation vs oracle {e:.2e}", flush=True)


@pytest.mark.parametrize("kind,p,nodes,kernel,k,petrov", [
    ("toeplitz", 4, "equi", "laplace", 0.0, False),
    ("symtoeplitz", 3, "equi", "laplace", 0.0, False),
    ("dense", 3, "cheb", "helmholtz", 10.0, False),
    ("symdense", 3, "cheb", "laplace", 0.0, True),
    ("toeplitz", 4, "equi", "helmholtz", 10.0, False),
    ("toeplitz", 3, "equi", "helmholtz", 2.0, True),
])
def test_map_built_on_device_equals_oracle(kind, p, nodes, kernel, k, petrov):
    import h2g
    from oracle import build as OB
    from oracle import h2 as O

    rng = np.random.default_rng(31)
    a = rng.random((2400, 3))
    b = rng.random((1700, 3)) if petrov else None
    core = OB.build_core(a, b, p=p, nodekind=nodes, m2l_kind=kind, kernel=kernel, wavenumber=k, leaf_size=40)
    arrays = h2g.build_map(a, b, p=p, nodekind=nodes, m2l_kind=kind, kernel=kernel, wavenumber=k, leaf_size=40,
                           plan="device", ops="device")
    A = h2g.H2Workspace(h2g.GpuH2Core(arrays), max_rhs=2)
    cplx = kernel == "helmholtz"
    for op in "NTC":
        M = {"N": A, "T": h2g.transpose(A), "C": h2g.adjoint(A)}[op]
        x = rng.standard_normal((M.shape[1], 2)) + (1j * rng.standard_normal((M.shape[1], 2)) if cplx else 0)
        y = np.zeros((M.shape[0], 2), dtype=x.dtype)
        h2g.mul_(y, M, x)
        ref = O.mul(core, x, op)
        ref = ref if cplx else ref.real
        err = np.linalg.norm(y - ref) / np.linalg.norm(ref)
        print(f"{kind} {kernel} op={op}: rel err vs oracle {err:.2e}", flush=True)
        assert err < 1e-12


def test_benchmark_map_operators_equal_host_builder():
    """The 1,580 Toeplitz spectra and the transfers of the N = 2M, p = 6 benchmark map."""
    from h2g import method as PM

    pts = np.random.default_rng(0).random((2_000_000, 3))
    dev = PM.build_map(pts, p=6, nodekind="equi", m2l_kind="toeplitz", leaf_size=64, plan="device", ops="device")
    host_ops = []
    for l in np.unique(dev.ops_keys[:, 0]):
        sel = dev.ops_keys[:, 0] == l
        host_ops.append(PM.toeplitz_spectra("laplace", 0.0, 6, 1.0 / 2.0 ** l, dev.ops_keys[sel, 1:]))
    host_ops = np.concatenate(host_ops)
    e = np.max(np.abs(dev.ops - host_ops), axis=(1, 2, 3)) / np.max(np.abs(host_ops), axis=(1, 2, 3))
    print(f"{len(host_ops)} spectra: largest deviation {e.max():.2e}", flush=True)
    assert e.max() < 1e-13
    tr = PM._transfers(dev.test.tree, PM.nodes1d(6, "equi"))
    assert np.max(np.abs(dev.test.transfers - tr)) < 1e-13
