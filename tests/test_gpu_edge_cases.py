"""Products on degenerate maps against the oracle, plain (1e-12) and mixed (the 1e-6 gate), every
orientation: a map that is a single leaf (no far field at all), 150 coincident points among 2,000
(a 21-level tree whose deep levels hold one box each), Petrov point sets far apart (no near
pair), and zero-column products (no-ops that leave y as it was)."""

import numpy as np
import pytest

from oracle import build as OB
from oracle import h2 as O

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def _products(core, cplx, nrhs=3, seed=0):
    import h2g
    from convert import to_arrays

    c = h2g.GpuH2Core(to_arrays(core), precisions=("plain", "mixed"))
    rng = np.random.default_rng(seed)
    for prec, tol in (("plain", 1e-12), ("mixed", 1e-6)):
        ws = h2g.H2Workspace(c, max_rhs=nrhs, precision=prec)
        for op in "NTC":
            M = {"N": ws, "T": h2g.transpose(ws), "C": h2g.adjoint(ws)}[op]
            x = rng.standard_normal((M.shape[1], nrhs))
            if cplx:
                x = x + 1j * rng.standard_normal((M.shape[1], nrhs))
            if prec == "mixed":
                x = x.astype(np.complex64 if cplx else np.float32)
            y = M @ x
            ref = O.mul(core, x.astype(np.complex128 if cplx else np.float64), op)
            ref = ref if cplx else ref.real
            err = relerr(y, ref)
            print(f"{prec} {op}: rel err vs oracle {err:.2e}", flush=True)
            assert err < tol, (prec, op, err)
        ws.close()
    return c


def test_single_leaf_map_has_no_far_field():
    pts = np.random.default_rng(1).random((37, 3))
    core = OB.build_core(pts, p=3, nodekind="equi", m2l_kind="toeplitz", leaf_size=64)
    assert len(core.pairs) == 0
    _products(core, cplx=False)


@pytest.mark.parametrize("kernel,kind,nodes", [("laplace", "toeplitz", "equi"), ("helmholtz", "dense", "cheb")])
def test_coincident_points(kernel, kind, nodes):
    rng = np.random.default_rng(2)
    pts = np.concatenate([np.full((150, 3), 0.3), rng.random((2000, 3))])
    core = OB.build_core(pts, p=3, nodekind=nodes, m2l_kind=kind, kernel=kernel,
                         wavenumber=2.0 if kernel == "helmholtz" else 0.0, leaf_size=40)
    assert core.test.n_levels >= 20
    _products(core, cplx=kernel == "helmholtz")


@pytest.mark.parametrize("alpha", [1.0, 0.7 - 0.2j])
def test_helmholtz_real_transfers_and_collections(alpha):
    """Helmholtz maps whose interpolation transfers are real run them on the real separable kernels
    over planar rows; with a real kernel factor their collections are kept real too (real leaf
    kernels), with a complex one they stay complex (the rows are converted after P2M and before
    L2P).  Petrov, Toeplitz couplings, every orientation and the 5-argument form."""
    import h2g
    from convert import to_arrays

    rng = np.random.default_rng(12)
    a = rng.random((1400, 3))
    b = rng.random((1000, 3))
    core = OB.build_core(a, b, p=3, nodekind="equi", m2l_kind="toeplitz", kernel="helmholtz", wavenumber=5.0,
                         leaf_size=24, alpha=alpha)
    _products(core, cplx=True, nrhs=5)
    c = h2g.GpuH2Core(to_arrays(core), precisions=("plain",))
    ws = h2g.H2Workspace(c, max_rhs=5)
    s1, s2 = 1.3 - 0.4j, -0.6 + 0.2j
    for op in "NTC":
        M = {"N": ws, "T": h2g.transpose(ws), "C": h2g.adjoint(ws)}[op]
        x = rng.standard_normal((M.shape[1], 5)) + 1j * rng.standard_normal((M.shape[1], 5))
        y0 = rng.standard_normal((M.shape[0], 5)) + 1j * rng.standard_normal((M.shape[0], 5))
        y = y0.copy()
        h2g.mul_(y, M, x, s1, s2)
        assert relerr(y, O.mul(core, x, op, s1, s2, y0)) < 1e-12, op
    ws.close()


def test_petrov_sets_without_near_pairs():
    rng = np.random.default_rng(3)
    a = 0.2 * rng.random((800, 3))
    b = 0.8 + 0.2 * rng.random((600, 3))
    core = OB.build_core(a, b, p=3, nodekind="cheb", m2l_kind="dense", leaf_size=32)
    assert len(core.near_pairs) == 0
    _products(core, cplx=False)


def test_zero_column_products_leave_y():
    import h2g

    pts = np.random.default_rng(4).random((500, 3))
    core = OB.build_core(pts, p=3, nodekind="equi", m2l_kind="toeplitz", leaf_size=32)
    c = _products(core, cplx=False, nrhs=2)
    ws = h2g.H2Workspace(c, max_rhs=2)
    x = np.zeros((500, 0))
    for fn in (h2g.mul_, h2g.farmul_, h2g.nearmul_):
        y = np.zeros((500, 0))
        fn(y, ws, x)
        fn(y, ws, x, 2.0, 3.0)
        assert y.shape == (500, 0)
    ws.close()
