"""Pins of the oracle's traversal: independent dense assembly, the reference's worked example,
the exact near/far partition and method accuracy against the O(N^2) direct sum."""

import numpy as np
import pytest

from oracle import build as OB
from oracle import h2 as O
from oracle import method as M
from oracle import tree as OT


def relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


RNG = np.random.default_rng(0)
PA = RNG.random((300, 3))
PB = RNG.random((240, 3))


@pytest.mark.parametrize("kind", ["dense", "toeplitz", "symdense", "symtoeplitz", "diagonal"])
@pytest.mark.parametrize("petrov", [False, True])
def test_traversal_equals_dense_assembly(kind, petrov):
    """Random complex operators: traversal in every orientation == composite-basis assembly."""
    core = OB.random_core(PA, PB if petrov else None, p=2, m2l_kind=kind, leaf_size=8, alpha=0.7 - 0.2j)
    A = O.dense_matrix(core)
    rng = np.random.default_rng(1)
    X = rng.standard_normal((A.shape[1], 2)) + 1j * rng.standard_normal((A.shape[1], 2))
    Xt = rng.standard_normal((A.shape[0], 2)) + 1j * rng.standard_normal((A.shape[0], 2))
    assert relerr(O.mul(core, X, "N"), A @ X) < 1e-13
    assert relerr(O.mul(core, Xt, "T"), A.T @ Xt) < 1e-13
    assert relerr(O.mul(core, Xt, "C"), A.conj().T @ Xt) < 1e-13
    a, b = 2 - 1j, 3 + 2j
    y0 = rng.standard_normal(A.shape[0]) + 0j
    assert relerr(O.mul(core, X[:, 0], "N", a, b, y0), a * (A @ X[:, 0]) + b * y0) < 1e-13


def test_worked_example_is_exact():
    """docs/src/implementing/example.md:205-256."""
    core, s, F = OB.worked_example_core()
    assert len(core.pairs) > 0 and not core.test.is_leaf[core.pairs[:, 1]].all()
    assert relerr(O.dense_matrix(core), F) < 1e-14
    x = np.sin(np.arange(1, 33)) + 1j * np.cos(2 * np.arange(1, 33))
    for op, R in (("N", F), ("T", F.T), ("C", F.conj().T)):
        assert relerr(O.mul(core, x, op), R @ x) < 1e-13
        out = np.full(32, 2.0 + 1j)
        assert relerr(O.mul(core, x, op, 2 - 1j, 3 + 2j, out), (2 - 1j) * (R @ x) + (3 + 2j) * out) < 1e-13


def test_partition_covers_every_pair_once():
    """Near blocks + far pairs (expanded to points) tile the full interaction matrix exactly once."""
    for petrov in (False, True):
        t = OT.build_tree(PA, 8)
        c = OT.build_tree(PB, 8) if petrov else t
        cover = np.zeros((t.points.shape[0], c.points.shape[0]), dtype=int)
        for b, cc in OT.near_pairs(t, c):
            cover[t.begin[b]:t.end[b], c.begin[cc]:c.end[cc]] += 1
        for b, cc in OT.m2l_pairs(t, c):
            cover[t.begin[b]:t.end[b], c.begin[cc]:c.end[cc]] += 1
        assert (cover == 1).all()


def test_near_block_values_are_the_kernel():
    pts = RNG.random((500, 3))
    core = OB.build_core(pts, p=3, leaf_size=32)
    G = M.kernel_matrix("laplace", 0.0, pts, pts)
    A = np.zeros_like(G)
    for (b, c), B in zip(core.near_pairs, core.near_values):
        A[np.ix_(core.test.perm[core.test.begin[b]:core.test.end[b]],
                 core.trial.perm[core.trial.begin[c]:core.trial.end[c]])] = B
    mask = A != 0
    assert np.array_equal(A[mask], G[mask])


@pytest.mark.parametrize("kind,p,nodes,tol", [("dense", 4, "cheb", 1e-3), ("toeplitz", 6, "equi", 1e-4),
                                              ("symtoeplitz", 6, "equi", 1e-4), ("symdense", 4, "equi", 1e-3)])
def test_method_accuracy_vs_direct_sum(kind, p, nodes, tol):
    pts = np.random.default_rng(2).random((2500, 3))
    x = np.random.default_rng(3).standard_normal(2500)
    core = OB.build_core(pts, p=p, nodekind=nodes, m2l_kind=kind, leaf_size=64)
    y = O.mul(core, x).real
    yd = O.direct_sum("laplace", 0.0, pts, pts, x)
    assert relerr(y, yd) < tol


def test_symmetry_reduction_is_exact_for_laplace():
    pts = np.random.default_rng(5).random((2500, 3))
    x = np.random.default_rng(6).standard_normal(2500)
    y0 = O.mul(OB.build_core(pts, p=4, nodekind="equi", m2l_kind="toeplitz"), x).real
    for kind in ("symtoeplitz", "symdense", "dense"):
        y = O.mul(OB.build_core(pts, p=4, nodekind="equi", m2l_kind=kind), x).real
        assert relerr(y, y0) < 1e-12, kind


def test_direct_sum_matches_double_loop():
    pts = RNG.random((40, 3))
    x = RNG.standard_normal(40)
    ref = np.array([sum(x[j] / (4 * np.pi * np.linalg.norm(pts[i] - pts[j])) for j in range(40) if j != i)
                    for i in range(40)])
    assert np.allclose(O.direct_sum("laplace", 0.0, pts, pts, x), ref, rtol=1e-13)


def test_reference_farmul_fixture():
    """test/farinteractions/test_farmul.jl:11-97: the hand-derived far matrix F (F[left, right] =
    F[right, left] = c = 2 - i) in every orientation and the 5-argument accumulation."""
    core, _, F = OB.farmul_fixture_core()
    rng = np.random.default_rng(0)
    x = rng.standard_normal(8) + 1j * rng.standard_normal(8)
    assert relerr(O.farmul(core, x, "N"), F @ x) < 1e-14
    assert relerr(O.farmul(core, x, "T"), F.T @ x) < 1e-14
    assert relerr(O.farmul(core, x, "C"), F.conj().T @ x) < 1e-14
    y = rng.standard_normal(8) + 1j * rng.standard_normal(8)
    a, b = 0.7 + 0.3j, -0.4 + 0.2j
    assert relerr(O.farmul(core, x, "N", a, b, y), a * F @ x + b * y) < 1e-14
    # a conjugation dropped from the adjoint would give F^T x instead
    assert relerr(F.T @ x, F.conj().T @ x) > 0.5


def test_farmul5_without_far_pairs_writes_no_row():
    """With no far pair every receiving leaf is uninitialised and the 5-argument farmul! leaves y
    exactly as it was -- not even scaled by beta (src/farinteractions/traversalhelpers.jl:103-112)."""
    core = OB.no_far_core()
    assert len(core.pairs) == 0 and len(core.near_pairs) > 0
    y0 = np.arange(5) + 1j
    assert np.array_equal(O.farmul(core, np.ones(5) + 0j, "N", 2.0, 3.0, y0), y0)
    assert np.array_equal(O.farmul(core, np.ones(5) + 0j, "N"), np.zeros(5))


def test_helmholtz_method_accuracy_vs_direct_sum():
    """Helmholtz k = 10 on the sphere, Chebyshev p = 5: the H2 product against the O(N^2) sum."""
    pts = np.random.default_rng(8).standard_normal((1500, 3))
    pts = 0.5 + 0.45 * pts / np.linalg.norm(pts, axis=1, keepdims=True)
    x = np.random.default_rng(3).standard_normal(1500) + 1j * np.random.default_rng(4).standard_normal(1500)
    core = OB.build_core(pts, p=5, nodekind="cheb", m2l_kind="dense", kernel="helmholtz", wavenumber=10.0,
                         leaf_size=32)
    for op, conj in (("N", False), ("C", True)):
        y = O.mul(core, x, op)
        yd = O.direct_sum("helmholtz", 10.0, pts, pts, x)
        if conj:  # the Helmholtz kernel matrix is symmetric: adjoint = conj(A)
            yd = np.conj(O.direct_sum("helmholtz", 10.0, pts, pts, np.conj(x)))
        assert relerr(y, yd) < 1e-3, op
