"""Pins of the oracle's method data against mathematics and the reference's own counts."""

import numpy as np
import pytest

from oracle import method as M


@pytest.mark.parametrize("kind", ["cheb", "equi"])
@pytest.mark.parametrize("p", [2, 4, 6])
def test_lagrange_reproduces_polynomials(kind, p):
    nodes = M.nodes1d(p, kind)
    t = np.linspace(-1, 1, 17)
    L = M.lagrange1d(nodes, t)
    for deg in range(p):
        assert np.allclose(L @ nodes ** deg, t ** deg, atol=1e-12)
    assert np.allclose(M.lagrange1d(nodes, nodes), np.eye(p), atol=1e-14)


@pytest.mark.parametrize("sector", range(8))
def test_transfer_is_parent_basis_at_child_nodes(sector):
    """T_s[k, j] = S^parent_k(child node j): moving a polynomial's samples from child to parent."""
    p = 3
    nodes = M.nodes1d(p, "cheb")
    T = M.transfer(nodes, sector)
    sig = np.array([(sector >> a) & 1 for a in range(3)]) * 2 - 1
    child = (M.grid_points(nodes) + sig[None, :]) / 2.0
    direct = M.tensor_basis(nodes, child)
    assert np.allclose(T, direct, atol=1e-14)
    # a degree-(p-1) polynomial sampled on the parent grid, evaluated at the child grid
    f = lambda x: 1 + x[:, 0] - 2 * x[:, 1] ** 2 + x[:, 0] * x[:, 2] + 0.5 * x[:, 2] ** 2
    assert np.allclose(T.T @ f(M.grid_points(nodes)), f(child), atol=1e-12)


@pytest.mark.parametrize("offset", [(2, 0, 0), (3, -2, 1), (-2, -3, 3), (0, 2, -2)])
def test_toeplitz_spectrum_equals_dense_kernel(offset):
    """Equispaced M2L: dense kernel matrix == multilevel Toeplitz of kappa == FFT application."""
    p, w = 4, 0.125
    nodes = M.nodes1d(p, "equi")
    dense = M.m2l_dense("laplace", 0.0, nodes, w, offset)
    kappa = M.toeplitz_kappa("laplace", 0.0, p, w, offset)
    assert np.allclose(M.toeplitz_dense(kappa, p), dense, rtol=1e-13, atol=0)
    D = np.fft.fftn(kappa)
    x = np.random.default_rng(0).standard_normal(p ** 3)
    y = M.modulate_out(M.hadamard_transform(D, "N") * M.modulate_in(x, p, 2 * p - 1, "N"), p, "N")
    assert np.allclose(y, dense @ x, rtol=1e-12, atol=1e-14)


def test_toeplitz_orientations_match_dense_transpose_and_adjoint():
    """test/H2Toeplitz/test_smoke.jl:27-78: transpose swaps the FFT plans, adjoint conjugates D;
    plus adjoint duality <y, T x> = <T^H y, x>."""
    rng = np.random.default_rng(1)
    p, L = 3, 5
    kappa = rng.standard_normal((L, L, L)) + 1j * rng.standard_normal((L, L, L))
    D = np.fft.fftn(kappa)
    Td = M.toeplitz_dense(kappa, p)
    x = rng.standard_normal(p ** 3) + 1j * rng.standard_normal(p ** 3)
    yv = rng.standard_normal(p ** 3) + 1j * rng.standard_normal(p ** 3)
    app = {op: M.modulate_out(M.hadamard_transform(D, op) * M.modulate_in(x, p, L, op), p, op) for op in "NTC"}
    assert np.allclose(app["N"], Td @ x, rtol=1e-12)
    assert np.allclose(app["T"], Td.T @ x, rtol=1e-12)
    assert np.allclose(app["C"], Td.conj().T @ x, rtol=1e-12)
    assert np.isclose(np.vdot(yv, Td @ x), np.vdot(Td.conj().T @ yv, x))


def test_symmetry_group_counts():
    """test/farinteractions/test_toeplitz3d.jl:107-116: 48 accepted / 48 distinct / 20 self-inverse
    on a (2n-1)^3 tensor; 16 accepted / 2 distinct / 2 self-inverse on (2n-1, 1, 1)."""
    L = 5
    g = M.symmetry_group()
    assert len(g) == 48
    perms = {tuple(M.toeplitz_permutation((L, L, L), s)) for s in g}
    selfinv = sum(1 for q in perms if np.array_equal(np.array(q)[np.array(q)], np.arange(L ** 3)))
    assert (len(perms), selfinv, len(perms) - selfinv) == (48, 20, 28)
    acc = []
    for s in g:
        try:
            acc.append(tuple(M.toeplitz_permutation((L, 1, 1), s)))
        except ValueError:
            pass
    dist = set(acc)
    assert (len(acc), len(dist), sum(1 for q in dist if np.array_equal(np.array(q)[np.array(q)], np.arange(L)))) == (16, 2, 2)


def test_canonical_orbits_of_far_offsets():
    """316 far offsets per level fall into 16 orbits: the 19.75x ceiling of
    test/farinteractions/test_toeplitz3d.jl:128-132."""
    offs = [o for o in np.ndindex(7, 7, 7) if max(abs(v - 3) for v in o) > 1]
    offs = [tuple(v - 3 for v in o) for o in offs]
    assert len(offs) == 316
    canon = {M.canonical_offset(o)[0] for o in offs}
    assert len(canon) == 16 and len(offs) / len(canon) == 19.75


@pytest.mark.parametrize("offset", [(3, -2, 1), (-2, 0, 2), (2, 3, -3)])
def test_permuted_spectra_and_dense_equal_direct(offset):
    """Symmetry-reduced storage adds no approximation for an invariant kernel: the permuted
    canonical spectrum / matrix equals the one built for the offset directly, and a chiral
    element applied with the wrong inverse convention does not (negative control)."""
    p, w = 3, 0.25
    L = 2 * p - 1
    nodes = M.nodes1d(p, "equi")
    canon, S = M.canonical_offset(offset)
    Dc = np.fft.fftn(M.toeplitz_kappa("laplace", 0.0, p, w, canon)).reshape(-1)
    Dd = np.fft.fftn(M.toeplitz_kappa("laplace", 0.0, p, w, offset)).reshape(-1)
    perm = M.toeplitz_permutation((L, L, L), S)
    assert np.allclose(Dc[perm], Dd, rtol=1e-12, atol=1e-12)
    Mc = M.m2l_dense("laplace", 0.0, nodes, w, canon)
    Md = M.m2l_dense("laplace", 0.0, nodes, w, offset)
    pi = M.lattice_grid_permutation((p, p, p), S)
    t = np.empty_like(Mc)
    t[np.ix_(pi, pi)] = Mc
    assert np.allclose(t, Md, rtol=1e-13)
    # P M Q^T form (test/i2otranslationcollection/test_reduceddense.jl:48-140)
    P = np.zeros((p ** 3, p ** 3))
    P[pi, np.arange(p ** 3)] = 1
    assert np.allclose(P @ Mc @ P.T, Md, rtol=1e-13)
    # negative control: the forward element in place of the inverse on a chiral symmetry
    if not np.array_equal(perm[perm], np.arange(L ** 3)):
        wrong = M.toeplitz_permutation((L, L, L), M.inverse_symmetry(S))
        assert not np.allclose(Dc[wrong], Dd, rtol=1e-6)


def test_helmholtz_kernel_is_the_spherical_hankel_function():
    """exp(ikr)/(4 pi r) = (ik/4 pi) h0(kr) with h0 = j0 + i y0 (scipy.special), 0 at r = 0; a sign
    slip (exp(-ikr)) or a missing 1/(4 pi) fails here."""
    from scipy.special import spherical_jn, spherical_yn

    r = np.random.default_rng(0).uniform(1e-3, 3.0, 200)
    for k in (2.0, 10.0, 37.5):
        h0 = spherical_jn(0, k * r) + 1j * spherical_yn(0, k * r)
        assert np.allclose(M.kernel_of_r("helmholtz", k, r), 1j * k / (4 * np.pi) * h0, rtol=1e-13, atol=0)
    xt = np.random.default_rng(1).random((30, 3))
    G = M.kernel_matrix("helmholtz", 10.0, xt, xt)
    d = np.linalg.norm(xt[:, None] - xt[None], axis=-1)
    off = ~np.eye(30, dtype=bool)
    assert np.allclose(G[off], M.kernel_of_r("helmholtz", 10.0, d[off]), rtol=1e-14)
    assert np.all(np.diag(G) == 0)


def test_helmholtz_tends_to_laplace():
    r = np.linspace(0.01, 2.0, 50)
    lap = M.kernel_of_r("laplace", 0.0, r)
    assert np.allclose(M.kernel_of_r("helmholtz", 1e-9, r), lap, rtol=1e-8, atol=0)
    assert np.array_equal(M.kernel_of_r("helmholtz", 0.0, r).real, lap)


@pytest.mark.parametrize("offset", [(2, 0, 0), (3, -2, 1)])
def test_helmholtz_toeplitz_spectrum_equals_dense_kernel(offset):
    """The Helmholtz branch of the equispaced M2L: dense == multilevel Toeplitz == FFT application;
    the transpose is the same (symmetric kernel), the adjoint conjugates."""
    p, w, k = 4, 0.125, 10.0
    nodes = M.nodes1d(p, "equi")
    dense = M.m2l_dense("helmholtz", k, nodes, w, offset)
    kappa = M.toeplitz_kappa("helmholtz", k, p, w, offset)
    assert np.allclose(M.toeplitz_dense(kappa, p), dense, rtol=1e-13, atol=0)
    D = np.fft.fftn(kappa)
    x = np.random.default_rng(0).standard_normal(p ** 3) + 1j * np.random.default_rng(1).standard_normal(p ** 3)
    for op, R in (("N", dense), ("T", dense.T), ("C", dense.conj().T)):
        y = M.modulate_out(M.hadamard_transform(D, op) * M.modulate_in(x, p, 2 * p - 1, op), p, op)
        assert np.allclose(y, R @ x, rtol=1e-12, atol=1e-14), op
