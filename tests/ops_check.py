"""Checks of the device-built operators (h2g.operators.device_operators) against the oracle's
operator builders: dense couplings, Toeplitz spectra (numpy's FFT of the oracle's kernel tensor),
sector transfers and symmetry permutations.  Shared by the CPU replay and the GPU tests."""

import numpy as np

from oracle import method as OM

KEYS = np.array([[2, 2, 0, 1], [3, -3, 2, -2], [4, 0, 0, 3], [5, 2, -3, 0], [6, -2, -2, -2], [6, 3, 3, 3]])


def check(p, nodekind, kernel, wavenumber, m2l_kind, root=1.0):
    """Largest relative deviation of the device operators from the oracle's (asserts 1e-13)."""
    from h2g.operators import device_operators

    nodes = OM.nodes1d(p, nodekind)
    ops, tr, perms = device_operators(kernel, wavenumber, p, nodes, root, m2l_kind, KEYS)
    worst = 0.0
    for i, (lv, *off) in enumerate(KEYS):
        w = root / 2.0 ** lv
        if "toeplitz" in m2l_kind:
            ref = np.fft.fftn(OM.toeplitz_kappa(kernel, wavenumber, p, w, tuple(off)))
        else:
            ref = OM.m2l_dense(kernel, wavenumber, nodes, w, tuple(off)).T  # stored column-major
        e = np.max(np.abs(ops[i] - ref)) / np.max(np.abs(ref))
        worst = max(worst, e)
        assert e < 1e-13, (i, e)
    for s in range(8):
        ref = OM.transfer(nodes, s).T
        e = np.max(np.abs(tr[s] - ref))
        worst = max(worst, e)
        assert e < 1e-13, (s, e)
    if m2l_kind == "symdense":
        g = OM.symmetry_group()
        for i, S in enumerate(g):
            assert np.array_equal(perms[i], OM.lattice_grid_permutation((p, p, p), S)), i
    elif m2l_kind == "symtoeplitz":
        L = 2 * p - 1
        g = OM.symmetry_group()
        for i, S in enumerate(g):
            assert np.array_equal(perms[i], OM.toeplitz_permutation((L, L, L), S)), i
    return worst


CASES = [(4, "cheb", "laplace", 0.0, "dense"), (3, "cheb", "helmholtz", 10.0, "symdense"),
         (4, "equi", "laplace", 0.0, "toeplitz"), (3, "equi", "helmholtz", 2.0, "symtoeplitz")]
