"""Core assembly for the oracle -- TEST INFRASTRUCTURE ONLY.

``build_core`` assembles the interpolation method (collections, transfers, M2L storage of
each kind, near blocks) over the octrees of ``tree.py``.  ``random_core`` fills the same
structure with random complex data, which makes every orientation rule observable (no
symmetry of the kernel can hide a transposition).  ``worked_example_core`` is the
reference's worked example (``docs/src/implementing/example.md:12-256``): 32 points on a
line, the exact two-term kernel ``1 + 2s + 3t + (4+i)st``, identity transfers and one
coupling ``[1 3; 2 4+i]``, for which the H2 map equals the dense kernel matrix.
"""

from __future__ import annotations

import numpy as np

from . import method as M
from .h2 import Core
from .tree import build_tree, m2l_pairs, near_pairs


def _level_transfers(tree, nodes):
    K = len(nodes) ** 3
    per_sector = [M.transfer(nodes, s) for s in range(8)]
    return [per_sector if l > 0 else [np.eye(K)] * 8 for l in range(tree.n_levels)]


def _collection(tree, nodes):
    K = len(nodes) ** 3
    n = tree.points.shape[0]
    C = np.zeros((K, n))
    pts = tree.points[tree.perm]
    for b in tree.leaves():
        w = tree.width(tree.level[b])
        C[:, tree.begin[b]:tree.end[b]] = M.leaf_collection(nodes, tree.center(b), w / 2, pts[tree.begin[b]:tree.end[b]])
    return C


def build_core(points_test, points_trial=None, p=4, nodekind="cheb", kernel="laplace",
               wavenumber=0.0, m2l_kind="dense", leaf_size=64, alpha=1.0, origin=(0.0, 0.0, 0.0), size=1.0):
    test = build_tree(points_test, leaf_size, origin, size)
    trial = test if points_trial is None else build_tree(points_trial, leaf_size, origin, size)
    nodes = M.nodes1d(p, nodekind)
    K = p ** 3
    pairs = m2l_pairs(test, trial)
    offs = trial.coord[pairs[:, 1]] - test.coord[pairs[:, 0]] if len(pairs) else np.zeros((0, 3), np.int64)
    lv = test.level[pairs[:, 0]] if len(pairs) else np.zeros(0, np.int64)
    dtype = np.complex128 if kernel == "helmholtz" else np.float64

    pair_op = np.zeros(len(pairs), dtype=np.int64)
    pair_sym = None
    perms = None
    ops = []
    keys = {}
    group = M.symmetry_group()
    sym_index = {g: i for i, g in enumerate(group)}
    if m2l_kind.startswith("sym"):
        pair_sym = np.zeros(len(pairs), dtype=np.int64)
    for k in range(len(pairs)):
        o = tuple(int(v) for v in offs[k])
        if m2l_kind.startswith("sym"):
            c, S = M.canonical_offset(o)
            key = (int(lv[k]), c)
            pair_sym[k] = sym_index[S]
        else:
            key = (int(lv[k]), o)
        if key not in keys:
            keys[key] = len(ops)
            w = test.width(key[0])
            if m2l_kind in ("dense", "symdense"):
                ops.append(M.m2l_dense(kernel, wavenumber, nodes, w, key[1]))
            else:
                ops.append(np.fft.fftn(M.toeplitz_kappa(kernel, wavenumber, p, w, key[1])))
        pair_op[k] = keys[key]
    if m2l_kind == "symdense":
        perms = np.stack([M.lattice_grid_permutation((p, p, p), g) for g in group])
    elif m2l_kind == "symtoeplitz":
        L = 2 * p - 1
        perms = np.stack([M.toeplitz_permutation((L, L, L), g) for g in group])
    if ops:
        ops = np.stack(ops)
    else:
        shape = (0, K, K) if m2l_kind in ("dense", "symdense") else (0, 2 * p - 1, 2 * p - 1, 2 * p - 1)
        ops = np.zeros(shape, dtype=np.complex128 if "toeplitz" in m2l_kind else dtype)

    npairs = near_pairs(test, trial)
    tp = test.points[test.perm]
    cp = trial.points[trial.perm]
    vals = [M.kernel_matrix(kernel, wavenumber, tp[test.begin[b]:test.end[b]], cp[trial.begin[c]:trial.end[c]])
            for b, c in npairs]
    return Core(
        test=test, trial=trial, K=K, p=p, dtype=dtype,
        coll_test=_collection(test, nodes),
        coll_trial=_collection(trial, nodes) if trial is not test else None,
        transfers_test=_level_transfers(test, nodes),
        transfers_trial=_level_transfers(trial, nodes) if trial is not test else None,
        pairs=pairs, pair_op=pair_op, m2l_kind=m2l_kind, ops=ops, pair_sym=pair_sym,
        perms=perms, near_pairs=npairs, near_values=vals, alpha=alpha,
    )._galerkin_fill()


def _galerkin_fill(self):
    if self.coll_trial is None:
        self.coll_trial = self.coll_test
    if self.transfers_trial is None:
        self.transfers_trial = self.transfers_test
    return self


Core._galerkin_fill = _galerkin_fill


def random_core(points_test, points_trial=None, p=2, m2l_kind="dense", leaf_size=8,
                seed=0, complex_data=True, alpha=1.0):
    """Same tree/plan structure with random (complex) operators of every kind, including the
    diagonal (elementwise, MLFMA-style) couplings: one random K-vector per translation id."""
    rng = np.random.default_rng(seed)
    test = build_tree(points_test, leaf_size)
    trial = test if points_trial is None else build_tree(points_trial, leaf_size)
    K = p ** 3
    L = 2 * p - 1

    def rnd(*shape):
        a = rng.standard_normal(shape)
        if complex_data:
            a = a + 1j * rng.standard_normal(shape)
        return a

    def transfers(tree):
        return [[rnd(K, K) / K for _ in range(8)] for _ in range(tree.n_levels)]

    def coll(tree):
        return rnd(K, tree.points.shape[0])

    pairs = m2l_pairs(test, trial)
    offs = trial.coord[pairs[:, 1]] - test.coord[pairs[:, 0]] if len(pairs) else np.zeros((0, 3), np.int64)
    lv = test.level[pairs[:, 0]] if len(pairs) else np.zeros(0, np.int64)
    group = M.symmetry_group()
    sym_index = {g: i for i, g in enumerate(group)}
    keys, ops = {}, []
    pair_op = np.zeros(len(pairs), dtype=np.int64)
    pair_sym = np.zeros(len(pairs), dtype=np.int64) if m2l_kind.startswith("sym") else None
    for k in range(len(pairs)):
        o = tuple(int(v) for v in offs[k])
        if pair_sym is not None:
            c, S = M.canonical_offset(o)
            key = (int(lv[k]), c)
            pair_sym[k] = sym_index[S]
        else:
            key = (int(lv[k]), o)
        if key not in keys:
            keys[key] = len(ops)
            if m2l_kind in ("dense", "symdense"):
                ops.append(rnd(K, K) / K)
            elif m2l_kind == "diagonal":
                ops.append(rnd(K))
            else:
                ops.append(np.fft.fftn(rnd(L, L, L) / K))
        pair_op[k] = keys[key]
    perms = None
    if m2l_kind == "symdense":
        perms = np.stack([M.lattice_grid_permutation((p, p, p), g) for g in group])
    elif m2l_kind == "symtoeplitz":
        perms = np.stack([M.toeplitz_permutation((L, L, L), g) for g in group])
    dtype = np.complex128 if complex_data else np.float64
    if ops:
        ops = np.stack(ops)
    else:
        shape = (0, K, K) if m2l_kind in ("dense", "symdense") else ((0, K) if m2l_kind == "diagonal" else (0, L, L, L))
        ops = np.zeros(shape, dtype=np.complex128 if (complex_data or "toeplitz" in m2l_kind) else np.float64)
    npairs = near_pairs(test, trial)
    vals = [rnd(test.end[b] - test.begin[b], trial.end[c] - trial.begin[c]) for b, c in npairs]
    ct = coll(test)
    tt = transfers(test)
    return Core(
        test=test, trial=trial, K=K, p=p, dtype=dtype,
        coll_test=ct, coll_trial=ct if trial is test else coll(trial),
        transfers_test=tt, transfers_trial=tt if trial is test else transfers(trial),
        pairs=pairs, pair_op=pair_op, m2l_kind=m2l_kind, ops=ops, pair_sym=pair_sym,
        perms=perms, near_pairs=npairs, near_values=vals, alpha=alpha,
    )


def worked_example_points(n=32):
    """32 points on a line inside the unit cube (the example's ``range(0, 1; length=32)``
    mapped to x in [0, 1) so that the octree splits along x)."""
    s = np.linspace(0.0, 1.0, n)
    pts = np.zeros((n, 3))
    pts[:, 0] = s * (1 - 1e-9)
    pts[:, 1] = 0.3
    pts[:, 2] = 0.6
    return s, pts


def worked_example_core(leaf_size=2):
    """The reference's worked example on the build's own octree: moments ``(sum x, sum s x)``,
    identity transfers, coupling ``C = [1 3; 2 4+i]`` for every far pair, near blocks from
    the kernel; ``Matrix(A) == [K(s_i, s_j)]`` exactly."""
    s, pts = worked_example_points()
    tree = build_tree(pts, leaf_size)
    kern = lambda a, b: 1 + 2 * a + 3 * b + (4 + 1j) * a * b
    sp = s[tree.perm]
    coll = np.vstack([np.ones(len(s)), sp]).astype(np.complex128)
    I2 = np.eye(2, dtype=np.complex128)
    transfers = [[I2] * 8 for _ in range(tree.n_levels)]
    pairs = m2l_pairs(tree, tree)
    C = np.array([[1, 3], [2, 4 + 1j]], dtype=np.complex128)
    npairs = near_pairs(tree, tree)
    vals = [kern(sp[tree.begin[b]:tree.end[b], None], sp[None, tree.begin[c]:tree.end[c]]) for b, c in npairs]
    core = Core(
        test=tree, trial=tree, K=2, p=0, dtype=np.complex128,
        coll_test=coll, coll_trial=coll, transfers_test=transfers, transfers_trial=transfers,
        pairs=pairs, pair_op=np.zeros(len(pairs), dtype=np.int64), m2l_kind="dense",
        ops=C[None], near_pairs=npairs, near_values=vals, alpha=1.0 + 0j,
    )
    F = kern(s[:, None], s[None, :])
    return core, s, F


def farmul_fixture_core():
    """The reference's hand-derived far-field fixture (``test/farinteractions/test_farmul.jl:11-97``):
    two clusters of four points on a line (x = 0, 0.1, 5, 5.1 and 40, 40.1, 45, 45.1, scaled here
    by 1/64 into the unit cube), each split into two sub-clusters.  Moments are the plain charge
    sum (K = 1, unit collections), transfers are the identity and every coupling is the scalar
    ``c = 2 - i``, so the far matrix is ``F[left, right] = F[right, left] = c`` and 0 elsewhere.
    On this build's octree (leaves of two points) the two clusters translate at level 2 and the
    sub-clusters of one cluster are adjacent leaves, as in the reference's tree; near values are
    zero (the fixture checks the far field only)."""
    s = np.array([0.0, 0.1, 5.0, 5.1, 40.0, 40.1, 45.0, 45.1]) / 64.0
    pts = np.zeros((8, 3))
    pts[:, 0] = s
    pts[:, 1] = 0.3
    pts[:, 2] = 0.6
    tree = build_tree(pts, 2)
    c = 2.0 - 1.0j
    one = np.ones((1, 8), dtype=np.complex128)
    I1 = np.eye(1, dtype=np.complex128)
    transfers = [[I1] * 8 for _ in range(tree.n_levels)]
    pairs = m2l_pairs(tree, tree)
    npairs = near_pairs(tree, tree)
    vals = [np.zeros((tree.end[b] - tree.begin[b], tree.end[q] - tree.begin[q]), dtype=np.complex128)
            for b, q in npairs]
    core = Core(
        test=tree, trial=tree, K=1, p=0, dtype=np.complex128,
        coll_test=one, coll_trial=one, transfers_test=transfers, transfers_trial=transfers,
        pairs=pairs, pair_op=np.zeros(len(pairs), dtype=np.int64), m2l_kind="dense",
        ops=np.array([[[c]]]), near_pairs=npairs, near_values=vals, alpha=1.0 + 0j,
    )
    F = np.zeros((8, 8), dtype=np.complex128)
    F[:4, 4:] = c
    F[4:, :4] = c
    return core, pts, F


def no_far_core():
    """Five points inside adjacent leaves with no far pair at all, random complex operators: every
    receiving leaf is uninitialised, so the 5-argument farmul! writes no row
    (``src/farinteractions/traversalhelpers.jl:103-112``)."""
    pts = np.array([[0.49, 0.5, 0.5], [0.48, 0.5, 0.5], [0.51, 0.5, 0.5], [0.52, 0.5, 0.5], [0.5, 0.51, 0.5]])
    return random_core(pts, p=2, m2l_kind="dense", leaf_size=2, seed=3)
