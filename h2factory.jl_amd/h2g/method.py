"""The point-kernel interpolation method that supplies a GPU core with its data.

H2Factory executes a fast method but defines none: the downstream method package supplies the
tree, leaf collections, level transfers, couplings and near operator
(``docs/src/index.md:3-5``; ``src/Interfaces.jl``).  This module is that method for the
benchmark workloads: tensor-product Lagrange interpolation of order ``p`` on Chebyshev or
equispaced nodes, Laplace ``1/(4 pi r)`` or Helmholtz ``exp(i k r)/(4 pi r)`` kernels
(0 on the diagonal), couplings stored dense, as Toeplitz spectra (equispaced nodes only),
or symmetry-reduced through the 48-element lattice group.

Layout conventions (shared with the C ABI): a moment is ``p x p x p`` with x fastest
(``k = kx + p*ky + p*p*kz``); K x K matrices are column-major; a Toeplitz difference tensor
has ``L = 2p-1`` entries per axis in wrap order ``0..p-1, -(p-1)..-1`` and
``T[i, j] = kappa[(i-j) mod L]`` per axis (``test/H2Toeplitz/test_smoke.jl:27-36``);
spectra are ``D = fftn(kappa)``, x fastest; a lattice symmetry ``(perm, signs)`` acts as
``(S v)[i] = signs[i] * v[perm[i]]``.
"""

from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np

from . import tree as T

FOUR_PI = 4.0 * np.pi
KINDS = {"dense": 0, "toeplitz": 1, "symdense": 2, "symtoeplitz": 3, "diagonal": 4}


def nodes1d(p: int, kind: str) -> np.ndarray:
    if kind == "cheb":
        return -np.cos((2 * np.arange(p) + 1) * np.pi / (2 * p))
    if kind == "equi":
        return np.linspace(-1.0, 1.0, p) if p > 1 else np.zeros(1)
    raise ValueError(kind)


def lagrange(nodes: np.ndarray, t: np.ndarray) -> np.ndarray:
    """``out[i, k] = l_k(t_i)``."""
    t = np.asarray(t, dtype=np.float64).reshape(-1)
    diff = t[:, None] - nodes[None, :]
    p = len(nodes)
    out = np.empty((t.size, p))
    for k in range(p):
        m = np.arange(p) != k
        out[:, k] = np.prod(diff[:, m] / (nodes[k] - nodes[m])[None, :], axis=1)
    return out


def transfer_matrix(nodes: np.ndarray, sector: int) -> np.ndarray:
    """Parent basis at the child's nodes: ``T[k, j]`` (K x K), child of ``sector``."""
    f = [lagrange(nodes, (((sector >> a) & 1) * 2 - 1 + nodes) / 2.0).T for a in range(3)]
    return np.kron(f[2], np.kron(f[1], f[0]))


def grid(nodes: np.ndarray) -> np.ndarray:
    p = len(nodes)
    kz, ky, kx = np.meshgrid(np.arange(p), np.arange(p), np.arange(p), indexing="ij")
    return np.stack([nodes[kx.ravel()], nodes[ky.ravel()], nodes[kz.ravel()]], axis=1)


def kernel_r(kind: str, wavenumber: float, r: np.ndarray):
    rr = np.where(r > 0, r, 1.0)
    if kind == "laplace":
        return np.where(r > 0, 1.0 / (FOUR_PI * rr), 0.0)
    if kind == "helmholtz":
        return np.where(r > 0, np.exp(1j * wavenumber * rr) / (FOUR_PI * rr), 0.0)
    raise ValueError(kind)


def dense_couplings(kind, wavenumber, nodes, width, offsets) -> np.ndarray:
    """``M[o][k, j] = G(x_k - y_j)``, receiving box at 0, translating box at ``offset*width``."""
    h = width / 2.0
    g = grid(nodes) * h
    d = g[None, :, None, :] - g[None, None, :, :] - width * np.asarray(offsets, dtype=np.float64)[:, None, None, :]
    return kernel_r(kind, wavenumber, np.sqrt(np.sum(d * d, axis=-1)))


def toeplitz_spectra(kind, wavenumber, p, width, offsets) -> np.ndarray:
    """``D[o] = fftn(kappa_o)`` with ``kappa_o[d] = G(delta*(d - (p-1)*o))``, indexed [z, y, x]."""
    L = 2 * p - 1
    delta = width / (p - 1)
    w = np.arange(L)
    d = np.where(w <= L // 2, w, w - L).astype(np.float64)
    o = np.asarray(offsets, dtype=np.float64)
    vx = delta * (d[None, None, None, :] - (p - 1) * o[:, 0, None, None, None])
    vy = delta * (d[None, None, :, None] - (p - 1) * o[:, 1, None, None, None])
    vz = delta * (d[None, :, None, None] - (p - 1) * o[:, 2, None, None, None])
    kappa = kernel_r(kind, wavenumber, np.sqrt(vx * vx + vy * vy + vz * vz))
    return np.fft.fftn(kappa, axes=(1, 2, 3))


# ------------------------------------------------------------------------------------------
# lattice symmetries
# ------------------------------------------------------------------------------------------
GROUP = [(perm, signs) for perm in itertools.permutations(range(3)) for signs in itertools.product((1, -1), repeat=3)]
GROUP_INDEX = {g: i for i, g in enumerate(GROUP)}


def _inverse(sym):
    perm, signs = sym
    inv = [0, 0, 0]
    for i in range(3):
        inv[perm[i]] = i
    return tuple(inv), tuple(signs[inv[j]] for j in range(3))


def canonicalize(offsets: np.ndarray):
    """Canonical offsets (sorted |o|) and the index of ``S`` in ``GROUP`` with ``S*c == o``."""
    o = np.asarray(offsets, dtype=np.int64)
    a = np.abs(o)
    order = np.argsort(a, axis=1, kind="stable")
    canon = np.take_along_axis(a, order, axis=1)
    perm = np.argsort(order, axis=1, kind="stable")
    signs = np.where(o < 0, -1, 1)
    sym = np.array([GROUP_INDEX[(tuple(int(v) for v in perm[i]), tuple(int(v) for v in signs[i]))]
                    for i in range(o.shape[0])], dtype=np.int64)
    return canon, sym


def toeplitz_perm(L: int, sym) -> np.ndarray:
    """``perm[j]`` = wrap index of ``S^-1 d_j`` (``kappa_{S o}[j] == kappa_o[perm[j]]``)."""
    inv = _inverse(sym)
    w = np.arange(L)
    d = np.where(w <= L // 2, w, w - L)
    dz, dy, dx = np.meshgrid(d, d, d, indexing="ij")
    v = np.stack([dx.ravel(), dy.ravel(), dz.ravel()], axis=1)
    img = np.stack([inv[1][i] * v[:, inv[0][i]] for i in range(3)], axis=1) % L
    return (img[:, 0] + L * (img[:, 1] + L * img[:, 2])).astype(np.int32)


def grid_perm(p: int, sym) -> np.ndarray:
    """``perm[i]`` = index of the image of grid point ``i`` (reversal ``u -> p-1-u`` per sign)."""
    perm, signs = sym
    g = np.arange(p)
    uz, uy, ux = np.meshgrid(g, g, g, indexing="ij")
    u = np.stack([ux.ravel(), uy.ravel(), uz.ravel()], axis=1)
    img = np.stack([u[:, perm[i]] if signs[i] > 0 else p - 1 - u[:, perm[i]] for i in range(3)], axis=1)
    return (img[:, 0] + p * (img[:, 1] + p * img[:, 2])).astype(np.int32)


# ------------------------------------------------------------------------------------------
# map description
# ------------------------------------------------------------------------------------------
@dataclass
class SideArrays:
    tree: T.Octree
    collection: np.ndarray = None      # (n_dofs, K) C-order == K x n col-major; None: device fill
    transfers: np.ndarray = None       # (n_levels, 8, K, K) holding T^T (col-major T)


@dataclass
class MapArrays:
    """Everything ``h2g_core_create`` takes, as numpy arrays (host-side core description)."""
    dtype: str
    K: int
    p: int
    test: SideArrays
    trial: SideArrays
    galerkin: bool
    m2l_kind: str
    pairs: np.ndarray
    pair_op: np.ndarray
    pair_sym: np.ndarray
    ops: np.ndarray
    perms: np.ndarray
    near_pairs: np.ndarray
    near_offset: np.ndarray
    n_near_values: int
    near_values: np.ndarray = None     # None: device fill
    ops_keys: np.ndarray = None        # (n_ops, 4): level and (canonical) offset of each translation id
    alpha: complex = 1.0
    # method data for device assembly
    kernel: str = "laplace"
    wavenumber: float = 0.0
    nodes: np.ndarray = None
    stats: dict = field(default_factory=dict)


def _side_collection(tree: T.Octree, nodes, K):
    pts = tree.points[tree.perm]
    C = np.zeros((tree.points.shape[0], K))
    cen = tree.centers()
    hw = tree.halfwidths()
    p = len(nodes)
    for b in tree.leaves():
        s, e = tree.begin[b], tree.end[b]
        t = (pts[s:e] - cen[b][None, :]) / hw[b]
        lx, ly, lz = (lagrange(nodes, t[:, a]) for a in range(3))
        C[s:e] = (lz[:, :, None, None] * ly[:, None, :, None] * lx[:, None, None, :]).reshape(e - s, p ** 3)
    return C


def _transfers(tree: T.Octree, nodes, per=None):
    """Per level and sector T^T (column-major T); ``per``: the 8 sector matrices already built."""
    K = len(nodes) ** 3
    if per is None:
        per = np.stack([transfer_matrix(nodes, s).T for s in range(8)])
    out = np.broadcast_to(per, (tree.n_levels, 8, K, K)).copy()
    out[0] = np.eye(K)
    return out


def translation_ids(test, trial, pairs, symmetric):
    """Translation id of each M2L pair: the rank of its (level, offset) -- (level, canonical
    offset) and a symmetry id when ``symmetric`` -- among the present ones.  Returns
    (pair_op, pair_sym or None, the (level, offset) rows of the ids)."""
    offs = trial.coord[pairs[:, 1]] - test.coord[pairs[:, 0]]
    lv = test.level[pairs[:, 0]]
    sym = None
    if symmetric:
        canon, sym = canonicalize(offs)
        keyarr = np.concatenate([lv[:, None], canon], axis=1)
    else:
        keyarr = np.concatenate([lv[:, None], offs], axis=1)
    # one integer per (level, offset): offsets lie in [-3, 3]^3 for the standard lists; the
    # generic path keeps np.unique over rows
    if len(keyarr) and np.abs(keyarr[:, 1:]).max() <= 3:
        code = ((keyarr[:, 0] * 8 + keyarr[:, 1] + 3) * 8 + keyarr[:, 2] + 3) * 8 + keyarr[:, 3] + 3
        ucode, first, pair_op = np.unique(code, return_index=True, return_inverse=True)
        uniq = keyarr[first]
    else:
        uniq, pair_op = np.unique(keyarr, axis=0, return_inverse=True)
    return pair_op.reshape(-1), sym, uniq


def build_map(points_test, points_trial=None, p=6, nodekind="equi", kernel="laplace", wavenumber=0.0,
              m2l_kind="toeplitz", leaf_size=64, alpha=1.0, host_values=False, origin=(0.0, 0.0, 0.0),
              size=1.0, plan="host", ops="host") -> MapArrays:
    """Tree, plans and operators for the interpolation method.

    ``host_values`` computes the near values and the leaf collections on the host (small
    problems, exact comparisons); otherwise they are filled on the device by
    ``h2g_method_fill`` after the core is created.  ``plan="device"`` builds the trees, the
    M2L pairs with their translation ids and the near pairs on the GPU (``h2g_plan_build``; the
    same arrays as the host builder, bit for bit).  ``ops="device"`` builds the couplings, the
    level transfers and the symmetry permutations on the GPU (``h2g_method_operators``).
    """
    if m2l_kind in ("toeplitz", "symtoeplitz") and nodekind != "equi":
        raise ValueError("Toeplitz couplings need equispaced nodes")
    if plan not in ("host", "device"):
        raise ValueError(f"plan must be 'host' or 'device', not {plan!r}")
    if ops not in ("host", "device"):
        raise ValueError(f"ops must be 'host' or 'device', not {ops!r}")
    nodes = nodes1d(p, nodekind)
    K = p ** 3
    dplan = None
    if plan == "device":
        from .plan import device_plan

        dplan = device_plan(points_test, points_trial, leaf_size, origin, size, symmetric=m2l_kind.startswith("sym"))
        test, trial = dplan.test, dplan.trial
        pairs, pair_op, sym, uniq = dplan.pairs, dplan.pair_op, dplan.pair_sym, dplan.keys
    else:
        test = T.build_octree(points_test, leaf_size, origin, size)
        trial = test if points_trial is None else T.build_octree(points_trial, leaf_size, origin, size)
        pairs = T.m2l_pairs(test, trial)
        pair_op, sym, uniq = translation_ids(test, trial, pairs, m2l_kind.startswith("sym"))
    dev_tr = None
    if ops == "device":
        from .operators import device_operators

        dev_ops, dev_tr, dev_perms = device_operators(kernel, wavenumber, p, nodes, size, m2l_kind, uniq)
    ops = []
    for l in (np.unique(uniq[:, 0]) if len(uniq) else []) if dev_tr is None else []:
        sel = uniq[:, 0] == l
        width = test.size / 2.0 ** l
        if m2l_kind in ("dense", "symdense"):
            m = dense_couplings(kernel, wavenumber, nodes, width, uniq[sel, 1:])
            ops.append(np.ascontiguousarray(np.swapaxes(m, 1, 2)))  # column-major K x K
        else:
            ops.append(toeplitz_spectra(kernel, wavenumber, p, width, uniq[sel, 1:]))
    is_complex = kernel == "helmholtz"
    if dev_tr is not None:
        ops = dev_ops
    elif ops:
        ops = np.concatenate(ops)
    else:
        ops = np.zeros((0, K, K) if m2l_kind in ("dense", "symdense") else (0, 2 * p - 1, 2 * p - 1, 2 * p - 1))
    if m2l_kind in ("toeplitz", "symtoeplitz"):
        ops = np.ascontiguousarray(ops.astype(np.complex128))
    else:
        ops = np.ascontiguousarray(ops.astype(np.complex128 if is_complex else np.float64))
    perms = None
    if dev_tr is not None:
        perms = dev_perms
    elif m2l_kind == "symdense":
        perms = np.stack([grid_perm(p, g) for g in GROUP])
    elif m2l_kind == "symtoeplitz":
        perms = np.stack([toeplitz_perm(2 * p - 1, g) for g in GROUP])

    npairs = dplan.near_pairs if dplan is not None else T.near_pairs(test, trial)
    nr = test.end[npairs[:, 0]] - test.begin[npairs[:, 0]]
    nc = trial.end[npairs[:, 1]] - trial.begin[npairs[:, 1]]
    sizes = nr * nc
    if dplan is not None:
        near_offset, n_near_values = dplan.near_offset, dplan.n_near_values
    else:
        near_offset = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64) if len(sizes) else np.zeros(0, np.int64)
        n_near_values = int(sizes.sum())
    near_values = None
    dt = np.complex128 if is_complex else np.float64
    ts = SideArrays(test, transfers=_transfers(test, nodes, dev_tr))
    cs = ts if trial is test else SideArrays(trial, transfers=_transfers(trial, nodes, dev_tr))
    if host_values:
        near_values = np.zeros(n_near_values, dtype=dt)
        tp = test.points[test.perm]
        cp = trial.points[trial.perm]
        for q, (b, c) in enumerate(npairs):
            x = tp[test.begin[b]:test.end[b]]
            y = cp[trial.begin[c]:trial.end[c]]
            d = x[:, None, :] - y[None, :, :]
            blk = kernel_r(kernel, wavenumber, np.sqrt(np.sum(d * d, axis=-1)))
            near_values[near_offset[q]:near_offset[q] + sizes[q]] = blk.T.ravel()  # column-major
        ts.collection = _side_collection(test, nodes, K).astype(dt)
        if cs is not ts:
            cs.collection = _side_collection(trial, nodes, K).astype(dt)
    if is_complex:
        ts.transfers = ts.transfers.astype(dt)
        if cs is not ts:
            cs.transfers = cs.transfers.astype(dt)
    return MapArrays(dtype="c64" if is_complex else "f64", K=K, p=p, test=ts, trial=cs, galerkin=trial is test,
                     m2l_kind=m2l_kind, pairs=pairs, pair_op=pair_op, pair_sym=sym, ops=ops, perms=perms,
                     near_pairs=npairs, near_offset=near_offset, n_near_values=n_near_values,
                     near_values=near_values, alpha=alpha, kernel=kernel, wavenumber=wavenumber, nodes=nodes,
                     ops_keys=np.asarray(uniq, dtype=np.int64).reshape(-1, 4),
                     stats=dict(n_boxes=test.n_boxes, n_leaves=int(test.is_leaf.sum()), n_pairs=len(pairs),
                                n_near_blocks=len(npairs), n_near_values=n_near_values, n_ops=len(ops)))
