"""Test helper: the oracle's core data in the product's host description (MapArrays).

Parity tests hand the HIP path exactly the arrays the oracle traverses, so a mismatch can only
come from the matvec itself.
"""

import numpy as np

from h2g.method import MapArrays, SideArrays


def _side(tree, coll, transfers, dt):
    K = coll.shape[0]
    tr = np.zeros((tree.n_levels, 8, K, K), dtype=dt)
    for l in range(tree.n_levels):
        for s in range(8):
            tr[l, s] = np.asarray(transfers[l][s]).T
    return SideArrays(tree=tree, collection=np.ascontiguousarray(coll.T, dtype=dt), transfers=tr)


def to_arrays(core) -> MapArrays:
    cplx = np.iscomplexobj(core.coll_test) or np.iscomplexobj(core.transfers_test[1][0] if len(core.transfers_test) > 1 else 0) \
        or (core.m2l_kind in ("dense", "symdense", "diagonal") and np.iscomplexobj(core.ops)) \
        or any(np.iscomplexobj(v) for v in core.near_values) or np.iscomplexobj(core.alpha) and complex(core.alpha).imag != 0
    dt = np.complex128 if cplx else np.float64
    ts = _side(core.test, core.coll_test, core.transfers_test, dt)
    cs = ts if core.galerkin else _side(core.trial, core.coll_trial, core.transfers_trial, dt)
    if core.m2l_kind in ("dense", "symdense"):
        ops = np.ascontiguousarray(np.swapaxes(np.asarray(core.ops), 1, 2), dtype=dt)
    elif core.m2l_kind == "diagonal":
        ops = np.ascontiguousarray(core.ops, dtype=dt)
    else:
        ops = np.ascontiguousarray(core.ops, dtype=np.complex128)
    sizes = np.array([v.size for v in core.near_values], dtype=np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64) if len(sizes) else np.zeros(0, np.int64)
    vals = np.concatenate([np.asarray(v, dtype=dt).T.ravel() for v in core.near_values]) if len(sizes) else np.zeros(0, dt)
    return MapArrays(dtype="c64" if cplx else "f64", K=core.K, p=core.p, test=ts, trial=cs, galerkin=core.galerkin,
                     m2l_kind=core.m2l_kind, pairs=core.pairs, pair_op=core.pair_op, pair_sym=core.pair_sym,
                     ops=ops, perms=core.perms, near_pairs=core.near_pairs, near_offset=off,
                     n_near_values=int(sizes.sum()), near_values=vals, alpha=core.alpha)
