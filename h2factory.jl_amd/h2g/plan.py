"""The octree and traversal plans built on the device (``h2g_plan_build``).

The same trees and pair lists as the host builder (``tree.build_octree``, ``tree.m2l_pairs``,
``tree.near_pairs`` and the translation ids of ``method.build_map``), bit for bit, computed by
the library's plan kernels: Morton codes and their sort, the level-by-level box split, the M2L
pairs with their translation ids (and lattice symmetries) and the near pairs with their value
offsets.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as H
from .tree import Octree


@dataclass
class DevicePlan:
    test: Octree
    trial: Octree
    pairs: np.ndarray          # (P, 2) (test box, trial box)
    pair_op: np.ndarray        # (P,) translation id
    pair_sym: np.ndarray | None
    keys: np.ndarray           # (n_ids, 4): level and (canonical) offset of each translation id
    near_pairs: np.ndarray     # (B, 2) (test leaf, trial leaf)
    near_offset: np.ndarray    # (B,)
    n_near_values: int


def _tree_out(n, nlev, nb):
    a = dict(perm=np.empty(n, np.int64), level_offsets=np.empty(nlev + 1, np.int64), level=np.empty(nb, np.int32),
             coord=np.empty((nb, 3), np.int32), key=np.empty(nb, np.uint64), parent=np.empty(nb, np.int32),
             sector=np.empty(nb, np.int32), begin=np.empty(nb, np.int64), end=np.empty(nb, np.int64),
             is_leaf=np.empty(nb, np.uint8))
    o = H.PlanTreeOut()
    o.perm, o.level_offsets = H.ptr(a["perm"], C.c_int64), H.ptr(a["level_offsets"], C.c_int64)
    o.level, o.coord = H.ptr(a["level"], C.c_int32), H.ptr(a["coord"], C.c_int32)
    o.key = H.ptr(a["key"], C.c_uint64)
    o.parent, o.sector = H.ptr(a["parent"], C.c_int32), H.ptr(a["sector"], C.c_int32)
    o.begin, o.end = H.ptr(a["begin"], C.c_int64), H.ptr(a["end"], C.c_int64)
    o.is_leaf = H.ptr(a["is_leaf"], C.c_uint8)
    return a, o


def _octree(points, origin, size, a):
    return Octree(points=points, origin=np.asarray(origin, dtype=np.float64), size=float(size), perm=a["perm"],
                  level_offsets=a["level_offsets"], level=a["level"].astype(np.int64),
                  coord=a["coord"].astype(np.int64), key=a["key"], parent=a["parent"].astype(np.int64),
                  sector=a["sector"].astype(np.int64), begin=a["begin"], end=a["end"],
                  is_leaf=a["is_leaf"].astype(bool))


def device_plan(points_test, points_trial=None, leaf_size=64, origin=(0.0, 0.0, 0.0), size=1.0,
                symmetric=False, max_level=0, device=0) -> DevicePlan:
    """Trees and pair lists of a Galerkin (``points_trial`` None) or Petrov map, built on GPU
    ``device``."""
    L = H.lib()
    pt = np.ascontiguousarray(points_test, dtype=np.float64)
    ps = None if points_trial is None else np.ascontiguousarray(points_trial, dtype=np.float64)
    opts = H.PlanOpts()
    opts.origin[:] = [float(v) for v in origin]
    opts.size, opts.leaf_size, opts.max_level = float(size), int(leaf_size), int(max_level)
    opts.symmetric, opts.device = int(bool(symmetric)), int(device)
    handle = C.c_void_p()
    H.check(L.h2g_plan_build(H.ptr(pt, C.c_double), pt.shape[0], H.ptr(ps, C.c_double) if ps is not None else None,
                             0 if ps is None else ps.shape[0], C.byref(opts), C.byref(handle)))
    try:
        cnt = H.PlanCounts()
        H.check(L.h2g_plan_info(handle, C.byref(cnt)))
        out = H.PlanOut()
        ta, out.test = _tree_out(cnt.n_dofs[0], cnt.n_levels[0], cnt.n_boxes[0])
        sa = None
        if not cnt.galerkin:
            sa, out.trial = _tree_out(cnt.n_dofs[1], cnt.n_levels[1], cnt.n_boxes[1])
        P, B = cnt.n_pairs, cnt.n_near_pairs
        pair_test, pair_trial, pair_op = (np.empty(P, np.int32) for _ in range(3))
        pair_sym = np.empty(P, np.int32) if symmetric else None
        keys = np.empty((cnt.n_keys, 4), np.int32)
        near_test, near_trial, near_offset = np.empty(B, np.int32), np.empty(B, np.int32), np.empty(B, np.int64)
        out.pair_test, out.pair_trial = H.ptr(pair_test, C.c_int32), H.ptr(pair_trial, C.c_int32)
        out.pair_op = H.ptr(pair_op, C.c_int32)
        out.pair_sym = H.ptr(pair_sym, C.c_int32) if symmetric else None
        out.keys = H.ptr(keys, C.c_int32)
        out.near_test, out.near_trial = H.ptr(near_test, C.c_int32), H.ptr(near_trial, C.c_int32)
        out.near_offset = H.ptr(near_offset, C.c_int64)
        H.check(L.h2g_plan_copy(handle, C.byref(out)))
        n_near_values = int(cnt.n_near_values)
    finally:
        L.h2g_plan_destroy(handle)
    test = _octree(pt, origin, size, ta)
    trial = test if sa is None else _octree(ps, origin, size, sa)
    return DevicePlan(test=test, trial=trial,
                      pairs=np.stack([pair_test, pair_trial], axis=1).astype(np.int64),
                      pair_op=pair_op.astype(np.int64),
                      pair_sym=None if pair_sym is None else pair_sym.astype(np.int64),
                      keys=keys.astype(np.int64),
                      near_pairs=np.stack([near_test, near_trial], axis=1).astype(np.int64),
                      near_offset=near_offset, n_near_values=n_near_values)
