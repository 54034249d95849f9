"""Octree, near/far partition and M2L interaction lists -- TEST INFRASTRUCTURE ONLY.

The reference takes its trees and traversal plans from H2Trees 0.4.2, which is not vendored
(call sites: ``src/maps/interactionstrategy.jl:50-60``, ``src/maps/schedules.jl:25-53``,
``src/farinteractions/traversalhelpers.jl:23,50,108``).  Its splitting rule and plan
construction cannot be imported, so this module states the build's own definition, which
the product builder must reproduce exactly:

* root cube ``[origin, origin + size)^3``; a box at level ``l`` with integer coordinate
  ``c`` covers ``origin + size*[c, c+1) / 2^l``; points get 21-bit integer coordinates;
* a box splits while it holds more than ``leaf_size`` points (and ``level < max_level``);
  only non-empty children exist; boxes of a level are in Morton order (children of one
  parent are contiguous, in sector order ``bx + 2*by + 4*bz``); points are in Morton
  order, so every box owns a contiguous range of the tree-ordered degrees of freedom;
* two same-level boxes are adjacent when their coordinates differ by at most 1 on every
  axis;
* M2L pairs (the ``DisaggregateTranslatePlan`` translating nodes): same-level test box
  ``b`` and trial box ``c`` with adjacent parents that are not adjacent themselves;
* near pairs: a test leaf and a trial leaf whose ancestors at the coarser of the two leaf
  levels are adjacent.
Every pair of points is then covered exactly once, by one near pair or by one M2L pair at
the first level where their ancestors separate (checked in ``tests/test_oracle_h2.py`` by
assembling the dense operator both ways).  Parity with H2Trees itself is unpinned; the
tree and plan semantics are the build's own.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

DEPTH = 21


def morton3(q: np.ndarray) -> np.ndarray:
    """Interleave 21-bit integer coordinates (x least significant in each 3-bit digit)."""
    q = q.astype(np.uint64)
    code = np.zeros(q.shape[0], dtype=np.uint64)
    for b in range(DEPTH - 1, -1, -1):
        code <<= np.uint64(3)
        code |= ((q[:, 2] >> np.uint64(b)) & np.uint64(1)) << np.uint64(2)
        code |= ((q[:, 1] >> np.uint64(b)) & np.uint64(1)) << np.uint64(1)
        code |= (q[:, 0] >> np.uint64(b)) & np.uint64(1)
    return code


@dataclass
class Tree:
    """One side's octree, boxes level-contiguous."""

    points: np.ndarray          # (N, 3) original order
    origin: np.ndarray
    size: float
    perm: np.ndarray            # tree position -> original index
    level_offsets: np.ndarray   # (n_levels + 1,)
    level: np.ndarray           # per box
    coord: np.ndarray           # (n_boxes, 3) int64
    parent: np.ndarray          # -1 for the root
    sector: np.ndarray          # -1 for the root
    begin: np.ndarray           # dof range in tree order
    end: np.ndarray
    is_leaf: np.ndarray
    child_begin: np.ndarray     # children range (box ids), empty for leaves
    child_end: np.ndarray
    lookup: list = field(default_factory=list)  # per level dict coord -> box

    @property
    def n_boxes(self):
        return len(self.level)

    @property
    def n_levels(self):
        return len(self.level_offsets) - 1

    def width(self, level):
        return self.size / (2 ** level)

    def center(self, b):
        w = self.width(self.level[b])
        return self.origin + w * (self.coord[b] + 0.5)

    def leaves(self):
        return np.nonzero(self.is_leaf)[0]


def build_tree(points, leaf_size=64, origin=(0.0, 0.0, 0.0), size=1.0, max_level=DEPTH) -> Tree:
    points = np.ascontiguousarray(points, dtype=np.float64)
    origin = np.asarray(origin, dtype=np.float64)
    n = points.shape[0]
    scale = float(2 ** DEPTH)
    q = np.floor((points - origin[None, :]) / size * scale)
    q = np.clip(q, 0, scale - 1).astype(np.int64)
    code = morton3(q)
    perm = np.argsort(code, kind="stable")
    qs = q[perm]

    level_l, coord_l, parent_l, sector_l, begin_l, end_l, leaf_l = [], [], [], [], [], [], []
    offsets = [0]
    # level 0
    cur = [(np.zeros(3, dtype=np.int64), -1, -1, 0, n)]
    lvl = 0
    while cur:
        nxt = []
        base = len(level_l)
        for i, (c, par, sec, b, e) in enumerate(cur):
            level_l.append(lvl)
            coord_l.append(c)
            parent_l.append(par)
            sector_l.append(sec)
            begin_l.append(b)
            end_l.append(e)
            split = (e - b) > leaf_size and lvl < max_level
            leaf_l.append(not split)
            if split:
                shift = DEPTH - (lvl + 1)
                sub = qs[b:e] >> shift
                digit = (sub[:, 0] & 1) + 2 * (sub[:, 1] & 1) + 4 * (sub[:, 2] & 1)
                for s in range(8):
                    idx = np.nonzero(digit == s)[0]
                    if idx.size == 0:
                        continue
                    cb, ce = b + int(idx[0]), b + int(idx[-1]) + 1
                    assert ce - cb == idx.size
                    cc = 2 * c + np.array([s & 1, (s >> 1) & 1, (s >> 2) & 1], dtype=np.int64)
                    nxt.append((cc, base + i, s, cb, ce))
        offsets.append(len(level_l))
        cur = nxt
        lvl += 1

    nb = len(level_l)
    parent = np.array(parent_l, dtype=np.int64)
    child_begin = np.zeros(nb, dtype=np.int64)
    child_end = np.zeros(nb, dtype=np.int64)
    for b in range(nb - 1, 0, -1):
        p = parent[b]
        if child_end[p] == 0:
            child_end[p] = b + 1
        child_begin[p] = b
    t = Tree(
        points=points,
        origin=origin,
        size=float(size),
        perm=perm.astype(np.int64),
        level_offsets=np.array(offsets, dtype=np.int64),
        level=np.array(level_l, dtype=np.int64),
        coord=np.array(coord_l, dtype=np.int64).reshape(-1, 3),
        parent=parent,
        sector=np.array(sector_l, dtype=np.int64),
        begin=np.array(begin_l, dtype=np.int64),
        end=np.array(end_l, dtype=np.int64),
        is_leaf=np.array(leaf_l, dtype=bool),
        child_begin=child_begin,
        child_end=child_end,
    )
    t.lookup = [
        {tuple(t.coord[b]): b for b in range(t.level_offsets[l], t.level_offsets[l + 1])}
        for l in range(t.n_levels)
    ]
    return t


def adjacent(c1, c2) -> bool:
    return bool(np.max(np.abs(np.asarray(c1) - np.asarray(c2))) <= 1)


def m2l_pairs(test: Tree, trial: Tree):
    """All (test box, trial box) M2L pairs, in test-box order then trial-box order."""
    pairs = []
    nl = min(test.n_levels, trial.n_levels)
    for l in range(1, nl):
        for b in range(test.level_offsets[l], test.level_offsets[l + 1]):
            pb = test.parent[b]
            pc = test.coord[pb]
            found = []
            for d in np.ndindex(3, 3, 3):
                nc = tuple(int(v) for v in pc + np.array(d) - 1)
                q = trial.lookup[l - 1].get(nc)
                if q is None:
                    continue
                for c in range(trial.child_begin[q], trial.child_end[q]):
                    if not adjacent(test.coord[b], trial.coord[c]):
                        found.append(c)
            for c in sorted(found):
                pairs.append((b, c))
    return np.array(pairs, dtype=np.int64).reshape(-1, 2)


def near_pairs(test: Tree, trial: Tree):
    """All (test leaf, trial leaf) near pairs, in test-leaf order then trial-leaf order."""
    tl = test.leaves()
    cl = trial.leaves()
    out = []
    for b in tl:
        lb = test.level[b]
        lc = trial.level[cl]
        m = np.minimum(lb, lc)
        cb = test.coord[b][None, :] >> (lb - m)[:, None]
        cc = trial.coord[cl] >> (lc - m)[:, None]
        adj = np.max(np.abs(cb - cc), axis=1) <= 1
        for c in cl[adj]:
            out.append((b, c))
    return np.array(out, dtype=np.int64).reshape(-1, 2)
