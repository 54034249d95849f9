"""Octree and traversal plans of the interpolation method, vectorised for millions of points.

Definitions (the build's own; H2Trees, which the reference uses, is not vendored):

* root cube ``[origin, origin + size)^3``; points get 21-bit integer coordinates and are
  ordered by their Morton code, so every box owns a contiguous range of tree-ordered dofs;
* a box is split while it holds more than ``leaf_size`` points; only non-empty children
  exist; boxes are stored level by level in Morton order; a child's sector is
  ``bx + 2*by + 4*bz``;
* M2L pairs: same-level boxes whose parents are adjacent (Chebyshev distance <= 1) and that
  are not adjacent themselves -- the translating nodes of a disaggregate-translate plan
  (reference: ``src/farinteractions/traversalhelpers.jl:167-219``);
* near pairs: a test leaf and a trial leaf whose ancestors at the coarser of the two leaf
  levels are adjacent (the ``isnear`` predicate, ``src/maps/interactionstrategy.jl:50-60``).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

DEPTH = 21


def _morton(q: np.ndarray) -> np.ndarray:
    q = q.astype(np.uint64)
    code = np.zeros(q.shape[0], dtype=np.uint64)
    one = np.uint64(1)
    for b in range(DEPTH - 1, -1, -1):
        bb = np.uint64(b)
        code = (code << np.uint64(3)) | (((q[:, 2] >> bb) & one) << np.uint64(2)) \
            | (((q[:, 1] >> bb) & one) << one) | ((q[:, 0] >> bb) & one)
    return code


def _deinterleave(key: np.ndarray, level: int) -> np.ndarray:
    """Integer coordinates of level-``level`` Morton keys."""
    key = key.astype(np.uint64)
    out = np.zeros((key.size, 3), dtype=np.int64)
    for b in range(level):
        digit = (key >> np.uint64(3 * b)) & np.uint64(7)
        out[:, 0] |= ((digit & np.uint64(1)).astype(np.int64)) << b
        out[:, 1] |= (((digit >> np.uint64(1)) & np.uint64(1)).astype(np.int64)) << b
        out[:, 2] |= (((digit >> np.uint64(2)) & np.uint64(1)).astype(np.int64)) << b
    return out


def _interleave(c: np.ndarray, level: int) -> np.ndarray:
    c = c.astype(np.uint64)
    key = np.zeros(c.shape[0], dtype=np.uint64)
    for b in range(level):
        bb = np.uint64(b)
        one = np.uint64(1)
        digit = ((c[:, 0] >> bb) & one) | (((c[:, 1] >> bb) & one) << one) | (((c[:, 2] >> bb) & one) << np.uint64(2))
        key |= digit << np.uint64(3 * b)
    return key


@dataclass
class Octree:
    points: np.ndarray
    origin: np.ndarray
    size: float
    perm: np.ndarray
    level_offsets: np.ndarray
    level: np.ndarray
    coord: np.ndarray
    key: np.ndarray            # Morton key of the box at its level
    parent: np.ndarray
    sector: np.ndarray
    begin: np.ndarray
    end: np.ndarray
    is_leaf: np.ndarray

    @property
    def n_boxes(self):
        return len(self.level)

    @property
    def n_levels(self):
        return len(self.level_offsets) - 1

    def width(self, level):
        return self.size / (2.0 ** np.asarray(level))

    def centers(self):
        w = self.width(self.level)
        return self.origin[None, :] + (self.coord + 0.5) * w[:, None]

    def halfwidths(self):
        return 0.5 * self.width(self.level)

    def leaves(self):
        return np.flatnonzero(self.is_leaf)

    def find(self, level: int, coords: np.ndarray) -> np.ndarray:
        """Box ids of level-``level`` boxes with integer ``coords`` (-1 where absent)."""
        lo, hi = self.level_offsets[level], self.level_offsets[level + 1]
        out = np.full(coords.shape[0], -1, dtype=np.int64)
        if hi == lo or coords.shape[0] == 0:
            return out
        n = 2 ** level
        ok = np.all((coords >= 0) & (coords < n), axis=1)
        keys = _interleave(coords[ok], level)
        lk = self.key[lo:hi]
        pos = np.searchsorted(lk, keys)
        pos_c = np.minimum(pos, hi - lo - 1)
        hit = lk[pos_c] == keys
        res = np.where(hit, lo + pos_c, -1)
        out[ok] = res
        return out


def build_octree(points, leaf_size=64, origin=(0.0, 0.0, 0.0), size=1.0, max_level=DEPTH) -> Octree:
    points = np.ascontiguousarray(points, dtype=np.float64)
    origin = np.asarray(origin, dtype=np.float64)
    n = points.shape[0]
    scale = float(2 ** DEPTH)
    q = np.clip(np.floor((points - origin[None, :]) / size * scale), 0, scale - 1).astype(np.int64)
    code = _morton(q)
    perm = np.argsort(code, kind="stable")
    codes = code[perm]

    lv_begin = [np.array([0], dtype=np.int64)]
    lv_end = [np.array([n], dtype=np.int64)]
    lv_key = [np.array([0], dtype=np.uint64)]
    lv_parent = [np.array([-1], dtype=np.int64)]
    offsets = [0, 1]
    l = 0
    while True:
        b, e = lv_begin[-1], lv_end[-1]
        split = ((e - b) > leaf_size) & (l < max_level)
        if not split.any():
            break
        sb, se = b[split], e[split]
        split_ids = np.flatnonzero(split) + offsets[l]
        lens = se - sb
        tot = int(lens.sum())
        starts_rep = np.repeat(sb - np.concatenate([[0], np.cumsum(lens)[:-1]]), lens)
        idx = np.arange(tot, dtype=np.int64) + starts_rep
        shift = np.uint64(3 * (DEPTH - (l + 1)))
        key = codes[idx] >> shift
        new = np.ones(tot, dtype=bool)
        new[1:] = key[1:] != key[:-1]
        st = np.flatnonzero(new)
        ln = np.diff(np.append(st, tot))
        cb = idx[st]
        owner = np.repeat(np.arange(len(sb)), lens)[st]
        lv_begin.append(cb)
        lv_end.append(cb + ln)
        lv_key.append(key[st])
        lv_parent.append(split_ids[owner])
        offsets.append(offsets[-1] + len(cb))
        l += 1

    begin = np.concatenate(lv_begin)
    end = np.concatenate(lv_end)
    key = np.concatenate(lv_key)
    parent = np.concatenate(lv_parent)
    level = np.repeat(np.arange(len(lv_begin)), [len(x) for x in lv_begin])
    coord = np.zeros((len(begin), 3), dtype=np.int64)
    for li in range(len(lv_begin)):
        lo, hi = offsets[li], offsets[li + 1]
        coord[lo:hi] = _deinterleave(key[lo:hi], li)
    sector = np.where(parent >= 0, (key & np.uint64(7)).astype(np.int64), -1)
    has_child = np.zeros(len(begin), dtype=bool)
    has_child[parent[parent >= 0]] = True
    return Octree(points=points, origin=origin, size=float(size), perm=perm.astype(np.int64),
                  level_offsets=np.array(offsets, dtype=np.int64), level=level.astype(np.int64), coord=coord,
                  key=key, parent=parent.astype(np.int64), sector=sector, begin=begin, end=end,
                  is_leaf=~has_child)


def _sector_offsets():
    """Per receiving sector: the 189 offsets to children of the parent's neighbours that are
    not adjacent."""
    out = []
    for s in range(8):
        bits = [(s >> a) & 1 for a in range(3)]
        rng = [np.arange(-2, 4) if bits[a] == 0 else np.arange(-3, 3) for a in range(3)]
        oz, oy, ox = np.meshgrid(rng[2], rng[1], rng[0], indexing="ij")
        o = np.stack([ox.ravel(), oy.ravel(), oz.ravel()], axis=1)
        keep = np.max(np.abs(o), axis=1) > 1
        out.append(o[keep])
    return out


SECTOR_OFFSETS = _sector_offsets()


def m2l_pairs(test: Octree, trial: Octree):
    """(P, 2) array of (test box, trial box) M2L pairs, grouped by test box, trial ascending."""
    res = []
    for l in range(1, min(test.n_levels, trial.n_levels)):
        lo, hi = test.level_offsets[l], test.level_offsets[l + 1]
        boxes = np.arange(lo, hi)
        sec = test.sector[lo:hi]
        for s in range(8):
            bs = boxes[sec == s]
            if bs.size == 0:
                continue
            offs = SECTOR_OFFSETS[s]
            cand = (test.coord[bs][:, None, :] + offs[None, :, :]).reshape(-1, 3)
            ids = trial.find(l, cand).reshape(bs.size, offs.shape[0])
            rows = np.repeat(bs, offs.shape[0]).reshape(bs.size, -1)
            ok = ids >= 0
            res.append(np.stack([rows[ok], ids[ok]], axis=1))
    if not res:
        return np.zeros((0, 2), dtype=np.int64)
    pairs = np.concatenate(res)
    order = np.lexsort((pairs[:, 1], pairs[:, 0]))
    return pairs[order]


_NEIGH = np.array([[dx, dy, dz] for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)], dtype=np.int64)


def _leaf_descendants(tree: Octree, boxes: np.ndarray):
    """For each box, the range [lo, hi) into tree.leaves() (sorted by dof begin) of the leaves
    of its subtree."""
    leaves = tree.leaves()
    lb = tree.begin[leaves]
    order = np.argsort(lb, kind="stable")
    leaves, lb = leaves[order], lb[order]
    lo = np.searchsorted(lb, tree.begin[boxes], side="left")
    hi = np.searchsorted(lb, tree.end[boxes], side="left")
    return leaves, lo, hi


def _expand(rows, leaves, lo, hi):
    cnt = hi - lo
    r = np.repeat(rows, cnt)
    start = np.repeat(lo - np.concatenate([[0], np.cumsum(cnt)[:-1]]), cnt)
    return r, leaves[np.arange(cnt.sum()) + start]


def near_pairs(test: Octree, trial: Octree):
    """(B, 2) array of (test leaf, trial leaf) near pairs, grouped by test leaf, trial ascending."""
    out = []

    def from_side(a: Octree, b: Octree, same_level_leaves: bool):
        la = a.leaves()
        for l in range(a.n_levels):
            ls = la[a.level[la] == l]
            if ls.size == 0 or l >= b.n_levels:
                continue
            cand = (a.coord[ls][:, None, :] + _NEIGH[None, :, :]).reshape(-1, 3)
            ids = b.find(l, cand)
            rows = np.repeat(ls, 27)
            ok = ids >= 0
            rows, ids = rows[ok], ids[ok]
            leaf = b.is_leaf[ids]
            if same_level_leaves:
                out_l = [np.stack([rows[leaf], ids[leaf]], axis=1)]
            else:
                out_l = []
            nl = ~leaf
            if nl.any():
                leaves, lo, hi = _leaf_descendants(b, ids[nl])
                r, d = _expand(rows[nl], leaves, lo, hi)
                out_l.append(np.stack([r, d], axis=1))
            yield from out_l

    for arr in from_side(test, trial, True):
        out.append(arr)
    for arr in from_side(trial, test, False):
        out.append(arr[:, ::-1])
    if not out:
        return np.zeros((0, 2), dtype=np.int64)
    pairs = np.concatenate(out)
    order = np.lexsort((pairs[:, 1], pairs[:, 0]))
    return pairs[order]
