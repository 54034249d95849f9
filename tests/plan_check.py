"""Checks of a device-built plan (h2g.plan.device_plan) against the oracle's tree and pair lists
and against the host builder.  Shared by the CPU replay and the GPU tests."""

import numpy as np

from oracle import method as OM
from oracle import tree as OT

TREE_FIELDS = ("perm", "level_offsets", "level", "coord", "parent", "sector", "begin", "end", "is_leaf")

# point sets: (name, test points, trial points or None, leaf size, origin, size, symmetric)
def point_sets(scale=1):
    rng = np.random.default_rng(5)
    cube = rng.random((3000 * scale, 3))
    sph = rng.standard_normal((2000 * scale, 3))
    sph /= np.linalg.norm(sph, axis=1, keepdims=True)
    blobs = np.concatenate([0.5 + 0.02 * rng.standard_normal((1500 * scale, 3)), rng.random((500 * scale, 3))])
    blobs = np.clip(blobs, 0.0, 0.999999)
    dup = np.concatenate([np.full((150, 3), 0.3), rng.random((2000 * scale, 3))])  # 150 equal points: depth 21
    return [
        ("cube", cube, None, 40, (0.0, 0.0, 0.0), 1.0, False),
        ("cube_symmetric", cube, None, 40, (0.0, 0.0, 0.0), 1.0, True),
        ("sphere", sph, None, 30, (-1.0, -1.0, -1.0), 2.0, False),
        ("clustered", blobs, None, 24, (0.0, 0.0, 0.0), 1.0, True),
        ("duplicates", dup, None, 40, (0.0, 0.0, 0.0), 1.0, False),
        ("petrov", cube[: 1700 * scale], rng.random((900 * scale, 3)), 20, (0.0, 0.0, 0.0), 1.0, True),
        ("single_leaf", cube[:37], None, 64, (0.0, 0.0, 0.0), 1.0, False),
    ]


def check_against_oracle(dp, pts_test, pts_trial, leaf, origin, size, symmetric):
    """Trees, pair lists, translation ids and near offsets equal to the oracle's."""
    ot = OT.build_tree(pts_test, leaf, origin, size)
    oc = ot if pts_trial is None else OT.build_tree(pts_trial, leaf, origin, size)
    for mine, ref in ((dp.test, ot), (dp.trial, oc)):
        for f in TREE_FIELDS:
            assert np.array_equal(getattr(mine, f), getattr(ref, f)), f
    pairs = OT.m2l_pairs(ot, oc)
    assert np.array_equal(dp.pairs, pairs)
    near = OT.near_pairs(ot, oc)
    assert np.array_equal(dp.near_pairs, near)
    # every translation id stands for one (level, offset) -- (level, canonical offset) and the
    # pair's symmetry for symmetric plans -- and every such key has one id
    group = OM.symmetry_group()
    assert len(np.unique(dp.keys, axis=0)) == len(dp.keys)
    assert np.array_equal(np.unique(dp.pair_op), np.arange(len(dp.keys)))
    for k in range(len(pairs)):
        a, b = pairs[k]
        o = tuple(int(v) for v in oc.coord[b] - ot.coord[a])
        key = tuple(int(v) for v in dp.keys[dp.pair_op[k]])
        if symmetric:
            c, S = OM.canonical_offset(o)
            assert key == (int(ot.level[a]),) + tuple(c)
            assert group[dp.pair_sym[k]] == S
        else:
            assert key == (int(ot.level[a]),) + o
    sizes = (ot.end - ot.begin)[near[:, 0]] * (oc.end - oc.begin)[near[:, 1]]
    assert np.array_equal(dp.near_offset, np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64))
    assert dp.n_near_values == int(sizes.sum())


def check_against_host_builder(dp, pts_test, pts_trial, leaf, origin, size, symmetric):
    """The same arrays as h2g.tree and method.translation_ids (large maps)."""
    from h2g import method as PM
    from h2g import tree as T

    test = T.build_octree(pts_test, leaf, origin, size)
    trial = test if pts_trial is None else T.build_octree(pts_trial, leaf, origin, size)
    for mine, ref in ((dp.test, test), (dp.trial, trial)):
        for f in TREE_FIELDS + ("key",):
            assert np.array_equal(getattr(mine, f), getattr(ref, f)), f
    pairs = T.m2l_pairs(test, trial)
    assert np.array_equal(dp.pairs, pairs)
    op, sym, uniq = PM.translation_ids(test, trial, pairs, symmetric)
    assert np.array_equal(dp.pair_op, op) and np.array_equal(dp.keys, uniq)
    assert (sym is None and dp.pair_sym is None) or np.array_equal(dp.pair_sym, sym)
    near = T.near_pairs(test, trial)
    assert np.array_equal(dp.near_pairs, near)
