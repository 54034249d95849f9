"""CPU checks of the product's host side: the method builder reproduces the oracle's trees,
plans and operators exactly, the C ABI library loads and exports every declared entry point,
and malformed core descriptions are rejected before any device call."""

import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import method as OM
from oracle import tree as OT

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("n,leaf,petrov", [(3000, 64, False), (2000, 16, True), (600, 8, False)])
def test_builder_matches_oracle(n, leaf, petrov):
    from h2g import tree as PT

    rng = np.random.default_rng(n)
    a = rng.random((n, 3))
    b = rng.random((n // 2 + 7, 3)) if petrov else None
    ot, pt = OT.build_tree(a, leaf), PT.build_octree(a, leaf)
    oc = OT.build_tree(b, leaf) if petrov else ot
    pc = PT.build_octree(b, leaf) if petrov else pt
    for f in ("perm", "level_offsets", "level", "coord", "parent", "sector", "begin", "end", "is_leaf"):
        assert np.array_equal(getattr(ot, f), getattr(pt, f)), f
    assert np.array_equal(OT.m2l_pairs(ot, oc), PT.m2l_pairs(pt, pc))
    assert np.array_equal(OT.near_pairs(ot, oc), PT.near_pairs(pt, pc))


def test_operators_match_oracle():
    from h2g import method as PM

    nodes = PM.nodes1d(6, "equi")
    assert np.allclose(nodes, OM.nodes1d(6, "equi"), atol=1e-15)
    for s in range(8):
        assert np.allclose(PM.transfer_matrix(nodes, s), OM.transfer(nodes, s), atol=1e-15)
    offs = np.array([[2, 0, 1], [-3, 2, -2], [0, 0, 3]])
    Dp = PM.toeplitz_spectra("laplace", 0.0, 6, 0.25, offs)
    Do = np.stack([np.fft.fftn(OM.toeplitz_kappa("laplace", 0.0, 6, 0.25, o)) for o in offs])
    assert np.allclose(Dp, Do, rtol=1e-14, atol=1e-12)
    Mp = PM.dense_couplings("helmholtz", 10.0, nodes, 0.25, offs)
    Mo = np.stack([OM.m2l_dense("helmholtz", 10.0, nodes, 0.25, o) for o in offs])
    assert np.allclose(Mp, Mo, rtol=1e-13)
    g = OM.symmetry_group()
    assert PM.GROUP == g
    for s in g:
        assert np.array_equal(PM.toeplitz_perm(11, s), OM.toeplitz_permutation((11, 11, 11), s))
        assert np.array_equal(PM.grid_perm(6, s), OM.lattice_grid_permutation((6, 6, 6), s))
    rng = np.random.default_rng(0)
    o = rng.integers(-3, 4, size=(50, 3))
    o = o[np.max(np.abs(o), axis=1) > 1]
    canon, sym = PM.canonicalize(o)
    for i in range(len(o)):
        c, S = OM.canonical_offset(o[i])
        assert tuple(canon[i]) == c and PM.GROUP[sym[i]] == S


def _declared_functions():
    txt = open(os.path.join(ROOT, "include", "h2g.h")).read()
    return sorted(set(re.findall(r"^(?:int|void \*)\s*(h2g_\w+)\(", txt, flags=re.M)))


def test_library_exports_every_declared_symbol():
    from h2g import _lib

    lib = _lib.lib()
    names = _declared_functions()
    assert len(names) >= 17
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.EXPORTED)
    assert lib.h2g_version() >= 100


def _tiny_desc(galerkin=True):
    """A valid 2-level description (root + 2 leaves) the tests then break one field at a time."""
    from h2g import _lib as H

    keep = {}
    keep["perm"] = np.array([0, 1, 2, 3], dtype=np.int64)
    keep["lo"] = np.array([0, 1, 3], dtype=np.int32)
    keep["par"] = np.array([-1, 0, 0], dtype=np.int32)
    keep["sec"] = np.array([-1, 0, 7], dtype=np.int32)
    keep["b"] = np.array([0, 0, 2], dtype=np.int64)
    keep["e"] = np.array([4, 2, 4], dtype=np.int64)
    keep["leaf"] = np.array([0, 1, 1], dtype=np.uint8)
    keep["coll"] = np.ones((4, 1))
    keep["tr"] = np.ones((2, 8, 1, 1))
    s = H.SideDesc()
    s.n_dofs, s.n_levels, s.n_boxes = 4, 2, 3
    s.perm = H.ptr(keep["perm"], C.c_int64)
    s.level_offsets = H.ptr(keep["lo"], C.c_int32)
    s.parent = H.ptr(keep["par"], C.c_int32)
    s.sector = H.ptr(keep["sec"], C.c_int32)
    s.dof_begin = H.ptr(keep["b"], C.c_int64)
    s.dof_end = H.ptr(keep["e"], C.c_int64)
    s.is_leaf = H.ptr(keep["leaf"], C.c_uint8)
    s.collection = H.ptr(keep["coll"])
    s.transfers = H.ptr(keep["tr"])
    d = H.CoreDesc()
    d.dtype, d.K, d.p, d.galerkin = H.H2G_F64, 1, 0, 1
    d.test = s
    d.m2l_kind = H.H2G_M2L_DENSE
    keep["nt"] = np.array([1, 2], dtype=np.int32)
    keep["nc"] = np.array([1, 2], dtype=np.int32)
    keep["no"] = np.array([0, 4], dtype=np.int64)
    keep["nv"] = np.ones(8)
    d.n_near_blocks = 2
    d.near_test, d.near_trial = H.ptr(keep["nt"], C.c_int32), H.ptr(keep["nc"], C.c_int32)
    d.near_offset = H.ptr(keep["no"], C.c_int64)
    d.n_near_values = 8
    d.near_values = H.ptr(keep["nv"])
    d.alpha_re = 1.0
    d.device = 0
    return d, keep


@pytest.mark.parametrize("breakage,code", [
    ("dtype", "EINVAL"), ("perm", "EINVAL"), ("levels", "EDIM"), ("near_range", "EDIM"),
    ("near_leaf", "EINVAL"), ("toeplitz_p", "EINVAL"), ("pair_range", "EINVAL"), ("sym_perm", "EDIM"),
])
def test_core_validation_errors(breakage, code):
    """Malformed descriptions fail with the mapped error (ArgumentError / DimensionMismatch in the
    reference) before any device work, so this runs without a GPU."""
    from h2g import _lib as H

    d, keep = _tiny_desc()
    if breakage == "dtype":
        d.dtype = 7
    elif breakage == "perm":
        keep["perm"][1] = 0
    elif breakage == "levels":
        keep["lo"][2] = 5
    elif breakage == "near_range":
        keep["no"][1] = 6
    elif breakage == "near_leaf":
        keep["nt"][0] = 0
    elif breakage == "toeplitz_p":
        d.m2l_kind, d.p = H.H2G_M2L_TOEPLITZ, 3
    elif breakage == "pair_range":
        keep["pt"] = np.array([5], dtype=np.int32)
        keep["pc"] = np.array([1], dtype=np.int32)
        keep["po"] = np.array([0], dtype=np.int32)
        keep["ops"] = np.ones(1)
        d.n_pairs, d.n_ops = 1, 1
        d.pair_test, d.pair_trial, d.pair_op = (H.ptr(keep[k], C.c_int32) for k in ("pt", "pc", "po"))
        d.ops = H.ptr(keep["ops"])
    elif breakage == "sym_perm":
        keep["pt"] = np.array([1], dtype=np.int32)
        keep["pc"] = np.array([2], dtype=np.int32)
        keep["po"] = np.array([0], dtype=np.int32)
        keep["ps"] = np.array([0], dtype=np.int32)
        keep["ops"] = np.ones(1)
        keep["perms"] = np.array([3], dtype=np.int32)
        d.m2l_kind = H.H2G_M2L_SYMDENSE
        d.n_pairs, d.n_ops, d.n_perms = 1, 1, 1
        d.pair_test, d.pair_trial, d.pair_op, d.pair_sym = (H.ptr(keep[k], C.c_int32) for k in ("pt", "pc", "po", "ps"))
        d.ops = H.ptr(keep["ops"])
        d.perms = H.ptr(keep["perms"], C.c_int32)
    h = C.c_void_p()
    rc = H.lib().h2g_core_create(C.byref(d), C.byref(h))
    assert rc == getattr(H, "H2G_" + code), (rc, H.last_error())
    assert h.value is None
    assert H.last_error()
