"""HIP path vs the oracle on identical core data (FP64 / ComplexF64 drop-in, rtol 1e-12).

Every product runs through libh2g.so; the oracle traverses the same arrays on the CPU.
"""

import numpy as np
import pytest

from oracle import build as OB
from oracle import h2 as O

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def relerr(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


def make_map(core, **ws):
    import h2g
    from convert import to_arrays

    arrays = to_arrays(core)
    c = h2g.GpuH2Core(arrays, precisions=("plain", "mixed"))
    return h2g.H2Workspace(c, **ws)


def orient(A, op):
    import h2g

    return {"N": A, "T": h2g.transpose(A), "C": h2g.adjoint(A)}[op]


RNG = np.random.default_rng(7)
PTS_A = RNG.random((260, 3))
PTS_B = RNG.random((190, 3))


@pytest.mark.parametrize("kind", ["dense", "toeplitz", "symdense", "symtoeplitz", "diagonal"])
@pytest.mark.parametrize("petrov", [False, True])
def test_random_core_all_products(gpu, kind, petrov):
    """Random complex operators: every orientation x product x 3/5-argument form."""
    import h2g

    core = OB.random_core(PTS_A, PTS_B if petrov else None, p=2, m2l_kind=kind, leaf_size=8, seed=11,
                          alpha=0.8 - 0.3j)
    A = make_map(core, max_rhs=2)
    rng = np.random.default_rng(5)
    a, b = 1.3 - 0.4j, -0.6 + 0.2j
    for op in "NTC":
        M = orient(A, op)
        n_out, n_in = M.shape
        for nrhs in (1, 3):
            shp_in = (n_in,) if nrhs == 1 else (n_in, nrhs)
            shp_out = (n_out,) if nrhs == 1 else (n_out, nrhs)
            x = rng.standard_normal(shp_in) + 1j * rng.standard_normal(shp_in)
            y0 = rng.standard_normal(shp_out) + 1j * rng.standard_normal(shp_out)
            for fn, ref in ((h2g.mul_, O.mul), (h2g.farmul_, O.farmul), (h2g.nearmul_, O.nearmul)):
                y = np.zeros(shp_out, dtype=np.complex128)
                fn(y, M, x)
                assert relerr(y, ref(core, x, op)) < RTOL, (op, nrhs, fn.__name__, "3-arg")
                y = y0.copy()
                fn(y, M, x, a, b)
                assert relerr(y, ref(core, x, op, a, b, y0)) < RTOL, (op, nrhs, fn.__name__, "5-arg")


def test_farmul5_leaves_uninitialised_rows(gpu):
    """farmul!(y, A, x, a, b) writes only leaves with a far contribution
    (src/farinteractions/traversalhelpers.jl:103-112)."""
    import h2g

    pts = np.random.default_rng(2).random((70, 3))
    core = OB.random_core(pts, p=2, m2l_kind="dense", leaf_size=8, seed=3)
    _, mask = O.far_apply(core, np.ones(70, dtype=np.complex128), "N")
    assert (~mask).any() and mask.any()
    A = make_map(core)
    x = np.random.default_rng(1).standard_normal(70) + 0j
    y0 = np.arange(70) + 1j
    y = y0.copy()
    h2g.farmul_(y, A, x, 2.0, 3.0)
    assert np.array_equal(y[~mask], y0[~mask])
    assert relerr(y, O.farmul(core, x, "N", 2.0, 3.0, y0)) < RTOL


def test_worked_example(gpu):
    """docs/src/implementing/example.md:205-256: forward, transpose, adjoint, Matrix(A) and
    scaled mul! in all orientations equal the dense kernel matrix to rtol 1e-12."""
    import h2g

    core, s, F = OB.worked_example_core()
    A = make_map(core, max_rhs=4)
    n = len(s)
    j = np.arange(1, n + 1)
    x = np.sin(j) + 1j * np.cos(2 * j)
    assert relerr(A @ x, F @ x) < RTOL
    assert relerr(h2g.transpose(A) @ x, F.T @ x) < RTOL
    assert relerr(h2g.adjoint(A) @ x, F.conj().T @ x) < RTOL
    assert relerr(h2g.Matrix(A), F) < RTOL
    for M, R in ((A, F), (h2g.transpose(A), F.T), (h2g.adjoint(A), F.conj().T)):
        out = np.full(n, 2.0 + 1j)
        expected = (2 - 1j) * (R @ x) + (3 + 2j) * out
        h2g.mul_(out, M, x, 2 - 1j, 3 + 2j)
        assert relerr(out, expected) < RTOL
    B = h2g.similarworkspace(A)
    assert B.core is A.core and B.handle.value != A.handle.value
    assert relerr(B @ x, F @ x) < RTOL


@pytest.mark.parametrize("kind,p,nodes", [("dense", 4, "cheb"), ("toeplitz", 4, "equi"),
                                          ("symtoeplitz", 4, "equi"), ("symdense", 3, "equi")])
def test_laplace_method(gpu, kind, p, nodes):
    """Real Laplace interpolation maps: GPU == oracle to 1e-12, forward and transpose."""
    import h2g

    pts = np.random.default_rng(4).random((3000, 3))
    core = OB.build_core(pts, p=p, nodekind=nodes, m2l_kind=kind, leaf_size=64)
    A = make_map(core, max_rhs=2)
    x = np.random.default_rng(9).standard_normal((3000, 2))
    for op in "NT":
        y = np.zeros((3000, 2))
        h2g.mul_(y, orient(A, op), x)
        assert relerr(y, O.mul(core, x, op).real) < RTOL, op


@pytest.mark.parametrize("kind,p,nodes", [("toeplitz", 6, "equi"), ("dense", 5, "cheb")])
def test_laplace_method_high_order(gpu, kind, p, nodes):
    """p = 5, 6 (separable transfers at their largest tile): plain == oracle to 1e-12, and the
    batched mixed form at 32 / 33 columns within the 1e-6 gate of the plain form."""
    import h2g

    pts = np.random.default_rng(14).random((6000, 3))
    core = OB.build_core(pts, p=p, nodekind=nodes, m2l_kind=kind, leaf_size=64)
    A = make_map(core, max_rhs=2)
    x = np.random.default_rng(15).standard_normal((6000, 2))
    for op in "NT":
        y = np.zeros((6000, 2))
        h2g.mul_(y, orient(A, op), x)
        assert relerr(y, O.mul(core, x, op).real) < RTOL, op
    from convert import to_arrays

    c = h2g.GpuH2Core(to_arrays(core), precisions=("plain", "mixed"))
    plain = h2g.H2Workspace(c, max_rhs=33, precision="plain")
    mixed = h2g.H2Workspace(c, max_rhs=33, precision="mixed", use_graph=True)
    X = np.random.default_rng(16).standard_normal((6000, 33))
    for nrhs in (32, 33):
        for op in "NT":
            y64 = np.zeros((6000, nrhs))
            h2g.mul_(y64, orient(plain, op), X[:, :nrhs])
            y32 = np.zeros((6000, nrhs), dtype=np.float32)
            h2g.mul_(y32, orient(mixed, op), X[:, :nrhs].astype(np.float32))
            assert relerr(y32, y64) < 1e-6, (op, nrhs)


def _tile_case():
    pts = np.random.default_rng(31).random((12000, 3))
    core = OB.build_core(pts, p=4, nodekind="equi", m2l_kind="toeplitz", leaf_size=32)
    return pts, core


@pytest.mark.parametrize("nrhs", [32, 40])
def test_tile_shapes_do_not_change_the_product(gpu, nrhs):
    """The workspace tile shapes (coupling-kernel groups per frequency, near-field columns per
    workgroup) change only the schedule: every shape gives the default shape's mixed product to
    the FP32 rounding level, and the plain FP64 product within the 1e-6 gate."""
    import h2g
    from convert import to_arrays

    pts, core = _tile_case()
    c = h2g.GpuH2Core(to_arrays(core), precisions=("plain", "mixed"))
    from h2g.diagnostics import h2summary

    assert h2summary(c).latticelevels > 0
    X = np.random.default_rng(32).standard_normal((12000, nrhs)).astype(np.float32)
    plain = h2g.H2Workspace(c, max_rhs=nrhs, precision="plain")
    for op in "NT":
        y64 = np.zeros((12000, nrhs))
        h2g.mul_(y64, orient(plain, op), X.astype(np.float64))
        ref = None
        for groups, cols in ((0, 0), (3, 64), (64, 16), (1024, 32), (100000, 0)):
            A = h2g.H2Workspace(c, max_rhs=nrhs, precision="mixed", use_graph=True, m2l_groups=groups,
                                near_cols=cols)
            y = np.zeros((12000, nrhs), dtype=np.float32)
            h2g.mul_(y, orient(A, op), X)
            A.close()
            assert relerr(y, y64) < 1e-6, (op, groups, cols)
            if ref is None:
                ref = y
            assert relerr(y, ref) < 1e-6, (op, groups, cols)


@pytest.mark.parametrize("nrhs", [4, 16, 32, 40])
def test_persistent_coupling_stage_equals_per_level(gpu, nrhs):
    """The coupling stage of all lattice levels in one persistent launch computes the per-level
    launches' products (the same task arithmetic, other workgroups): equal to FP32 rounding, in
    every orientation, and within the 1e-6 gate of the plain FP64 product."""
    import h2g
    from convert import to_arrays

    pts, core = _tile_case()
    c = h2g.GpuH2Core(to_arrays(core), precisions=("plain", "mixed"))
    X = np.random.default_rng(34).standard_normal((12000, nrhs)).astype(np.float32)
    plain = h2g.H2Workspace(c, max_rhs=nrhs, precision="plain")
    for op in "NTC":
        y64 = np.zeros((12000, nrhs))
        h2g.mul_(y64, orient(plain, op), X.astype(np.float64))
        ys = []
        for persistent, groups in ((0, 0), (1, 0), (1, 64)):
            A = h2g.H2Workspace(c, max_rhs=nrhs, precision="mixed", use_graph=True, persistent=persistent,
                                m2l_groups=groups)
            y = np.zeros((12000, nrhs), dtype=np.float32)
            for _ in range(2):  # the second replay starts from counters the first one left behind
                h2g.mul_(y, orient(A, op), X)
            A.close()
            ys.append(y)
            assert relerr(y, y64) < 1e-6, (op, persistent, groups)
        # each block's sums run in the same order in either launch form and any grouping
        assert np.array_equal(ys[1], ys[0]) and np.array_equal(ys[2], ys[0]), op


def test_tile_shape_arguments_are_checked(gpu):
    import h2g
    from convert import to_arrays

    _, core = _tile_case()
    c = h2g.GpuH2Core(to_arrays(core), precisions=("plain", "mixed"))
    for kw in ({"near_cols": 24}, {"m2l_groups": -1}, {"persistent": 2}, {"dense_pairs": 2}):
        with pytest.raises(h2g.ArgumentError):
            h2g.H2Workspace(c, max_rhs=8, precision="mixed", **kw)


def test_autotune_picks_a_measured_shape(gpu, tmp_path):
    """tune() times every candidate of each axis, returns the fastest, caches it per map, and the
    tuned workspace computes the default workspace's product."""
    import h2g
    from convert import to_arrays
    from h2g.autotune import tune, tuned_workspace

    _, core = _tile_case()
    c = h2g.GpuH2Core(to_arrays(core), precisions=("plain", "mixed"))
    cache = str(tmp_path / "tune.json")
    best, table = tune(c, nrhs=32, precision="mixed", blocks=(16, 32), m2l_groups=(128, 512), near_cols=(16, 32),
                       repeats=1, cache_path=cache)
    assert best["max_rhs"] in (16, 32) and best["m2l_groups"] in (128, 512) and best["near_cols"] in (16, 32)
    tried = {(t["max_rhs"], t["m2l_groups"], t["near_cols"]) for t, _ in table}
    assert (best["max_rhs"], best["m2l_groups"], best["near_cols"]) in tried
    assert all(ms > 0 for _, ms in table) and len(table) >= 5
    best2, table2 = tune(c, nrhs=32, precision="mixed", cache_path=cache)  # served from the cache
    assert best2 == best and len(table2) == len(table)
    X = np.random.default_rng(33).standard_normal((12000, 32)).astype(np.float32)
    y0 = np.zeros((12000, 32), dtype=np.float32)
    y1 = np.zeros((12000, 32), dtype=np.float32)
    h2g.mul_(y0, h2g.H2Workspace(c, max_rhs=32, precision="mixed"), X)
    T = tuned_workspace(c, nrhs=32, precision="mixed", cache_path=cache)
    h2g.mul_(y1, T, X)
    assert relerr(y1, y0) < 1e-6


@pytest.mark.skip(reason="faulted the GPU twice (memory aperture violation) with the separable-transfer kernel "
                         "for ComplexF64; that path is switched off for complex maps and this case stays off the "
                         "GPU until the cause is shown")
def test_helmholtz_complex(gpu):
    import h2g

    pts = np.random.default_rng(8).standard_normal((1500, 3))
    pts = 0.5 + 0.45 * pts / np.linalg.norm(pts, axis=1, keepdims=True)
    core = OB.build_core(pts, p=3, nodekind="cheb", m2l_kind="dense", kernel="helmholtz", wavenumber=10.0,
                         leaf_size=32)
    A = make_map(core)
    x = np.random.default_rng(2).standard_normal(1500) + 1j * np.random.default_rng(3).standard_normal(1500)
    for op in "NTC":
        y = np.zeros(1500, dtype=np.complex128)
        h2g.mul_(y, orient(A, op), x)
        assert relerr(y, O.mul(core, x, op)) < RTOL, op


def test_busy_flag(gpu):
    """A held workspace rejects products (ConcurrentWorkspaceError), views share the flag."""
    import h2g

    core, s, F = OB.worked_example_core()
    A = make_map(core)
    x = np.ones(len(s), dtype=np.complex128)
    A.acquire()
    with pytest.raises(h2g.ConcurrentWorkspaceError):
        A @ x
    with pytest.raises(h2g.ConcurrentWorkspaceError):
        h2g.transpose(A) @ x
    A.release()
    assert relerr(A @ x, F @ x) < RTOL


def test_graph_replays_match_eager(gpu):
    """hipGraph-captured passes replayed several times, eager work in between, equal eager."""
    import h2g

    core = OB.random_core(PTS_A, p=2, m2l_kind="toeplitz", leaf_size=8, seed=2)
    A = make_map(core, max_rhs=3)
    G = make_map(core, max_rhs=3, use_graph=True)
    rng = np.random.default_rng(0)
    for it in range(4):
        x = rng.standard_normal((260, 3)) + 1j * rng.standard_normal((260, 3))
        ye, yg = A @ x, G @ x
        assert relerr(yg, ye) < 1e-15, it
        assert relerr(h2g.adjoint(G) @ x, h2g.adjoint(A) @ x) < 1e-15, it


def test_device_tensors(gpu):
    """torch tensors on the GPU go through the same entry point in place."""
    import torch

    import h2g

    core, s, F = OB.worked_example_core()
    A = make_map(core, max_rhs=2)
    x = np.random.default_rng(0).standard_normal((32, 2)) + 0j
    xt = torch.tensor(x, device="cuda").t().contiguous().t()
    yt = torch.zeros((32, 2), dtype=torch.complex128, device="cuda").t().contiguous().t()
    h2g.mul_(yt, A, xt)
    assert relerr(yt.cpu().numpy(), F @ x) < RTOL


@pytest.mark.parametrize("kind", ["toeplitz", "symtoeplitz"])
@pytest.mark.parametrize("real", [True, False])
def test_lattice_toeplitz_matches_generic(gpu, kind, real):
    """The lattice-blocked Toeplitz kernels (frequency-major spectra, half spectra for real maps)
    against the generic per-pair kernels and the oracle, every orientation, 1..33 RHS."""
    import h2g
    from convert import to_arrays

    pts = np.random.default_rng(12).random((4000, 3))
    if real:
        core = OB.build_core(pts, p=4, nodekind="equi", m2l_kind=kind, leaf_size=32)
    else:
        core = OB.random_core(pts[:900], p=2, m2l_kind=kind, leaf_size=8, seed=4)
    arrays = to_arrays(core)
    lat = h2g.H2Workspace(h2g.GpuH2Core(arrays, precisions=("plain", "mixed")), max_rhs=33)
    gen = h2g.H2Workspace(h2g.GpuH2Core(arrays, precisions=("plain",), lattice=False), max_rhs=33)
    n = arrays.test.tree.points.shape[0]
    rng = np.random.default_rng(3)
    for nrhs in (1, 5, 33):
        x = rng.standard_normal((n, nrhs))
        if not real:
            x = x + 1j * rng.standard_normal((n, nrhs))
        for op in "NTC":
            yl = orient(lat, op) @ x
            yg = orient(gen, op) @ x
            ref = O.mul(core, x, op)
            ref = ref.real if real else ref
            assert relerr(yl, ref) < RTOL, (op, nrhs)
            assert relerr(yl, yg) < RTOL, (op, nrhs)
    # mixed form: narrow batches (1, 3, 8 columns) and wide ones (16, 20, 40: per-block slot
    # lists; 40 takes a second, partly filled chunk of 32 columns)
    mixed = h2g.H2Workspace(lat.core, max_rhs=40, precision="mixed")
    for nrhs in (1, 3, 8, 16, 20, 40):
        x = rng.standard_normal((n, nrhs)) + (0 if real else 1j * rng.standard_normal((n, nrhs)))
        for op in "NTC":
            ref = O.mul(core, x, op)
            assert relerr(orient(mixed, op) @ x, ref.real if real else ref) < 1e-6, (op, nrhs)


def test_mfma_f64_lane_layout(gpu):
    """The lane layout the FP64 MFMA near kernel assumes for v_mfma_f64_16x16x4f64."""
    import ctypes as C

    from h2g import _lib as H

    rng = np.random.default_rng(0)
    A = rng.standard_normal((16, 4))
    B = rng.standard_normal((4, 16))
    D = np.zeros((16, 16))
    lib = H.lib()
    lib.h2g_probe_mfma_f64.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    assert lib.h2g_probe_mfma_f64(H.ptr(A), H.ptr(B), H.ptr(D)) == 0
    assert np.allclose(D, A @ B, rtol=1e-14, atol=1e-14)


def test_mfma_f32_lane_layout(gpu):
    """The lane layout the list M2L kernel assumes for v_mfma_f32_16x16x4f32 (result rows
    4 (lane / 16) + element); small integers, so every product and sum is exact."""
    import ctypes as C

    from h2g import _lib as H

    rng = np.random.default_rng(1)
    A = rng.integers(-4, 5, (16, 4)).astype(np.float32)
    B = rng.integers(-4, 5, (4, 16)).astype(np.float32)
    D = np.zeros((16, 16), dtype=np.float32)
    lib = H.lib()
    lib.h2g_probe_mfma_f32.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    assert lib.h2g_probe_mfma_f32(H.ptr(A), H.ptr(B), H.ptr(D)) == 0
    assert np.array_equal(D, A @ B)


@pytest.mark.parametrize("petrov,leaf", [(False, 40), (True, 70)])
def test_wide_near_field_all_orientations(gpu, petrov, leaf):
    """Near field alone at 4..70 RHS (FP64 MFMA path for real maps) against the oracle.

    Leaves of up to 70 points give output leaves of 1..4 row tiles and leaves split into two
    row-block tasks; distinct test and trial points give rectangular blocks, whose transposes
    read rows with a stride other than the column count.  70 RHS take two column groups."""
    import h2g

    rng = np.random.default_rng(21)
    pts = rng.random((2500, 3))
    pts_b = rng.random((2100, 3)) if petrov else None
    core = OB.build_core(pts, pts_b, p=3, nodekind="cheb", m2l_kind="dense", leaf_size=leaf)
    A = make_map(core, max_rhs=70)
    for nrhs in (4, 7, 16, 21, 40, 70):
        for op in "NT":
            M = orient(A, op)
            n_out, n_in = M.shape
            x = rng.standard_normal((n_in, nrhs))
            y = np.zeros((n_out, nrhs))
            h2g.nearmul_(y, M, x)
            assert relerr(y, O.nearmul(core, x, op)) < RTOL, (op, nrhs)


@pytest.mark.parametrize("petrov,leaf", [(False, 40), (True, 70)])
def test_wide_near_field_mixed(gpu, petrov, leaf):
    """Near field alone in the mixed form at 4..70 RHS: the FP32 copy of the values (rounded on
    the device after the upload) on FP64 MFMA, against the oracle on the same FP32 inputs.
    Rounding the values to FP32 bounds the error by about 1e-7; reading the FP64 array as FP32
    (or the reverse) or a missing copy would give errors of order one."""
    import h2g

    rng = np.random.default_rng(22)
    pts = rng.random((2500, 3))
    pts_b = rng.random((2100, 3)) if petrov else None
    core = OB.build_core(pts, pts_b, p=3, nodekind="cheb", m2l_kind="dense", leaf_size=leaf)
    A = make_map(core, max_rhs=70, precision="mixed")
    for nrhs in (4, 7, 16, 21, 40, 70):
        for op in "NT":
            M = orient(A, op)
            n_out, n_in = M.shape
            x = rng.standard_normal((n_in, nrhs)).astype(np.float32)
            y = np.zeros((n_out, nrhs), dtype=np.float32)
            h2g.nearmul_(y, M, x)
            assert relerr(y, O.nearmul(core, x.astype(np.float64), op)) < 1e-6, (op, nrhs)


def test_reference_farmul_fixture(gpu):
    """test/farinteractions/test_farmul.jl:11-97 through the C ABI: the hand-derived far matrix
    (c = 2 - i between the two clusters) in every orientation, and the 5-argument form."""
    import h2g

    core, _, F = OB.farmul_fixture_core()
    A = make_map(core, max_rhs=2)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(8) + 1j * rng.standard_normal(8)
    for M, R in ((A, F), (h2g.transpose(A), F.T), (h2g.adjoint(A), F.conj().T)):
        y = np.zeros(8, dtype=np.complex128)
        h2g.farmul_(y, M, x)
        assert relerr(y, R @ x) < RTOL
    y0 = rng.standard_normal(8) + 1j * rng.standard_normal(8)
    a, b = 0.7 + 0.3j, -0.4 + 0.2j
    y = y0.copy()
    h2g.farmul_(y, A, x, a, b)
    assert relerr(y, a * F @ x + b * y0) < RTOL


def test_farmul5_without_far_pairs_writes_no_row(gpu):
    """No far pair: the 5-argument farmul! leaves y bit for bit (traversalhelpers.jl:103-112)."""
    import h2g

    core = OB.no_far_core()
    A = make_map(core)
    y0 = np.arange(5) + 1j
    y = y0.copy()
    h2g.farmul_(y, A, np.ones(5) + 0j, 2.0, 3.0)
    assert np.array_equal(y, y0)


@pytest.mark.parametrize("kind,p,nodes,petrov", [("dense", 4, "cheb", False), ("toeplitz", 6, "equi", False),
                                                 ("toeplitz", 4, "equi", True)])
def test_device_filled_map_matches_oracle(gpu, kind, p, nodes, petrov):
    """The build path the benchmark takes: near values and leaf collections filled on the device
    (h2g_method_fill, FP64 and their FP32 copies) against the oracle's host assembly on the same
    points -- plain form to 1e-12 (near field alone and the full product, forward and transpose),
    mixed form within the 1e-6 gate."""
    import h2g
    from h2g import method as PM

    rng = np.random.default_rng(31)
    pts = rng.random((6000, 3))
    pts_b = rng.random((4500, 3)) if petrov else None
    core = OB.build_core(pts, pts_b, p=p, nodekind=nodes, m2l_kind=kind, leaf_size=64)
    arrays = PM.build_map(pts, pts_b, p=p, nodekind=nodes, m2l_kind=kind, leaf_size=64, host_values=False)
    assert arrays.near_values is None and arrays.test.collection is None
    c = h2g.GpuH2Core(arrays, precisions=("plain", "mixed"))
    A = h2g.H2Workspace(c, max_rhs=4)
    Am = h2g.H2Workspace(c, max_rhs=32, precision="mixed")
    for op in "NT":
        n_out, n_in = orient(A, op).shape
        x = rng.standard_normal((n_in, 4))
        y = np.zeros((n_out, 4))
        h2g.nearmul_(y, orient(A, op), x)
        assert relerr(y, O.nearmul(core, x, op)) < RTOL, (op, "near")
        y = np.zeros((n_out, 4))
        h2g.mul_(y, orient(A, op), x)
        assert relerr(y, O.mul(core, x, op).real) < RTOL, (op, "mul")
        X = rng.standard_normal((n_in, 32)).astype(np.float32)
        ym = orient(Am, op) @ X
        assert relerr(ym, O.mul(core, X.astype(np.float64), op).real) < 1e-6, (op, "mixed")


def test_helmholtz_device_near_field(gpu):
    """Helmholtz k = 10 on the sphere: the near field filled on the device (ComplexF64 kernel
    values) and applied alone, every orientation, against the oracle (no transfer kernel runs)."""
    import h2g
    from h2g import method as PM

    pts = np.random.default_rng(8).standard_normal((3000, 3))
    pts = 0.5 + 0.45 * pts / np.linalg.norm(pts, axis=1, keepdims=True)
    core = OB.build_core(pts, p=3, nodekind="cheb", m2l_kind="dense", kernel="helmholtz", wavenumber=10.0,
                         leaf_size=48)
    arrays = PM.build_map(pts, p=3, nodekind="cheb", m2l_kind="dense", kernel="helmholtz", wavenumber=10.0,
                          leaf_size=48, host_values=False)
    A = h2g.H2Workspace(h2g.GpuH2Core(arrays), max_rhs=3)
    rng = np.random.default_rng(5)
    for nrhs in (1, 3):
        x = rng.standard_normal((3000, nrhs)) + 1j * rng.standard_normal((3000, nrhs))
        for op in "NTC":
            y = np.zeros((3000, nrhs), dtype=np.complex128)
            h2g.nearmul_(y, orient(A, op), x)
            assert relerr(y, O.nearmul(core, x, op)) < RTOL, (op, nrhs)


def test_mfma_bf16_lane_layout(gpu):
    """The lane layout the split-BF16 dense M2L kernel assumes for v_mfma_f32_16x16x32_bf16 (lane
    l holds A[l % 16][8 (l / 16) + u], B[8 (l / 16) + u][l % 16]; D[m][n] in lane n + 16 (m / 4),
    element m % 4); small integers, exact in BF16 and in the FP32 sums."""
    import ctypes as C

    from h2g import _lib as H

    rng = np.random.default_rng(2)
    A = rng.integers(-4, 5, (16, 32)).astype(np.float32)
    B = rng.integers(-4, 5, (32, 16)).astype(np.float32)
    D = np.zeros((16, 16), dtype=np.float32)
    lib = H.lib()
    lib.h2g_probe_mfma_bf16.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    assert lib.h2g_probe_mfma_bf16(H.ptr(A), H.ptr(B), H.ptr(D)) == 0
    assert np.array_equal(D, A @ B)


@pytest.mark.parametrize("case", ["laplace", "helmholtz"])
def test_split_bf16_dense_m2l(gpu, case):
    """Dense couplings on split-BF16 MFMA against the plain FP64 form: hi + mid + lo (six plane
    products) within the 1e-6 gate; hi + lo (three products) within its error budget of 1e-4
    (each operand kept to about 2^-17 relative, the dropped products about 2^-16)."""
    import h2g
    from convert import to_arrays

    rng = np.random.default_rng(41)
    if case == "laplace":
        pts = rng.random((6000, 3))
        core = OB.build_core(pts, p=5, nodekind="cheb", m2l_kind="dense", leaf_size=64)
        x = rng.standard_normal((6000, 20))
    else:
        pts = rng.standard_normal((5000, 3))
        pts = 0.5 + 0.45 * pts / np.linalg.norm(pts, axis=1, keepdims=True)
        core = OB.build_core(pts, p=4, nodekind="cheb", m2l_kind="dense",
                             kernel="helmholtz", wavenumber=10.0, leaf_size=48)
        x = rng.standard_normal((5000, 20)) + 1j * rng.standard_normal((5000, 20))
    c = h2g.GpuH2Core(to_arrays(core), precisions=("plain", "mixed", "bf16x2", "bf16x3"))
    plain = h2g.H2Workspace(c, max_rhs=20)
    errs = {}
    for prec, tol in (("mixed", 1e-6), ("bf16x3", 1e-6), ("bf16x2", 1e-4)):
        ws = h2g.H2Workspace(c, max_rhs=20, precision=prec)
        for op in ("NTC" if case == "helmholtz" else "NT"):
            xx = x.astype(np.complex64 if case == "helmholtz" else np.float32)
            y = orient(ws, op) @ xx
            ref = orient(plain, op) @ x.astype(xx.dtype).astype(np.complex128 if case == "helmholtz" else np.float64)
            errs[(prec, op)] = relerr(y, ref)
            assert errs[(prec, op)] < tol, (prec, op, errs[(prec, op)])
        ws.close()
    print("split-BF16 errors vs plain:", {k: f"{v:.2e}" for k, v in errs.items()})


def test_helmholtz_unit_sphere_p6(gpu):
    """The Helmholtz workload's shape at small N: points on the unit sphere (octree on [-1, 1]^3),
    k = 10, Chebyshev p = 6, dense complex couplings, 16 right-hand sides.  Plain ComplexF64 ==
    oracle to 1e-12 in every orientation; ComplexF32 and split BF16 (hi + mid + lo) within the 1e-6
    gate of the plain form, forward and adjoint (device-filled near values and collections)."""
    import h2g
    from h2g import method as PM

    rng = np.random.default_rng(44)
    pts = rng.standard_normal((4000, 3))
    pts /= np.linalg.norm(pts, axis=1, keepdims=True)
    kw = dict(p=6, nodekind="cheb", m2l_kind="dense", kernel="helmholtz", wavenumber=10.0, leaf_size=64,
              origin=(-1.0, -1.0, -1.0), size=2.0)
    core = OB.build_core(pts, **kw)
    arrays = PM.build_map(pts, host_values=False, **kw)
    c = h2g.GpuH2Core(arrays, precisions=("plain", "mixed", "bf16x3"))
    plain = h2g.H2Workspace(c, max_rhs=16)
    x = rng.standard_normal((4000, 16)) + 1j * rng.standard_normal((4000, 16))
    for op in "NTC":
        y = orient(plain, op) @ x
        assert relerr(y, O.mul(core, x, op)) < RTOL, op
    x32 = x.astype(np.complex64)
    for prec in ("mixed", "bf16x3"):
        ws = h2g.H2Workspace(c, max_rhs=16, precision=prec, use_graph=True)
        for op in "NC":
            ref = orient(plain, op) @ x32.astype(np.complex128)
            assert relerr(orient(ws, op) @ x32, ref) < 1e-6, (prec, op)
        ws.close()


def test_diagnostics_and_timing_records(gpu):
    """h2summary on the lattice and dense-MFMA paths; timing records carry the orientation."""
    import h2g
    from convert import to_arrays
    from h2g import diagnostics as D

    pts = np.random.default_rng(30).random((4000, 3))
    lat = OB.build_core(pts, p=4, nodekind="equi", m2l_kind="toeplitz", leaf_size=32)
    c = h2g.GpuH2Core(to_arrays(lat), precisions=("plain", "mixed"))
    s = D.h2summary(c)
    assert s.latticelevels > 0 and s.farpairs == len(lat.pairs) and s.m2l_kind == "toeplitz"
    assert D.sanitycheck(c) == []
    ws = h2g.H2Workspace(c, max_rhs=2, timing=True)
    x = np.random.default_rng(1).standard_normal((4000, 2))
    h2g.adjoint(ws) @ x
    recs = D.timing_records(ws)
    assert recs and all(r.mode == "adjoint" for r in recs)
    assert {"near", "m2l_apply", "l2p"} <= {r.operation for r in recs}
    dense = OB.build_core(pts[:2000], p=3, nodekind="cheb", m2l_kind="dense", leaf_size=32)
    assert D.h2summary(h2g.GpuH2Core(to_arrays(dense))).dense_mfma


@pytest.mark.parametrize("kind,nodes,p", [("dense", "cheb", 3), ("symdense", "equi", 3), ("dense", "equi", 6)])
@pytest.mark.parametrize("petrov", [False, True])
@pytest.mark.parametrize("coords", [True, False])
def test_dense_receiver_pairs(gpu, kind, nodes, p, petrov, coords):
    """Dense couplings of real maps on receiving boxes two by two (``dense_pairs=1``: lattice
    neighbours paired with tree coordinates, consecutive boxes without), and one by one
    (``dense_pairs=-1``), against the oracle: plain to 1e-12 (3- and 5-argument), mixed within the
    1e-6 gate, at 1, 5 and 17 columns (two 16-column chunks per pair), forward and transpose.  p = 3
    runs 4 waves of one row tile, p = 6 8 waves of two."""
    import h2g
    from convert import to_arrays

    rng = np.random.default_rng(21)
    a = rng.random((2500, 3))
    b = rng.random((1800, 3)) if petrov else None
    core = OB.build_core(a, b, p=p, nodekind=nodes, m2l_kind=kind, leaf_size=48)
    c = h2g.GpuH2Core(to_arrays(core), precisions=("plain", "mixed"), lattice=coords)
    s1, s2 = 1.7, -0.4
    for R in (1, 5, 17):
        for op in "NT":
            n_out, n_in = (c.shape if op == "N" else c.shape[::-1])
            x = rng.standard_normal((n_in, R))
            y0 = rng.standard_normal((n_out, R))
            ref = O.mul(core, x, op).real
            ref5 = O.mul(core, x, op, s1, s2, y0).real
            for prec, tol in (("plain", 1e-12), ("mixed", 1e-6)):
                for pairs in (1, -1):
                    W = h2g.H2Workspace(c, max_rhs=R, precision=prec, dense_pairs=pairs)
                    M = orient(W, op)
                    xx = x if prec == "plain" else x.astype(np.float32)
                    y = M @ xx
                    err = relerr(y, ref)
                    assert err < tol, (R, op, prec, pairs, err)
                    if prec == "plain":
                        y = y0.copy()
                        h2g.mul_(y, M, x, s1, s2)
                        assert relerr(y, ref5) < tol, (R, op, "5-arg", pairs)
                    W.close()
