"""Serving loop: admission against the HBM budget, continuous batching, and each request's
result the same alone and in the mix."""

import numpy as np
import pytest

from oracle import h2 as O


def _maps():
    from h2g import method as PM

    rng = np.random.default_rng(0)
    cube = rng.random((3000, 3))
    sph = rng.standard_normal((1600, 3))
    sph = 0.5 + 0.45 * sph / np.linalg.norm(sph, axis=1, keepdims=True)
    lap = PM.build_map(cube, p=4, nodekind="equi", m2l_kind="toeplitz", leaf_size=32, host_values=True)
    hel = PM.build_map(sph, p=3, nodekind="cheb", m2l_kind="dense", kernel="helmholtz", wavenumber=2.0,
                       leaf_size=32, host_values=True)
    return lap, hel


def test_admission_rejects_over_budget():
    """A map that does not fit the budget is refused before anything is built."""
    from h2g.serve import AdmissionError, Server

    lap, _ = _maps()
    srv = Server(budget_bytes=1_000_000)
    with pytest.raises(AdmissionError):
        srv.add_map("laplace", lap)
    assert not srv.maps


def test_admission_falls_back_to_one_workspace_then_narrower_batches():
    """When the preferred workspaces do not fit beside the core, a map gets one workspace of the
    full batch width, then one of half the width; below that it is refused."""
    from h2g.serve import Server, workspace_bytes

    lap, hel = _maps()
    for arrays in (lap, hel):
        w32, w16 = workspace_bytes(arrays, 32), workspace_bytes(arrays, 16)
        assert w16 < w32
        assert Server(budget_bytes=3 * w32, max_batch=32)._workspace_plan(arrays, 0) == (2, 32)
        assert Server(budget_bytes=int(1.5 * w32), max_batch=32)._workspace_plan(arrays, 0) == (1, 32)
        assert Server(budget_bytes=(w16 + w32) // 2, max_batch=32)._workspace_plan(arrays, 0) == (1, 16)
        assert Server(budget_bytes=w16 - 1, max_batch=32)._workspace_plan(arrays, 0) is None


@pytest.mark.gpu
def test_workspace_estimate_bounds_the_library(gpu):
    """The admission estimate of a workspace (host-pointer staging included) is at least what the
    library allocates and within a small factor of it."""
    import h2g
    from h2g import _lib as H
    from h2g.serve import workspace_bytes

    for arrays in _maps():
        core = h2g.GpuH2Core(arrays, precisions=("plain", "mixed"))
        cplx = arrays.dtype == "c64"
        for nb in (1, 8, 32):
            ws = h2g.H2Workspace(core, max_rhs=nb, precision="mixed")
            n = max(core.shape)
            dt = np.complex64 if cplx else np.float32
            x = np.zeros((n, nb), dtype=dt, order="F")
            y = np.zeros((n, nb), dtype=dt, order="F")
            code = H.H2G_C32 if cplx else H.H2G_F32
            H.check(H.lib().h2g_matvec(ws.handle, H.H2G_OP_N, H.H2G_MUL, code, nb, H.ptr(x), n, H.ptr(y), n, None,
                                       None, H.H2G_HOST_PTRS))
            actual, est = ws.device_bytes(), workspace_bytes(arrays, nb)
            print(f"{arrays.dtype} {nb} columns: library {actual} B, estimate {est} B", flush=True)
            assert actual <= est <= 2 * actual, (arrays.dtype, nb, actual, est)
            ws.close()
        core.close()


@pytest.mark.gpu
def test_continuous_batching_same_alone_and_in_mix(gpu):
    import h2g
    from h2g.serve import Request, Server

    lap, hel = _maps()
    # four workspaces on a pool of two streams: completion is told apart per workspace
    srv = Server(budget_bytes=50 * 2 ** 30, max_batch=8, ws_per_map=2, n_streams=2)
    srv.add_map("lap", lap)
    srv.add_map("hel", hel)
    rng = np.random.default_rng(1)
    reqs = []
    for i in range(60):
        name = "lap" if i % 3 else "hel"
        op = "NTC"[i % 3] if name == "hel" else "NT"[i % 2]
        n = 3000 if name == "lap" else 1600
        x = rng.standard_normal(n).astype(np.float32)
        if name == "hel":
            x = (x + 1j * rng.standard_normal(n)).astype(np.complex64)
        reqs.append(Request(name, op, x))
    stats = srv.run_stream(reqs, rate_hz=20000.0, seed=3)
    assert stats["requests"] == 60 and all(r.result is not None for r in reqs)
    assert stats["mean_batch"] > 1.0  # requests were batched
    # alone: the same request as a one-column product on a fresh workspace of the same map
    alone = {name: h2g.H2Workspace(srv.maps[name].core, max_rhs=1, precision="mixed") for name in ("lap", "hel")}
    for r in reqs:
        A = alone[r.map_name]
        M = {"N": A, "T": h2g.transpose(A), "C": h2g.adjoint(A)}[r.op]
        y = M @ r.x
        assert np.linalg.norm(r.result - y) <= 1e-6 * np.linalg.norm(y), (r.map_name, r.op, r.batch_size)
    for A in alone.values():
        A.close()
    # and every served result against the FP64 oracle on the same points, within the 1e-6 gate
    ref = _oracle_cores()
    for r in reqs:
        y64 = O.mul(ref[r.map_name], r.x.astype(np.complex128 if r.map_name == "hel" else np.float64), r.op)
        y64 = y64 if r.map_name == "hel" else y64.real
        assert np.linalg.norm(r.result - y64) <= 1e-6 * np.linalg.norm(y64), (r.map_name, r.op, r.batch_size)
    srv.close()
    assert not srv.maps


def _oracle_cores():
    from oracle import build as OB

    rng = np.random.default_rng(0)
    cube = rng.random((3000, 3))
    sph = rng.standard_normal((1600, 3))
    sph = 0.5 + 0.45 * sph / np.linalg.norm(sph, axis=1, keepdims=True)
    return {"lap": OB.build_core(cube, p=4, nodekind="equi", m2l_kind="toeplitz", leaf_size=32),
            "hel": OB.build_core(sph, p=3, nodekind="cheb", m2l_kind="dense", kernel="helmholtz", wavenumber=2.0,
                                 leaf_size=32)}
