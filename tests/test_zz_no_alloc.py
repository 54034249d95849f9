"""h2g_matvec_noalloc (the product without allocation, buffers in the library's static device
scratch) against h2g_matvec on a workspace, through the C ABI: two calls in a row, the second on a
smaller map with fewer columns, plain (1e-12) and mixed (1e-6 gate) precision."""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _map(n, seed):
    import h2g
    from h2g import method as PM

    pts = np.random.default_rng(seed).random((n, 3))
    arrays = PM.build_map(pts, p=4, nodekind="equi", m2l_kind="toeplitz", leaf_size=48, host_values=False)
    return h2g.GpuH2Core(arrays, precisions=("plain", "mixed"))


def _noalloc(core, precision, op, X, Y):
    import torch

    from h2g import _lib as H

    dt = H.H2G_F64 if X.dtype == torch.float64 else H.H2G_F32
    nrhs = X.shape[1]
    stream = torch.cuda.current_stream().cuda_stream
    rc = H.lib().h2g_matvec_noalloc(core.handle, precision, op, H.H2G_MUL, dt, nrhs, C.c_void_p(X.data_ptr()),
                                    X.stride(1), C.c_void_p(Y.data_ptr()), Y.stride(1), None, None,
                                    C.c_void_p(stream), 0)
    assert rc == H.H2G_OK, H.last_error()


def _colmajor(a, torch):
    return torch.from_numpy(np.ascontiguousarray(a.T)).cuda().t()


def test_noalloc_matches_workspace_product(gpu):
    import torch

    import h2g
    from h2g import _lib as H

    rng = np.random.default_rng(0)
    for n, nrhs, seed in ((6000, 20, 1), (2500, 3, 2)):  # the second call: a smaller input
        core = _map(n, seed)
        for precision, name, tol, dt in ((H.H2G_PREC_PLAIN, "plain", 1e-12, np.float64),
                                         (H.H2G_PREC_MIXED, "mixed", 1e-6, np.float32)):
            ws = h2g.H2Workspace(core, max_rhs=nrhs, precision=name)
            for op in (H.H2G_OP_N, H.H2G_OP_T):
                x = rng.standard_normal((n, nrhs)).astype(dt)
                X = _colmajor(x, torch)
                Y = _colmajor(np.full((n, nrhs), np.nan, dtype=dt), torch)
                _noalloc(core, precision, op, X, Y)
                M = ws if op == H.H2G_OP_N else h2g.transpose(ws)
                ref = M @ x
                y = Y.cpu().numpy()
                assert np.isfinite(y).all(), (n, name, op)
                assert relerr(y, ref) < tol, (n, name, op, relerr(y, ref))
            ws.close()
        core.close()
