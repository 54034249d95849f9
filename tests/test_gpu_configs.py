"""The benchmark statement's single-map configurations, end to end on the GPU with maps built on
the device, against the oracle, the C/OpenMP restatement and the O(N^2) direct sum:

* the reference's CPU-runnable case: N = 20,000 uniform in the unit cube, Chebyshev p = 4 (dense
  FP64 couplings), leaves of at most 64 points, one FP64 column;
* N = 1,000,000, equispaced p = 6 with Toeplitz couplings, one column, plain FP64 and FP32
  moments, against the restatement and the direct sum on 1,000 targets."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def direct_sum_gpu(pts, x, rows):
    """y_i = sum_j G(p_i - p_j) x_j (Laplace, 0 on the diagonal) on the given rows, FP64 on the GPU,
    from explicit coordinate differences in blocks of 512 rows x 100,000 columns."""
    import torch

    P = torch.from_numpy(pts).cuda()
    X = torch.from_numpy(np.asarray(x, dtype=np.float64).reshape(len(pts), -1)).cuda()
    out = []
    for r0 in range(0, len(rows), 512):
        T = P[torch.from_numpy(rows[r0:r0 + 512]).cuda()]
        y = torch.zeros((T.shape[0], X.shape[1]), dtype=torch.float64, device="cuda")
        for c0 in range(0, len(pts), 100_000):
            diff = T[:, None, :] - P[None, c0:c0 + 100_000, :]
            d = torch.sqrt((diff * diff).sum(-1))
            G = torch.where(d > 0, 1.0 / (4 * np.pi * torch.where(d > 0, d, torch.ones_like(d))), torch.zeros_like(d))
            y += G @ X[c0:c0 + 100_000]
        out.append(y)
    return torch.cat(out).cpu().numpy()


def test_reference_cpu_case_n20k_chebyshev_p4():
    import h2g
    from oracle import build as OB
    from oracle import h2 as O

    pts = np.random.default_rng(0).random((20_000, 3))
    x = np.random.default_rng(1).standard_normal(20_000)
    core = OB.build_core(pts, p=4, nodekind="cheb", m2l_kind="dense", leaf_size=64)
    arrays = h2g.build_map(pts, p=4, nodekind="cheb", m2l_kind="dense", leaf_size=64, plan="device", ops="device")
    c = h2g.GpuH2Core(arrays, precisions=("plain", "mixed"))
    A = h2g.H2Workspace(c, max_rhs=1)
    for op in "NTC":
        M = {"N": A, "T": h2g.transpose(A), "C": h2g.adjoint(A)}[op]
        y = np.zeros(20_000)
        h2g.mul_(y, M, x)
        e = relerr(y, O.mul(core, x, op).real)
        print(f"N=20k p=4 dense op={op}: rel err vs oracle {e:.2e}", flush=True)
        assert e < 1e-12
    y = A @ x
    yd = direct_sum_gpu(pts, x, np.arange(20_000))[:, 0]
    ed = relerr(y, yd)
    Am = h2g.H2Workspace(c, max_rhs=1, precision="mixed")
    ym = Am @ x
    em = relerr(ym, y)
    print(f"N=20k p=4 dense: plain vs direct sum (all targets) {ed:.2e}, mixed vs plain {em:.2e}", flush=True)
    assert ed < 1e-3 and em < 1e-6


def test_toeplitz_n1m_p6_one_column():
    import h2g
    from oracle.cpu import CpuH2

    pts = np.random.default_rng(0).random((1_000_000, 3))
    x = np.random.default_rng(1).standard_normal(1_000_000)
    arrays = h2g.build_map(pts, p=6, nodekind="equi", m2l_kind="toeplitz", leaf_size=64, plan="device", ops="device")
    c = h2g.GpuH2Core(arrays, precisions=("plain", "mixed"))
    y = h2g.H2Workspace(c, max_rhs=1) @ x
    ym = h2g.H2Workspace(c, max_rhs=1, precision="mixed") @ x
    cpu = CpuH2(arrays)
    yc = cpu.mul(x)
    cpu.close()
    rows = np.random.default_rng(2).choice(1_000_000, 1000, replace=False)
    yd = direct_sum_gpu(pts, x, rows)[:, 0]
    e_cpu, e_mixed = relerr(y, yc), relerr(ym, y)
    e_d, e_dm = relerr(y[rows], yd), relerr(ym[rows], yd)
    print(f"N=1M p=6 Toeplitz: plain vs CPU restatement {e_cpu:.2e}, mixed vs plain {e_mixed:.2e}, "
          f"direct sum (1000 targets): plain {e_d:.2e}, mixed {e_dm:.2e}", flush=True)
    assert e_cpu < 1e-12 and e_mixed < 1e-6 and e_d < 1e-4 and e_dm < 1e-4
