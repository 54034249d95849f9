"""The C/OpenMP restatement (oracle/h2cpu.c) against the NumPy oracle on the same points: the
CPU baseline computes the same product (forward, FP64, 1e-12)."""

import numpy as np
import pytest

from oracle import build as OB
from oracle import h2 as O


@pytest.mark.parametrize("n,p,leaf", [(3000, 4, 64), (5000, 6, 64), (2500, 3, 20)])
def test_c_restatement_matches_numpy_oracle(n, p, leaf):
    from h2g import method as PM
    from oracle.cpu import CpuH2

    pts = np.random.default_rng(n).random((n, 3))
    x = np.random.default_rng(1).standard_normal(n)
    arrays = PM.build_map(pts, p=p, nodekind="equi", m2l_kind="toeplitz", leaf_size=leaf)
    cpu = CpuH2(arrays)
    y = cpu.mul(x)
    ref = O.mul(OB.build_core(pts, p=p, nodekind="equi", m2l_kind="toeplitz", leaf_size=leaf), x).real
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-12
    cpu.close()
