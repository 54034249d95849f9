"""Operators of the interpolation method built on the device (``h2g_method_operators``): the
couplings of the translation ids (dense kernel matrices or FP64 Toeplitz kernel spectra), the
eight sector transfers and the symmetry permutation tables, in the layouts of ``method.py``'s
host builders."""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as H

_KINDS = {"dense": H.H2G_M2L_DENSE, "toeplitz": H.H2G_M2L_TOEPLITZ, "symdense": H.H2G_M2L_SYMDENSE,
          "symtoeplitz": H.H2G_M2L_SYMTOEPLITZ}


def device_operators(kernel, wavenumber, p, nodes, root_size, m2l_kind, keys, device=0, want=("ops", "transfers",
                                                                                                 "perms")):
    """(ops, transfers, perms): ops per key (rows (level, ox, oy, oz) of ``keys``) as
    ``method.build_map`` stores them; transfers (8, K, K) holding T_s^T (column-major T_s);
    perms (48, K or L^3) for the symmetric kinds, else None."""
    K, L = p ** 3, 2 * p - 1
    keys = np.ascontiguousarray(keys, dtype=np.int32).reshape(-1, 4)
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    d = H.OpsDesc()
    d.kernel = {"laplace": 0, "helmholtz": 1}[kernel]
    d.wavenumber, d.p, d.nodes, d.root_size = float(wavenumber), int(p), H.ptr(nodes, C.c_double), float(root_size)
    d.m2l_kind, d.n_keys, d.keys, d.device = _KINDS[m2l_kind], keys.shape[0], H.ptr(keys, C.c_int32), int(device)
    toep = "toeplitz" in m2l_kind
    if toep:
        ops = np.empty((keys.shape[0], L, L, L), np.complex128)
    else:
        ops = np.empty((keys.shape[0], K, K), np.complex128 if kernel == "helmholtz" else np.float64)
    tr = np.empty((8, K, K), np.float64)
    perms = None
    if m2l_kind.startswith("sym"):
        perms = np.empty((48, L ** 3 if toep else K), np.int32)
    o = H.OpsOut()
    o.ops = H.ptr(ops) if "ops" in want and keys.shape[0] else None
    o.transfers = H.ptr(tr, C.c_double) if "transfers" in want else None
    o.perms = H.ptr(perms, C.c_int32) if perms is not None and "perms" in want else None
    H.check(H.lib().h2g_method_operators(C.byref(d), C.byref(o)))
    return ops, tr, perms
