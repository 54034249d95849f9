"""Diagnostics of GPU maps with the reference's names (``src/diagnostics/``):
``translation_counts``, ``h2summary``, ``sanitycheck`` and the timing records of a product.

They read the library (``h2g_core_summary``, ``h2g_core_sanitycheck``, ``h2g_ws_timing``); the
records keep the reference's fields (``src/diagnostics/timing.jl:11-20``, ``structure.jl:40-310``).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

from . import _lib as H

PHASES = ["gather", "near", "p2m", "m2m", "m2l_fft_in", "m2l_apply", "m2l_fft_out", "l2l", "l2p", "scatter"]
MODES = {H.H2G_OP_N: "forward", H.H2G_OP_T: "transpose", H.H2G_OP_C: "adjoint"}
KIND_NAMES = {0: "dense", 1: "toeplitz", 2: "symdense", 3: "symtoeplitz", 4: "diagonal"}
ISSUES = {1: "no_far_pairs", 2: "self_far_pair", 3: "unresolved_translation", 4: "missing_moment",
          5: "nonuniform_moment_size", 6: "generic_coupling_level"}
SEVERITY = {0: "info", 1: "warning", 2: "error"}


@dataclass
class TranslationCounts:
    farpairs: int
    uniquetranslations: int
    reuse: float


@dataclass
class H2Summary:
    maptype: str
    size: tuple
    eltype: str
    numlevels: int
    numleaves: int
    numnodes: int
    minleafdofs: int
    maxleafdofs: int
    meanleafdofs: float
    farpairs: int
    uniquetranslations: int
    m2l_kind: str
    nearblocks: int
    nearvalues: int
    latticelevels: int
    dense_mfma: bool
    device_bytes: int


@dataclass
class DiagnosticIssue:
    severity: str
    code: str
    count: int
    level: int


@dataclass
class TimingRecord:
    operation: str
    phase: str
    level: object      # int or None (missing)
    node: object       # None: records are per level
    mode: str
    seconds: float


def _core(A):
    """The core of a map, a transpose / adjoint view of one, or a core itself."""
    from .maps import GpuH2Core

    return A if isinstance(A, GpuH2Core) else A._ws().core


def _summary(A) -> H.Summary:
    s = H.Summary()
    H.check(H.lib().h2g_core_summary(_core(A).handle, C.byref(s)))
    return s


def translation_counts(A) -> TranslationCounts:
    """Far pairs, distinct coupling operators (translation id, symmetry) and their reuse."""
    s = _summary(A)
    return TranslationCounts(s.far_pairs, s.unique_translations, s.reuse)


def h2summary(A) -> H2Summary:
    """Map size, test-tree shape, leaf occupancy, far-pair counts (and the GPU's coupling path)."""
    s = _summary(A)
    return H2Summary("GpuH2Map", (s.rows, s.cols), "ComplexF64" if s.dtype == H.H2G_C64 else "Float64", s.n_levels,
                     s.n_leaves, s.n_nodes, s.min_leaf_dofs, s.max_leaf_dofs, s.mean_leaf_dofs, s.far_pairs,
                     s.unique_translations, KIND_NAMES.get(s.m2l_kind, "?"), s.near_blocks, s.near_values,
                     s.lattice_levels, bool(s.dense_mfma), s.device_bytes)


def sanitycheck(A) -> list:
    """Issues of the core's plans and moments (empty list: nothing found)."""
    n = C.c_int32()
    core = _core(A)
    H.check(H.lib().h2g_core_sanitycheck(core.handle, None, 0, C.byref(n)))
    arr = (H.Issue * max(n.value, 1))()
    H.check(H.lib().h2g_core_sanitycheck(core.handle, arr, n.value, C.byref(n)))
    return [DiagnosticIssue(SEVERITY[i.severity], ISSUES.get(i.code, str(i.code)), i.count,
                            None if i.level < 0 else i.level) for i in arr[:n.value]]


def timing_records(ws) -> list:
    """TimingRecords of the last product of a workspace created with ``timing=True``."""
    n = C.c_int32()
    H.check(H.lib().h2g_ws_timing(ws.handle, None, 0, C.byref(n)))
    arr = (H.TimingRecord * max(n.value, 1))()
    H.check(H.lib().h2g_ws_timing(ws.handle, arr, n.value, C.byref(n)))
    return [TimingRecord(PHASES[r.phase], "level" if r.level >= 0 else "total", r.level if r.level >= 0 else None,
                         None if r.node < 0 else r.node, MODES.get(r.mode, "forward"), r.ms * 1e-3)
            for r in arr[:n.value]]
