"""Autotuner for the batched product: picks the RHS block (columns per pass) of a workspace.

A multi-RHS product over ``nrhs`` columns runs as ceil(nrhs / block) passes of the whole
hierarchy.  Wider blocks amortise every read of the near field, the collections and the
coupling spectra over more columns and give the per-frequency Hadamard more lanes; narrower
blocks need less workspace memory (moments and spectra scale with the block) and keep the
spectra slices of one frequency L2-resident.  The best block depends on the map (N, p, tree
depth, near-field density), so it is measured: each candidate block runs the same product on
synthetic columns, timed on the device, and the fastest per column wins.  Results are cached
per map signature in a JSON file, so a serving process tunes a map once.
"""

from __future__ import annotations

import hashlib
import json
import os
import time

import numpy as np

from .maps import H2Workspace, mul_

DEFAULT_CANDIDATES = (1, 4, 8, 16, 32)


def signature(core, precision: str, nrhs: int) -> str:
    a = core.arrays
    key = json.dumps({"n": int(a.test.tree.points.shape[0]), "boxes": int(a.test.tree.n_boxes),
                      "pairs": int(len(a.pairs)), "near": int(a.n_near_values), "K": int(a.K),
                      "kind": a.m2l_kind, "dtype": a.dtype, "precision": precision, "nrhs": int(nrhs)},
                     sort_keys=True)
    return hashlib.sha1(key.encode()).hexdigest()[:16]


def tune(core, nrhs: int = 32, precision: str = "mixed", candidates=DEFAULT_CANDIDATES, repeats: int = 3,
         cache_path: str | None = None, use_graph: bool = True):
    """Returns (best_block, {block: ms_per_column}); measured on the core's device."""
    import torch

    sig = signature(core, precision, nrhs)
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as fh:
            cache = json.load(fh)
        if sig in cache:
            hit = cache[sig]
            return int(hit["best"]), {int(k): v for k, v in hit["ms_per_column"].items()}
    n_out, n_in = core.shape
    cplx = core.dtype == "c64"
    tdt = torch.complex64 if cplx else torch.float32
    g = torch.Generator(device="cpu").manual_seed(7)
    X = torch.randn((nrhs, n_in), generator=g, dtype=torch.float32)
    if cplx:
        X = X + 1j * torch.randn((nrhs, n_in), generator=g, dtype=torch.float32)
    X = X.to(tdt).cuda().t()
    Y = torch.zeros((nrhs, n_out), dtype=tdt, device="cuda").t()
    results = {}
    for block in candidates:
        if block > nrhs:
            continue
        ws = H2Workspace(core, max_rhs=block, precision=precision, use_graph=use_graph)
        mul_(Y, ws, X)  # warm (captures graphs)
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(repeats):
            t0 = time.perf_counter()
            mul_(Y, ws, X)
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        results[block] = 1e3 * best / nrhs
        ws.close()
    best_block = min(results, key=results.get)
    if cache_path:
        cache = {}
        if os.path.exists(cache_path):
            with open(cache_path) as fh:
                cache = json.load(fh)
        cache[sig] = {"best": best_block, "ms_per_column": {str(k): v for k, v in results.items()}}
        with open(cache_path, "w") as fh:
            json.dump(cache, fh, indent=1)
    return best_block, results


def tuned_workspace(core, nrhs: int = 32, precision: str = "mixed", **kw):
    """A workspace with the tuned RHS block."""
    block, _ = tune(core, nrhs=nrhs, precision=precision, **kw)
    return H2Workspace(core, max_rhs=block, precision=precision, use_graph=True)


def _candidates_fit(core, candidates, budget_bytes):
    """Drop blocks whose workspace would not fit ``budget_bytes`` (estimated)."""
    a = core.arrays
    es = 8
    out = []
    for b in candidates:
        ws = b * (a.test.tree.n_boxes * a.K * es * 2 + 2 * (2 * a.p - 1) ** 3 * a.test.tree.n_boxes * es)
        if ws <= budget_bytes:
            out.append(b)
    return tuple(out) or (min(candidates),)
