"""Autotuner for the batched product: picks a workspace's RHS block and tile shapes per map.

What is tuned (each a workspace option of ``h2g_ws_opts``):

* ``max_rhs`` -- the RHS block, columns per pass.  A product over ``nrhs`` columns runs as
  ceil(nrhs / block) passes of the whole hierarchy.  Wider blocks amortise every read of the
  near field, the collections and the couplings over more columns; narrower blocks need less
  workspace memory (moments and spectra scale with the block).
* ``m2l_groups`` -- workgroups per frequency of the lattice Toeplitz coupling kernel (groups of
  Morton-consecutive parent blocks).  Fewer, larger groups load each frequency's coupling table
  into LDS fewer times; more, smaller groups keep fewer frequencies in flight per XCD, so the
  source spectra of a frequency stay in that XCD's L2.
* ``near_cols`` -- columns per workgroup of the near-field kernel (one, two or four 16-column
  MFMA tiles per wave): wider tiles reuse every matrix value for more columns, narrower ones
  run more workgroups with fewer registers each.
* ``persistent`` -- the coupling stage of all lattice levels in one persistent launch (per-XCD
  task lists, the small levels filling the tail of the large ones) or one launch per level; the
  persistent form keeps the spectra of all levels at once (more workspace memory).
* ``dense_pairs`` -- dense couplings of real maps on receiving boxes two by two (each operator
  the two share read once) or one by one: pairs halve the operator stream where the operators
  come from HBM, single boxes keep more waves per CU.

The best values depend on the map (N, p, tree depth, near-field density, coupling kind) and the
device, so they are measured: each candidate runs the same product on synthetic columns, timed
on the device, and the fastest per column wins.  The search is coordinate-wise (RHS block with
the default shapes, then the coupling groups, the near-field columns, the persistent stage and the
receiver pairs, each where the map has it), which needs a few
products per candidate instead of the full grid.  Results are cached per map signature in a
JSON file, so a serving process tunes a map once.
"""

from __future__ import annotations

import hashlib
import json
import os
import time

from .maps import H2Workspace, mul_

DEFAULT_BLOCKS = (8, 16, 32)
DEFAULT_M2L_GROUPS = (256, 512, 768)
DEFAULT_NEAR_COLS = (16, 32)
DEFAULT_PERSISTENT = (0, 1)
DEFAULT_DENSE_PAIRS = (-1, 1)


def signature(core, precision: str, nrhs: int) -> str:
    a = core.arrays
    key = json.dumps({"n": int(a.test.tree.points.shape[0]), "boxes": int(a.test.tree.n_boxes),
                      "pairs": int(len(a.pairs)), "near": int(a.n_near_values), "K": int(a.K),
                      "kind": a.m2l_kind, "dtype": a.dtype, "precision": precision, "nrhs": int(nrhs)},
                     sort_keys=True)
    return hashlib.sha1(key.encode()).hexdigest()[:16]


def _dense_real(core):
    """Whether the map's couplings are dense real operators (dense_pairs matters)."""
    return core.arrays.m2l_kind in ("dense", "symdense") and core.dtype != "c64"


def _lattice(core):
    """Whether the map's couplings run on the lattice Toeplitz kernels (m2l_groups matters)."""
    return core.arrays.m2l_kind in ("toeplitz", "symtoeplitz")


def tune(core, nrhs: int = 32, precision: str = "mixed", blocks=DEFAULT_BLOCKS, m2l_groups=DEFAULT_M2L_GROUPS,
         near_cols=DEFAULT_NEAR_COLS, persistent=DEFAULT_PERSISTENT, dense_pairs=DEFAULT_DENSE_PAIRS,
         repeats: int = 3, cache_path: str | None = None, use_graph: bool = True):
    """Returns (best, table): ``best`` = {"max_rhs", "m2l_groups", "near_cols", "persistent",
    "dense_pairs"}
    (0 = the library's default) and ``table`` = [(config, ms per column)] for every candidate
    timed."""
    import torch

    sig = signature(core, precision, nrhs)
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as fh:
            cache = json.load(fh)
        if sig in cache:
            hit = cache[sig]
            return dict(hit["best"]), [(dict(c), ms) for c, ms in hit["table"]]
    n_out, n_in = core.shape
    cplx = core.dtype == "c64"
    tdt = torch.complex64 if cplx else torch.float32
    g = torch.Generator(device="cpu").manual_seed(7)
    X = torch.randn((nrhs, n_in), generator=g, dtype=torch.float32)
    if cplx:
        X = X + 1j * torch.randn((nrhs, n_in), generator=g, dtype=torch.float32)
    X = X.to(tdt).cuda().t()
    Y = torch.zeros((nrhs, n_out), dtype=tdt, device="cuda").t()
    table = []
    timed = {}

    def measure(cfg):
        key = (cfg["max_rhs"], cfg["m2l_groups"], cfg["near_cols"], cfg["persistent"], cfg["dense_pairs"])
        if key in timed:
            return timed[key]
        ws = H2Workspace(core, precision=precision, use_graph=use_graph, **cfg)
        try:
            mul_(Y, ws, X)  # warm (captures the graphs)
            torch.cuda.synchronize()
            best = float("inf")
            for _ in range(repeats):
                t0 = time.perf_counter()
                mul_(Y, ws, X)
                torch.cuda.synchronize()
                best = min(best, time.perf_counter() - t0)
        finally:
            ws.close()
        ms = 1e3 * best / nrhs
        timed[key] = ms
        table.append((dict(cfg), ms))
        return ms

    best = {"max_rhs": min(nrhs, max(blocks)), "m2l_groups": 0, "near_cols": 0, "persistent": 0, "dense_pairs": 0}
    for axis, values in (("max_rhs", [b for b in blocks if b <= nrhs] or [nrhs]),
                         ("m2l_groups", list(m2l_groups) if _lattice(core) else []),
                         ("near_cols", list(near_cols) if not cplx else []),
                         ("persistent", list(persistent) if _lattice(core) else []),
                         ("dense_pairs", list(dense_pairs) if _dense_real(core) else [])):
        if not values:
            continue
        scores = {}
        for v in values:
            cfg = dict(best, **{axis: v})
            scores[v] = measure(cfg)
        best[axis] = min(scores, key=scores.get)
    if cache_path:
        cache = {}
        if os.path.exists(cache_path):
            with open(cache_path) as fh:
                cache = json.load(fh)
        cache[sig] = {"best": best, "table": table}
        with open(cache_path, "w") as fh:
            json.dump(cache, fh, indent=1)
    return best, table


def tuned_workspace(core, nrhs: int = 32, precision: str = "mixed", use_graph: bool = True, stream=None, **kw):
    """A workspace with the tuned RHS block and tile shapes."""
    best, _ = tune(core, nrhs=nrhs, precision=precision, **kw)
    return H2Workspace(core, precision=precision, use_graph=use_graph, stream=stream, **best)
