"""Serving loop: resident H2 maps, continuous per-map batching, memory-budget admission.

An iterative solver issues one matvec at a time per map; many solvers (or many right-hand
sides of one) run against a few resident maps.  The server keeps the built cores in HBM and
turns the incoming stream into multi-RHS products:

* a map is admitted when its core plus its workspaces fit the HBM budget
  (``AdmissionError`` otherwise); each map owns ``ws_per_map`` workspaces of ``max_batch``
  columns, or, when those do not fit, one workspace of ``max_batch`` and then of
  ``max_batch // 2`` columns (the spectra of a Toeplitz map's largest level dominate a
  workspace), so a batch needs no new memory once its map is admitted;
* requests queue per (map, orientation); whenever a stream has room (``stream_depth`` batches
  in flight, default 1) and a workspace of some map is free, up to ``max_batch`` queued
  requests of the (map, orientation) with the oldest waiting request form one product --
  requests that arrive while the streams are busy join the next batch instead of waiting
  behind it in a stream's queue (continuous batching; a map's product costs nearly the same
  for 1 as for 8 columns, so batches formed late are the cheap ones);
* the workspaces of all maps share a pool of ``n_streams`` HIP streams (default 4, the number
  of hardware queues a process gets by default; more streams than queues would only serialise
  on the queues); each batch goes to the stream with room expected to drain first (per-map
  product times measured at admission, a 1-column and a ``max_batch``-column product)
  (``h2g_ws_set_stream`` moves the workspace, ordered after its previous product);
  ``h2g_matvec`` with ``H2G_HOST_PTRS | H2G_ASYNC`` on page-locked buffers enqueues
  host->device copy, product and device->host copy back to back, so batches on different
  streams overlap on the device while the host forms the next batch;
* completion is polled with ``h2g_ws_query`` (one completion event per workspace, so
  workspaces sharing a stream are told apart); latency = completion time - arrival time;
* nothing else records work on a workspace's stream, and ``close`` waits for every product
  before it releases the pinned buffers and destroys the workspaces (a stream must not be
  destroyed while events recorded on it are still alive).

Each result is the product of its own column: the kernels treat the columns of a batch
independently, so a request's result is the same alone and in the mix (the near-field kernel
for narrow batches sums in a different order than the one for wide batches; the difference is
at the FP64 rounding level).
"""

from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _lib as H
from .maps import GpuH2Core, H2Workspace, _raise

OPS = {"N": H.H2G_OP_N, "T": H.H2G_OP_T, "C": H.H2G_OP_C}


def _spectrum_values(arrays):
    """Stored frequencies per box of a Toeplitz map (L^3; real maps about half: one value of each
    conjugate pair), the estimate of the workspaces' spectrum buffers; 0 for dense couplings."""
    if arrays.m2l_kind not in ("toeplitz", "symtoeplitz"):
        return 0
    L = 2 * arrays.p - 1
    return L ** 3 if arrays.dtype == "c64" else (L * L + 1) // 2 + (arrays.p - 1) * L * L


def workspace_bytes(arrays, max_rhs, precision="mixed", staging=True):
    """Device bytes of one workspace of ``max_rhs`` columns, by the library's layout: tree-ordered
    x and y in the core type, the moment and local pools in the compute type, the FP32 copy of x
    that the mixed near field of real maps reads, for Toeplitz couplings the source and receiver
    spectra of the largest level (an upper bound: every box of the level counted), and with
    ``staging`` the device copies of host-pointer requests (FP32 / ComplexF32 x and y)."""
    cplx = arrays.dtype == "c64"
    mixed = precision != "plain"
    se = 16 if cplx else 8
    ce = se // 2 if mixed else se
    trees = [arrays.test.tree] + ([] if arrays.galerkin else [arrays.trial.tree])
    nmax = max(t.points.shape[0] for t in trees)
    bmax = max(t.n_boxes for t in trees)
    R = int(max_rhs)
    b = 2 * nmax * R * se + 2 * bmax * arrays.K * R * ce
    if staging:
        b += 2 * nmax * R * (8 if cplx else 4)
    if mixed and not cplx:
        w = 16 if R <= 16 else 32 if R <= 32 else 64
        b += nmax * ((R + w - 1) // w * w) * 4
    F = _spectrum_values(arrays)
    if F:
        lv = max(int(np.diff(np.asarray(t.level_offsets)).max()) for t in trees)
        cc = 8 if mixed else 16
        rin = (R + 31) // 32 * 32 if mixed and R >= 4 else R
        b += lv * F * (rin + R) * cc
    return int(b)


class AdmissionError(RuntimeError):
    """A map does not fit the server's HBM budget."""


@dataclass
class Request:
    map_name: str
    op: str
    x: np.ndarray
    t_arrival: float = 0.0
    t_done: float = None
    result: np.ndarray = None
    batch_size: int = 0

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_arrival


@dataclass
class _Slot:
    ws: H2Workspace
    xhost: np.ndarray      # pinned, column-major n x max_batch
    yhost: np.ndarray
    keep: tuple            # the pinned allocations backing xhost / yhost
    batch: list = field(default_factory=list)
    n_out: int = 0
    stream: int = -1       # pool index of the stream the batch in flight runs on
    draining: object = None  # future of the copies of a finished batch's results out of yhost


def _pinned(shape, dtype):
    """Page-locked host array (column-major) and the tensor that owns its memory."""
    import torch

    tdt = {np.dtype(np.float32): torch.float32, np.dtype(np.complex64): torch.complex64}[np.dtype(dtype)]
    t = torch.zeros((shape[1], shape[0]), dtype=tdt).pin_memory()
    return t.numpy().T, t


class _ResidentMap:
    def __init__(self, name, core, n_ws, max_batch, precision, dtype, streams, use_graph, tile=None):
        self.name = name
        self.core = core
        self.queues = {op: deque() for op in OPS}
        self.slots = []
        self.tile = dict(tile or {})
        n_out, n_in = core.shape
        n = max(n_out, n_in)
        shapes = {k: self.tile.get(k, 0) for k in ("m2l_groups", "near_cols")}
        for _ in range(n_ws):
            ws = H2Workspace(core, max_rhs=max_batch, precision=precision, use_graph=use_graph, stream=streams(),
                             **shapes)
            xh, tx = _pinned((n, max_batch), dtype)
            yh, ty = _pinned((n, max_batch), dtype)
            self.slots.append(_Slot(ws, xh, yh, (tx, ty)))
        self.n = n
        self.dtype = dtype
        self.t1 = self.t32 = 1e-3  # product times (s) of 1 and max_batch columns, from calibrate()
        self.max_batch = max_batch

    def calibrate(self):
        """Times a 1-column and a full-width product on the first workspace (also loads the code)."""
        import time as _t

        slot = self.slots[0]
        dt = H.H2G_C32 if self.dtype == np.complex64 else H.H2G_F32
        n_out, n_in = self.core.shape
        for nb, attr in ((1, "t1"), (self.max_batch, "t32")):
            best = None
            for _ in range(2):
                t0 = _t.perf_counter()
                _raise(H.lib().h2g_matvec(slot.ws.handle, H.H2G_OP_N, H.H2G_MUL, dt, nb, H.ptr(slot.xhost),
                                          slot.xhost.shape[0], H.ptr(slot.yhost), slot.yhost.shape[0], None, None,
                                          H.H2G_HOST_PTRS))
                d = _t.perf_counter() - t0
                best = d if best is None else min(best, d)
            setattr(self, attr, best)
        # the other workspaces: one product each (their first host-pointer product allocates the
        # device staging buffers, which must not happen in the middle of the stream)
        for s in self.slots[1:]:
            _raise(H.lib().h2g_matvec(s.ws.handle, H.H2G_OP_N, H.H2G_MUL, dt, 1, H.ptr(s.xhost), s.xhost.shape[0],
                                      H.ptr(s.yhost), s.yhost.shape[0], None, None, H.H2G_HOST_PTRS))

    def cost(self, nb):
        return self.t1 + (self.t32 - self.t1) * (nb - 1) / max(1, self.max_batch - 1)

    def busy(self, slot):
        if not slot.batch:
            return False
        rc = H.lib().h2g_ws_query(slot.ws.handle)
        if rc == H.H2G_EAGAIN:
            return True
        _raise(rc)
        return False

    def device_bytes(self):
        return self.core.device_bytes() + sum(s.ws.device_bytes() for s in self.slots)

    def close(self):
        # every product must have finished before the pinned buffers or the workspaces go away
        for s in self.slots:
            _raise(H.lib().h2g_ws_sync(s.ws.handle))
        for s in self.slots:
            s.ws.close()
        self.slots = []


def estimate_core_bytes(arrays, precisions=("plain", "mixed")):
    """Device bytes of a core before it is built (admission check): near values (and for real maps
    with a reduced precision their FP32 row-major copies per orientation, columns padded to 16)
    and task tables, collections, couplings, transfers, and for dense couplings and non-separable
    (complex) transfers the MFMA fragment matrices per orientation and representation."""
    cplx = arrays.dtype == "c64"
    es = 16 if cplx else 8
    reduced = any(p != "plain" for p in precisions)
    n = arrays.test.tree.points.shape[0] + (0 if arrays.galerkin else arrays.trial.tree.points.shape[0])
    b = arrays.n_near_values * (es + (2 * 4 * 1.05 if reduced and not cplx else 0))
    coll = n * arrays.K * es
    b += coll * (1.5 if reduced else 1.0)
    b += arrays.ops.nbytes * (1.5 if reduced else 1.0)
    b += arrays.test.transfers.nbytes * 1.5
    b += len(arrays.near_pairs) * 40 + len(arrays.pairs) * 24 + n * 24
    # the near field's row-block task tables: a 16-byte entry per source point of every task (at
    # most 64 output rows, 32 for complex maps), in each orientation
    if len(arrays.near_pairs):
        t = arrays.test.tree
        r = t if arrays.galerkin else arrays.trial.tree
        near = np.asarray(arrays.near_pairs, dtype=np.int64)
        st = (np.asarray(t.end) - np.asarray(t.begin))[near[:, 0]]
        sr = (np.asarray(r.end) - np.asarray(r.begin))[near[:, 1]]
        T = 32 if cplx else 64
        b += 16 * int((sr * -(-st // T)).sum() + (st * -(-sr // T)).sum())
    krp2 = (16 * ((arrays.K * (2 if cplx else 1) + 15) // 16)) ** 2
    per = ("plain" in precisions) * 8 + reduced * 4 + any(p.startswith("bf16") for p in precisions) * 6
    if arrays.m2l_kind in ("dense", "symdense"):
        if arrays.m2l_kind == "dense":
            n_fops = len(arrays.ops)
        else:
            n_fops = len(np.unique(np.asarray(arrays.pair_op, dtype=np.int64) * 64 + np.asarray(arrays.pair_sym)))
        b += n_fops * krp2 * per * (3 if cplx else 2)
    if cplx:  # transfers on the dense kernel: 4 orientations, FP64 / FP32
        b += arrays.test.tree.n_levels * 8 * krp2 * (("plain" in precisions) * 8 + reduced * 4) * 4
    return int(b)


class Server:
    """Continuous-batching request loop over resident maps."""

    def __init__(self, budget_bytes: int, max_batch: int = 32, ws_per_map: int = 2, precision: str = "mixed",
                 n_streams: int = 4, use_graph: bool = False, autotune: bool = False, stream_depth: int = 1):
        """``use_graph``: capture each (orientation, batch width) product of a workspace in a hipGraph
        on first use (worth it for maps whose products are launch-bound; a stream of single
        requests produces every width from 1 to ``max_batch``).  ``autotune``: measure each
        admitted map's tile shapes (``h2g.autotune.tune`` at ``max_batch`` columns) before its
        workspaces are made.  ``stream_depth``: batches in flight per stream; a batch is formed
        only when a stream has room, so requests that arrive while every stream is busy join the
        batch rather than waiting behind it in a stream's queue."""
        self.budget = int(budget_bytes)
        self.max_batch = max_batch
        self.ws_per_map = ws_per_map
        self.precision = precision
        self.use_graph = use_graph
        self.autotune = autotune
        self.maps = {}
        self.done = []
        self._n_streams = max(1, n_streams)
        self._pool = None  # the streams, created with the first admitted map
        self._next_stream = 0
        self._free_at = [0.0] * self._n_streams  # expected drain time of each stream
        self._depth = max(1, stream_depth)
        self._inflight = [0] * self._n_streams  # batches in flight per stream
        # results are copied out of the pinned buffers by worker threads (numpy releases the GIL
        # for the copies), so the loop keeps forming and launching batches meanwhile
        from concurrent.futures import ThreadPoolExecutor

        self._copier = ThreadPoolExecutor(max_workers=4)
        # accounting of a stream: host seconds staging batches in / results out, and the device
        # seconds the batches were expected to take (per-map cost model)
        self.host_launch_s = self.host_harvest_s = self.device_est_s = 0.0

    @property
    def _streams(self):
        if self._pool is None:
            import torch

            self._pool = [torch.cuda.Stream() for _ in range(self._n_streams)]
        return self._pool

    def _stream(self):
        s = self._streams[self._next_stream % len(self._streams)]
        self._next_stream += 1
        return s.cuda_stream

    def used_bytes(self):
        return sum(m.device_bytes() for m in self.maps.values())

    def _workspace_plan(self, arrays, core_bytes):
        """(workspaces, columns) of the first configuration that fits the budget beside a core of
        ``core_bytes``: ``ws_per_map`` x ``max_batch``, then 1 x ``max_batch``, then 1 x
        ``max_batch // 2``; None when none does."""
        options = [(self.ws_per_map, self.max_batch), (1, self.max_batch), (1, max(1, self.max_batch // 2))]
        left = self.budget - self.used_bytes() - core_bytes
        for n_ws, nb in options:
            if n_ws * workspace_bytes(arrays, nb, self.precision) <= left:
                return n_ws, nb
        return None

    def _admit(self, name, core, n_ws, nb):
        dtype = np.complex64 if core.arrays.dtype == "c64" else np.float32
        rm = _ResidentMap(name, core, n_ws, nb, self.precision, dtype, self._stream, self.use_graph,
                          self._tile(core, nb))
        self.maps[name] = rm
        if self.used_bytes() > self.budget:
            rm.close()
            del self.maps[name]
            raise AdmissionError(f"map {name!r} exceeded the budget once built")
        rm.calibrate()
        return rm

    def add_map(self, name, arrays):
        """Admit and build a map, or raise AdmissionError when it does not fit the budget."""
        est = estimate_core_bytes(arrays)
        plan = self._workspace_plan(arrays, est)
        if plan is None:
            need = est + workspace_bytes(arrays, max(1, self.max_batch // 2), self.precision)
            raise AdmissionError(f"map {name!r} needs ~{need / 1e9:.1f} GB, "
                                 f"{(self.budget - self.used_bytes()) / 1e9:.1f} GB of the budget left")
        core = GpuH2Core(arrays, precisions=("plain", self.precision))
        try:
            return self._admit(name, core, *plan)
        except AdmissionError:
            core.close()
            raise

    def add_core(self, name, core):
        """Admit an already built core (its workspaces are created here), or raise AdmissionError."""
        plan = self._workspace_plan(core.arrays, 0)
        if plan is None:
            need = workspace_bytes(core.arrays, max(1, self.max_batch // 2), self.precision)
            raise AdmissionError(f"map {name!r} needs ~{need / 1e9:.1f} GB of workspaces, "
                                 f"{(self.budget - self.used_bytes()) / 1e9:.1f} GB of the budget left")
        return self._admit(name, core, *plan)

    def _tile(self, core, nb):
        if not self.autotune:
            return None
        from .autotune import tune

        # one launch per coupling level: the admission estimate sizes the spectra for the largest level
        best, _ = tune(core, nrhs=nb, precision=self.precision, blocks=(nb,), repeats=2,
                       use_graph=self.use_graph, persistent=(0,))
        return best

    def remove_map(self, name):
        rm = self.maps.pop(name)
        for slot in rm.slots:
            if slot.draining is not None:
                self._drained(slot)
        rm.close()

    def close(self):
        self._finish_drains()
        for name in list(self.maps):
            self.remove_map(name)
        self._copier.shutdown(wait=True)

    def _finish_drains(self):
        for rm in self.maps.values():
            for slot in rm.slots:
                if slot.draining is not None:
                    self._drained(slot)

    def _drained(self, slot: _Slot):
        self.host_harvest_s += slot.draining.result()  # (re-raises a failed copy)
        self.done.extend(slot.batch)
        slot.draining = None
        slot.batch = []

    def submit(self, req: Request):
        if req.map_name not in self.maps:
            raise KeyError(req.map_name)
        self.maps[req.map_name].queues[req.op].append(req)

    def pending(self):
        return sum(len(q) for m in self.maps.values() for q in m.queues.values()) + \
            sum(len(s.batch) for m in self.maps.values() for s in m.slots)

    def _launch(self, rm: _ResidentMap, slot: _Slot, op: str, si: int):
        t_host = time.perf_counter()
        q = rm.queues[op]
        batch = [q.popleft() for _ in range(min(rm.max_batch, len(q)))]
        nb = len(batch)
        n_out, n_in = rm.core.shape
        if op != "N":
            n_out, n_in = n_in, n_out
        for j, r in enumerate(batch):
            slot.xhost[:n_in, j] = r.x
        now = time.perf_counter()
        self._free_at[si] = max(now, self._free_at[si]) + rm.cost(nb)
        self._inflight[si] += 1
        slot.stream = si
        self.device_est_s += rm.cost(nb)
        slot.ws.set_stream(self._streams[si].cuda_stream)
        dt = H.H2G_C32 if rm.dtype == np.complex64 else H.H2G_F32
        # host -> device copy, product and device -> host copy, all on the workspace stream
        _raise(H.lib().h2g_matvec(slot.ws.handle, OPS[op], H.H2G_MUL, dt, nb, H.ptr(slot.xhost), slot.xhost.shape[0],
                                  H.ptr(slot.yhost), slot.yhost.shape[0], None, None,
                                  H.H2G_HOST_PTRS | H.H2G_ASYNC))
        for r in batch:
            r.batch_size = nb
        slot.batch = batch
        slot.n_out = n_out
        self.host_launch_s += time.perf_counter() - t_host

    def _harvest(self, slot: _Slot):
        """The batch's product has finished: its stream has room again; its results are copied out
        by a worker, after which the workspace takes a new batch."""
        self._inflight[slot.stream] -= 1
        slot.stream = -1
        batch, n_out, y = slot.batch, slot.n_out, slot.yhost

        def copy_out():
            t0 = time.perf_counter()
            for j, r in enumerate(batch):
                r.result = y[:n_out, j].copy()
                r.t_done = time.perf_counter()
            return time.perf_counter() - t0

        slot.draining = self._copier.submit(copy_out)

    def step(self):
        """One scheduler pass: harvest finished batches; then, while a stream has room, launch the
        (map, orientation) whose oldest waiting request is the oldest of all, on a free workspace of
        that map, on the stream with room expected to drain first."""
        progressed = False
        for rm in self.maps.values():
            for slot in rm.slots:
                if slot.draining is not None:
                    if slot.draining.done():
                        self._drained(slot)
                        progressed = True
                elif slot.batch and not rm.busy(slot):
                    self._harvest(slot)
                    progressed = True
        while True:
            room = [i for i in range(self._n_streams) if self._inflight[i] < self._depth]
            if not room:
                break
            best = None
            for rm in self.maps.values():
                slot = next((sl for sl in rm.slots if not sl.batch), None)
                if slot is None:
                    continue
                for op, q in rm.queues.items():
                    if q and (best is None or q[0].t_arrival < best[0]):
                        best = (q[0].t_arrival, rm, slot, op)
            if best is None:
                break
            now = time.perf_counter()
            si = min(room, key=lambda i: max(now, self._free_at[i]))
            self._launch(best[1], best[2], best[3], si)
            progressed = True
        return progressed

    def drain(self, timeout_s=600.0):
        t0 = time.perf_counter()
        while self.pending():
            if not self.step():
                time.sleep(20e-6)
            if time.perf_counter() - t0 > timeout_s:
                raise TimeoutError("requests still pending")

    def run_stream(self, requests, rate_hz, seed=0):
        """Replay ``requests`` with Poisson arrivals of mean rate ``rate_hz`` (seeded)."""
        self.host_launch_s = self.host_harvest_s = self.device_est_s = 0.0
        rng = np.random.default_rng(seed)
        gaps = rng.exponential(1.0 / rate_hz, size=len(requests))
        t0 = time.perf_counter()
        arrivals = t0 + np.cumsum(gaps)
        i = 0
        while i < len(requests) or self.pending():
            now = time.perf_counter()
            while i < len(requests) and arrivals[i] <= now:
                requests[i].t_arrival = arrivals[i]
                self.submit(requests[i])
                i += 1
            if not self.step():
                time.sleep(20e-6)
        wall = time.perf_counter() - t0
        self.last_accounting = {"host_launch_s": self.host_launch_s, "host_harvest_s": self.host_harvest_s,
                                "device_est_s": self.device_est_s}
        lat = np.array([r.latency for r in requests])
        return {"requests": len(requests), "wall_s": wall, "matvecs_per_s": len(requests) / wall,
                "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
                "mean_batch": float(np.mean([r.batch_size for r in requests]))}
