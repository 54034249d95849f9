"""Serving loop: resident H2 maps, continuous per-map batching, memory-budget admission.

An iterative solver issues one matvec at a time per map; many solvers (or many right-hand
sides of one) run against a few resident maps.  The server keeps the built cores in HBM and
turns the incoming stream into multi-RHS products:

* a map is admitted when its core plus its workspaces fit the HBM budget
  (``AdmissionError`` otherwise); each map owns ``ws_per_map`` workspaces of ``max_batch``
  columns, so a batch needs no new memory once its map is admitted;
* requests queue per (map, orientation); whenever a workspace of that map is free, up to
  ``max_batch`` queued requests form one product -- requests that arrived while earlier
  batches were on the GPU join the next batch (continuous batching);
* the workspaces of all maps share a pool of ``n_streams`` HIP streams (default 4, the number
  of hardware queues a process gets by default; more streams than queues would only serialise
  on the queues); each batch goes to the stream expected to drain first (per-map product times
  measured at admission, a 1-column and a 32-column product), so a small map's batch does not
  wait in order behind a large map's product (``h2g_ws_set_stream`` moves the workspace, ordered
  after its previous product); ``h2g_matvec`` with ``H2G_HOST_PTRS | H2G_ASYNC`` on page-locked
  buffers enqueues host->device copy, product and device->host copy back to back, so batches on
  different streams overlap on the device while the host forms the next batch;
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
    with a reduced precision their FP32 row-major copies per orientation, columns padded to 16),
    collections, couplings, transfers, and for dense couplings and non-separable (complex)
    transfers the MFMA fragment matrices per orientation and representation."""
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
                 n_streams: int = 4, use_graph: bool = False, autotune: bool = False):
        """``use_graph``: capture each (orientation, batch width) product of a workspace in a hipGraph
        on first use (worth it for maps whose products are launch-bound; a stream of single
        requests produces every width from 1 to ``max_batch``).  ``autotune``: measure each
        admitted map's tile shapes (``h2g.autotune.tune`` at ``max_batch`` columns) before its
        workspaces are made."""
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

    def add_map(self, name, arrays):
        """Admit and build a map, or raise AdmissionError when it does not fit the budget."""
        est = estimate_core_bytes(arrays)
        n = arrays.test.tree.points.shape[0]
        es = 16 if arrays.dtype == "c64" else 8
        ws_est = self.ws_per_map * (n * self.max_batch * (4 * es) + arrays.test.tree.n_boxes * arrays.K *
                                    self.max_batch * es + 2 * _spectrum_values(arrays) *
                                    arrays.test.tree.n_boxes * self.max_batch * es)
        if self.used_bytes() + est + ws_est > self.budget:
            raise AdmissionError(f"map {name!r} needs ~{(est + ws_est) / 1e9:.1f} GB, "
                                 f"{(self.budget - self.used_bytes()) / 1e9:.1f} GB of the budget left")
        core = GpuH2Core(arrays, precisions=("plain", self.precision))
        dtype = np.complex64 if arrays.dtype == "c64" else np.float32
        rm = _ResidentMap(name, core, self.ws_per_map, self.max_batch, self.precision, dtype, self._stream,
                          self.use_graph, self._tile(core))
        self.maps[name] = rm
        if self.used_bytes() > self.budget:
            rm.close()
            del self.maps[name]
            raise AdmissionError(f"map {name!r} exceeded the budget once built")
        rm.calibrate()
        return rm

    def add_core(self, name, core):
        """Admit an already built core (its workspaces are created here), or raise AdmissionError."""
        a = core.arrays
        es = 16 if a.dtype == "c64" else 8
        n = max(core.shape)
        ws_est = self.ws_per_map * (n * self.max_batch * (4 * es) + a.test.tree.n_boxes * a.K * self.max_batch * es +
                                    2 * _spectrum_values(a) * a.test.tree.n_boxes * self.max_batch * es)
        if self.used_bytes() + ws_est > self.budget:
            raise AdmissionError(f"map {name!r} needs ~{ws_est / 1e9:.1f} GB of workspaces, "
                                 f"{(self.budget - self.used_bytes()) / 1e9:.1f} GB of the budget left")
        dtype = np.complex64 if a.dtype == "c64" else np.float32
        rm = _ResidentMap(name, core, self.ws_per_map, self.max_batch, self.precision, dtype, self._stream,
                          self.use_graph, self._tile(core))
        self.maps[name] = rm
        rm.calibrate()
        return rm

    def _tile(self, core):
        if not self.autotune:
            return None
        from .autotune import tune

        # one launch per coupling level: the admission estimate sizes the spectra for the largest level
        best, _ = tune(core, nrhs=self.max_batch, precision=self.precision, blocks=(self.max_batch,), repeats=2,
                       use_graph=self.use_graph, persistent=(0,))
        return best

    def remove_map(self, name):
        self.maps.pop(name).close()

    def close(self):
        for name in list(self.maps):
            self.remove_map(name)

    def submit(self, req: Request):
        if req.map_name not in self.maps:
            raise KeyError(req.map_name)
        self.maps[req.map_name].queues[req.op].append(req)

    def pending(self):
        return sum(len(q) for m in self.maps.values() for q in m.queues.values()) + \
            sum(len(s.batch) for m in self.maps.values() for s in m.slots)

    def _launch(self, rm: _ResidentMap, slot: _Slot, op: str):
        t_host = time.perf_counter()
        q = rm.queues[op]
        batch = [q.popleft() for _ in range(min(self.max_batch, len(q)))]
        nb = len(batch)
        n_out, n_in = rm.core.shape
        if op != "N":
            n_out, n_in = n_in, n_out
        for j, r in enumerate(batch):
            slot.xhost[:n_in, j] = r.x
        # the stream expected to drain first
        now = time.perf_counter()
        si = min(range(len(self._streams)), key=lambda i: max(now, self._free_at[i]))
        self._free_at[si] = max(now, self._free_at[si]) + rm.cost(nb)
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

    def _harvest(self, slot: _Slot, now: float):
        for j, r in enumerate(slot.batch):
            r.t_done = now
            r.result = slot.yhost[:slot.n_out, j].copy()
            self.done.append(r)
        slot.batch = []
        self.host_harvest_s += time.perf_counter() - now

    def step(self):
        """One scheduler pass: harvest finished batches, launch new ones on free workspaces."""
        progressed = False
        for rm in self.maps.values():
            for slot in rm.slots:
                if slot.batch and not rm.busy(slot):
                    self._harvest(slot, time.perf_counter())
                    progressed = True
            for slot in rm.slots:
                if slot.batch:
                    continue
                # the orientation with the oldest waiting request goes first
                heads = [(q[0].t_arrival, op) for op, q in rm.queues.items() if q]
                if not heads:
                    break
                _, op = min(heads)
                self._launch(rm, slot, op)
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
