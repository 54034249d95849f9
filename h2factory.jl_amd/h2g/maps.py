"""GPU H2 maps with H2Factory's operator interface.

Mirrors the reference's map surface (``src/H2Factory.jl:155-230``):

* ``GpuH2Core`` -- the read-only core (``GalerkinH2Core`` / ``PetrovGalerkinH2Core``,
  ``src/maps/core.jl:17-54``) built from a method description;
* ``H2Workspace(core, ...)`` -- the executable map (``src/maps/workspace.jl:81-100``); each
  workspace owns its device buffers and runs one product at a time
  (``ConcurrentWorkspaceError`` otherwise, ``src/maps/workspace.jl:284-343``);
  ``similarworkspace`` shares the core (``src/maps/workspace.jl:109-113``);
* ``transpose(A)`` / ``adjoint(A)`` -- lazy orientation views sharing the workspace;
* ``mul_(y, A, x[, alpha, beta])``, ``farmul_``, ``nearmul_`` -- the 3- and 5-argument
  products (``src/maps/abstracth2map.jl:42-86``, ``src/farinteractions/farmul.jl:8-70``,
  ``src/nearinteractions/nearmul.jl:8-14``), for a vector or an n x nrhs matrix;
* ``Matrix(A)`` -- column-by-column reconstruction.

Vectors are numpy arrays (host, copied through the library) or torch tensors on the GPU
(used in place).  Every product runs in ``libh2g.so``; there is no other execution path.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as H
from .method import KINDS, MapArrays


class ConcurrentWorkspaceError(RuntimeError):
    """A workspace was used by two products at once."""


class ArgumentError(ValueError):
    pass


class DimensionMismatch(ValueError):
    pass


def _raise(code: int):
    if code == H.H2G_OK:
        return
    msg = H.last_error()
    if code == H.H2G_EBUSY:
        raise ConcurrentWorkspaceError(msg)
    if code == H.H2G_EDIM:
        raise DimensionMismatch(msg)
    if code in (H.H2G_EINVAL, H.H2G_EUNSUPPORTED):
        raise ArgumentError(msg)
    raise H.H2GError(code, msg)


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


class GpuH2Core:
    """Read-only device core built from a ``MapArrays`` description."""

    def __init__(self, arrays: MapArrays, precisions=("plain",), device: int = 0, fill=True, lattice=True):
        self.arrays = arrays
        self.use_coords = lattice
        self._keep = []
        d = H.CoreDesc()
        cdt = np.complex128 if arrays.dtype == "c64" else np.float64
        d.dtype = H.H2G_C64 if arrays.dtype == "c64" else H.H2G_F64
        d.K = arrays.K
        d.p = arrays.p if arrays.m2l_kind in ("toeplitz", "symtoeplitz") else 0
        d.galerkin = 1 if arrays.galerkin else 0
        d.test = self._side(arrays.test, cdt)
        if not arrays.galerkin:
            d.trial = self._side(arrays.trial, cdt)
        d.m2l_kind = KINDS[arrays.m2l_kind]
        pt, pc = _i32(arrays.pairs[:, 0]), _i32(arrays.pairs[:, 1])
        po = _i32(arrays.pair_op)
        ps = _i32(arrays.pair_sym)
        toep = arrays.m2l_kind in ("toeplitz", "symtoeplitz")
        ops = np.ascontiguousarray(arrays.ops, dtype=np.complex128 if toep else cdt)
        perms = _i32(arrays.perms)
        self._keep += [pt, pc, po, ps, ops, perms]
        d.n_pairs = len(arrays.pairs)
        d.pair_test, d.pair_trial, d.pair_op = H.ptr(pt, C.c_int32), H.ptr(pc, C.c_int32), H.ptr(po, C.c_int32)
        d.pair_sym = H.ptr(ps, C.c_int32) if ps is not None else None
        d.n_ops = len(ops)
        d.ops = H.ptr(ops)
        d.n_perms = 0 if perms is None else perms.shape[0]
        d.perms = H.ptr(perms, C.c_int32) if perms is not None else None
        nt, nc = _i32(arrays.near_pairs[:, 0]), _i32(arrays.near_pairs[:, 1])
        no = _i64(arrays.near_offset)
        nv = None if arrays.near_values is None else np.ascontiguousarray(arrays.near_values, dtype=cdt)
        self._keep += [nt, nc, no, nv]
        d.n_near_blocks = len(arrays.near_pairs)
        d.near_test, d.near_trial, d.near_offset = H.ptr(nt, C.c_int32), H.ptr(nc, C.c_int32), H.ptr(no, C.c_int64)
        d.n_near_values = arrays.n_near_values
        d.near_values = H.ptr(nv) if nv is not None else None
        a = complex(arrays.alpha)
        d.alpha_re, d.alpha_im = a.real, a.imag
        d.precisions = sum(1 << H.PRECISIONS[p] for p in set(precisions))
        d.device = device
        self.desc = d
        self.dtype = arrays.dtype
        h = C.c_void_p()
        _raise(H.lib().h2g_core_create(C.byref(d), C.byref(h)))
        self.handle = h
        self._keep.clear()
        if fill and (arrays.near_values is None or arrays.test.collection is None):
            self._fill()

    def _side(self, s, cdt):
        t = s.tree
        sd = H.SideDesc()
        arrs = dict(perm=_i64(t.perm), lo=_i32(t.level_offsets), par=_i32(t.parent), sec=_i32(t.sector),
                    b=_i64(t.begin), e=_i64(t.end), leaf=np.ascontiguousarray(t.is_leaf, dtype=np.uint8),
                    coll=None if s.collection is None else np.ascontiguousarray(s.collection, dtype=cdt),
                    tr=np.ascontiguousarray(s.transfers, dtype=cdt),
                    coord=None if getattr(t, "coord", None) is None else np.ascontiguousarray(t.coord, dtype=np.int32))
        self._keep.append(arrs)
        sd.n_dofs = t.points.shape[0]
        sd.n_levels = t.n_levels
        sd.n_boxes = t.n_boxes
        sd.perm = H.ptr(arrs["perm"], C.c_int64)
        sd.level_offsets = H.ptr(arrs["lo"], C.c_int32)
        sd.parent = H.ptr(arrs["par"], C.c_int32)
        sd.sector = H.ptr(arrs["sec"], C.c_int32)
        sd.dof_begin = H.ptr(arrs["b"], C.c_int64)
        sd.dof_end = H.ptr(arrs["e"], C.c_int64)
        sd.is_leaf = H.ptr(arrs["leaf"], C.c_uint8)
        sd.collection = H.ptr(arrs["coll"]) if arrs["coll"] is not None else None
        sd.transfers = H.ptr(arrs["tr"])
        sd.coord = H.ptr(arrs["coord"], C.c_int32) if (arrs["coord"] is not None and self.use_coords) else None
        return sd

    def _fill(self):
        a = self.arrays
        m = H.MethodDesc()
        m.kernel = {"laplace": 0, "helmholtz": 1}[a.kernel]
        m.wavenumber = float(a.wavenumber)
        m.p = a.p
        nodes = np.ascontiguousarray(a.nodes, dtype=np.float64)
        keep = [nodes]

        def side(s):
            t = s.tree
            pts = np.ascontiguousarray(t.points, dtype=np.float64)
            cen = np.ascontiguousarray(t.centers(), dtype=np.float64)
            hw = np.ascontiguousarray(t.halfwidths(), dtype=np.float64)
            keep.extend([pts, cen, hw])
            sd = H.MethodSideDesc()
            sd.points, sd.box_center, sd.box_halfwidth = H.ptr(pts, C.c_double), H.ptr(cen, C.c_double), H.ptr(hw, C.c_double)
            return sd

        m.nodes = H.ptr(nodes, C.c_double)
        m.test = side(a.test)
        if not a.galerkin:
            m.trial = side(a.trial)
        m.fill_near = 1 if a.near_values is None else 0
        m.fill_collections = 1 if a.test.collection is None else 0
        _raise(H.lib().h2g_method_fill(self.handle, C.byref(m)))

    @property
    def shape(self):
        r, c = C.c_int64(), C.c_int64()
        _raise(H.lib().h2g_core_size(self.handle, C.byref(r), C.byref(c)))
        return (r.value, c.value)

    def device_bytes(self):
        b = C.c_int64()
        _raise(H.lib().h2g_core_device_bytes(self.handle, C.byref(b)))
        return b.value

    def close(self):
        if getattr(self, "handle", None):
            H.lib().h2g_core_release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_DT = {"f64": H.H2G_F64, "c64": H.H2G_C64, "f32": H.H2G_F32, "c32": H.H2G_C32}


def _dtype_code(a):
    name = str(a.dtype)
    if "complex128" in name:
        return H.H2G_C64
    if "complex64" in name:
        return H.H2G_C32
    if "float64" in name:
        return H.H2G_F64
    if "float32" in name:
        return H.H2G_F32
    raise ArgumentError(f"unsupported element type {a.dtype}")


class H2Workspace:
    """The executable GPU map (workspace over a shared core)."""

    op = H.H2G_OP_N

    def __init__(self, core: GpuH2Core, max_rhs=1, precision="plain", use_graph=False, timing=False, stream=None,
                 m2l_groups=0, near_cols=0, persistent=0, dense_pairs=0):
        """``stream``: a HIP stream handle (int) the workspace runs on, shared with other workspaces,
        or None for streams of its own.  ``m2l_groups`` / ``near_cols``: tile shapes (0: the
        library's defaults; ``h2g.autotune.tune`` measures them for a map).  ``persistent``: the
        lattice coupling stage of all levels in one persistent launch.  ``dense_pairs``: dense
        couplings of real maps on receiving boxes two by two (1), one by one (-1), or the library's
        choice (0)."""
        self.core = core
        o = H.WsOpts()
        o.stream = stream
        o.precision = H.PRECISIONS[precision]
        o.max_rhs = int(max_rhs)
        o.use_graph = int(bool(use_graph))
        o.timing = int(bool(timing))
        o.m2l_groups = int(m2l_groups)
        o.near_cols = int(near_cols)
        o.persistent = int(persistent)
        o.dense_pairs = int(dense_pairs)
        self.opts = dict(max_rhs=max_rhs, precision=precision, use_graph=use_graph, timing=timing, stream=stream,
                         m2l_groups=m2l_groups, near_cols=near_cols, persistent=persistent,
                         dense_pairs=dense_pairs)
        h = C.c_void_p()
        _raise(H.lib().h2g_ws_create(core.handle, C.byref(o), C.byref(h)))
        self.handle = h

    @property
    def shape(self):
        return self.core.shape

    def size(self, dim=None):
        return self.shape if dim is None else self.shape[dim - 1]

    @property
    def T(self):
        return _OrientedMap(self, H.H2G_OP_T)

    @property
    def H(self):
        return _OrientedMap(self, H.H2G_OP_C)

    def _ws(self):
        return self

    def device_bytes(self):
        b = C.c_int64()
        _raise(H.lib().h2g_ws_device_bytes(self.handle, C.byref(b)))
        return b.value

    def stream(self):
        return H.lib().h2g_ws_stream(self.handle)

    def set_stream(self, stream):
        """Move a workspace created with ``stream=`` to another caller stream (ordered after its
        last product)."""
        _raise(H.lib().h2g_ws_set_stream(self.handle, C.c_void_p(stream)))

    def acquire(self):
        _raise(H.lib().h2g_ws_acquire(self.handle))

    def release(self):
        _raise(H.lib().h2g_ws_release(self.handle))

    def sync(self):
        _raise(H.lib().h2g_ws_sync(self.handle))

    def timings(self):
        """Per-phase device times of the last product: (phase, level, ms) records."""
        n = C.c_int32()
        _raise(H.lib().h2g_ws_timing(self.handle, None, 0, C.byref(n)))
        arr = (H.TimingRecord * max(n.value, 1))()
        _raise(H.lib().h2g_ws_timing(self.handle, arr, n.value, C.byref(n)))
        return [(r.phase, r.level, r.ms) for r in arr[:n.value]]

    def close(self):
        if getattr(self, "handle", None):
            H.lib().h2g_ws_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __matmul__(self, x):
        return _apply_new(self, self.op, x)


class _OrientedMap:
    """``transpose(A)`` / ``adjoint(A)``: shares the workspace (and its busy flag)."""

    def __init__(self, ws: H2Workspace, op: int):
        self.ws = ws
        self.op = op

    @property
    def shape(self):
        r, c = self.ws.shape
        return (c, r)

    def _ws(self):
        return self.ws

    def __matmul__(self, x):
        return _apply_new(self, self.op, x)


def transpose(A):
    """``transpose(A)``; ``transpose(transpose(A)) is A``."""
    if A.op == H.H2G_OP_N:
        return _OrientedMap(A._ws(), H.H2G_OP_T)
    if A.op == H.H2G_OP_T:
        return A._ws()
    raise ArgumentError("transpose(adjoint(A)) (= conj(A)) is not a supported orientation")


def adjoint(A):
    """``adjoint(A)``; ``adjoint(adjoint(A)) is A``."""
    if A.op == H.H2G_OP_N:
        return _OrientedMap(A._ws(), H.H2G_OP_C)
    if A.op == H.H2G_OP_C:
        return A._ws()
    raise ArgumentError("adjoint(transpose(A)) (= conj(A)) is not a supported orientation")


def _is_torch(a):
    return type(a).__module__.startswith("torch")


def _call(A, which, y, x, alpha, beta):
    ws = A._ws()
    op = A.op
    n_out, n_in = A.shape
    xd = _dtype_code(x)
    if _dtype_code(y) != xd:
        raise ArgumentError("x and y must have the same element type")
    if x.ndim not in (1, 2) or y.ndim != x.ndim:
        raise DimensionMismatch("x and y must both be vectors or both matrices")
    nrhs = 1 if x.ndim == 1 else x.shape[1]
    if x.shape[0] != n_in or y.shape[0] != n_out or (x.ndim == 2 and y.shape[1] != nrhs):
        raise DimensionMismatch(f"map is {n_out}x{n_in}, x is {tuple(x.shape)}, y is {tuple(y.shape)}")
    five = alpha is not None
    sc = {H.H2G_F64: np.float64, H.H2G_F32: np.float32, H.H2G_C64: np.complex128, H.H2G_C32: np.complex64}[xd]
    a = np.array([alpha if five else 1], dtype=sc)
    b = np.array([beta if five else 0], dtype=sc)
    ap = H.ptr(a) if five else None
    bp = H.ptr(b) if five else None
    if _is_torch(x):
        import torch

        xx = x if x.ndim == 1 else x
        if xx.ndim == 2 and xx.stride(0) != 1:
            xx = xx.t().contiguous().t()
        elif xx.ndim == 1 and xx.stride(0) != 1:
            xx = xx.contiguous()
        yy = y
        copy_back = False
        if (yy.ndim == 2 and yy.stride(0) != 1) or (yy.ndim == 1 and yy.stride(0) != 1):
            yy = y.t().contiguous().t() if y.ndim == 2 else y.contiguous()
            copy_back = True
        ldx = xx.shape[0] if xx.ndim == 1 else xx.stride(1)
        ldy = yy.shape[0] if yy.ndim == 1 else yy.stride(1)
        if nrhs == 1 and x.ndim == 2:
            ldx, ldy = max(ldx, n_in), max(ldy, n_out)
        torch.cuda.current_stream().synchronize()
        _raise(H.lib().h2g_matvec(ws.handle, op, which, xd, nrhs, C.c_void_p(xx.data_ptr()), ldx,
                                  C.c_void_p(yy.data_ptr()), ldy, ap, bp, 0))
        if copy_back:
            y.copy_(yy)
        return y
    xa = np.asfortranarray(x)
    ya = y if (y.ndim == 1 or y.flags["F_CONTIGUOUS"]) else np.asfortranarray(y)
    if not ya.flags["WRITEABLE"]:
        raise ArgumentError("y is read-only")
    _raise(H.lib().h2g_matvec(ws.handle, op, which, xd, nrhs, H.ptr(xa), n_in, H.ptr(ya), n_out, ap, bp,
                              H.H2G_HOST_PTRS))
    if ya is not y:
        y[...] = ya
    return y


def mul_(y, A, x, alpha=None, beta=None):
    """``mul!(y, A, x[, alpha, beta])``."""
    return _call(A, H.H2G_MUL, y, x, alpha, beta)


def farmul_(y, A, x, alpha=None, beta=None):
    """``farmul!(y, A, x[, alpha, beta])``."""
    return _call(A, H.H2G_FARMUL, y, x, alpha, beta)


def nearmul_(y, A, x, alpha=None, beta=None):
    """``nearmul!(y, A, x[, alpha, beta])``."""
    return _call(A, H.H2G_NEARMUL, y, x, alpha, beta)


def _apply_new(A, op, x):
    n_out = A.shape[0]
    if _is_torch(x):
        import torch

        y = torch.zeros((n_out,) + tuple(x.shape[1:]), dtype=x.dtype, device=x.device)
        if y.ndim == 2:
            y = y.t().contiguous().t()
    else:
        y = np.zeros((n_out,) + x.shape[1:], dtype=x.dtype, order="F")
    return mul_(y, A, x)


def similarworkspace(A):
    ws = A._ws()
    return H2Workspace(ws.core, **ws.opts)


def Matrix(A):
    """Dense matrix of the map, column by column (host; small maps only)."""
    n_out, n_in = A.shape
    dt = np.complex128 if A._ws().core.dtype == "c64" else np.float64
    eye = np.eye(n_in, dtype=dt, order="F")
    out = np.zeros((n_out, n_in), dtype=dt, order="F")
    return mul_(out, A, eye)
