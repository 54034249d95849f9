"""ctypes binding of ``libh2g.so`` (declarations in ``include/h2g.h``).

The library is loaded from ``h2factory.jl_amd/lib/libh2g.so``; a missing library raises
immediately -- there is no fallback path.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lib", "libh2g.so")

H2G_OK, H2G_EBUSY, H2G_EINVAL, H2G_EDIM, H2G_ENOMEM, H2G_EHIP, H2G_EUNSUPPORTED, H2G_EAGAIN = range(8)
H2G_F64, H2G_C64, H2G_F32, H2G_C32 = range(4)
H2G_OP_N, H2G_OP_T, H2G_OP_C = range(3)
H2G_NEARMUL, H2G_FARMUL, H2G_MUL = 1, 2, 3
H2G_M2L_DENSE, H2G_M2L_TOEPLITZ, H2G_M2L_SYMDENSE, H2G_M2L_SYMTOEPLITZ = range(4)
H2G_PREC_PLAIN, H2G_PREC_MIXED, H2G_PREC_BF16X2, H2G_PREC_BF16X3 = 0, 1, 2, 3
PRECISIONS = {"plain": H2G_PREC_PLAIN, "mixed": H2G_PREC_MIXED, "bf16x2": H2G_PREC_BF16X2, "bf16x3": H2G_PREC_BF16X3}
H2G_HOST_PTRS, H2G_ASYNC = 1, 2

EXPORTED = [
    "h2g_core_create", "h2g_core_retain", "h2g_core_release", "h2g_core_size", "h2g_core_device_bytes",
    "h2g_ws_create", "h2g_ws_destroy", "h2g_ws_device_bytes", "h2g_ws_stream", "h2g_ws_acquire",
    "h2g_ws_release", "h2g_matvec", "h2g_ws_sync", "h2g_ws_timing", "h2g_last_error", "h2g_version",
    "h2g_method_fill", "h2g_ws_query", "h2g_matvec_noalloc", "h2g_ws_set_stream", "h2g_core_summary",
    "h2g_core_sanitycheck", "h2g_plan_build", "h2g_plan_info", "h2g_plan_copy", "h2g_plan_destroy",
    "h2g_method_operators",
]

VP = C.c_void_p
I64P = C.POINTER(C.c_int64)
I32P = C.POINTER(C.c_int32)
U8P = C.POINTER(C.c_uint8)
F64P = C.POINTER(C.c_double)


class SideDesc(C.Structure):
    _fields_ = [
        ("n_dofs", C.c_int64), ("n_levels", C.c_int32), ("n_boxes", C.c_int32),
        ("perm", I64P), ("level_offsets", I32P), ("parent", I32P), ("sector", I32P),
        ("dof_begin", I64P), ("dof_end", I64P), ("is_leaf", U8P),
        ("collection", VP), ("transfers", VP), ("coord", I32P),
    ]


class CoreDesc(C.Structure):
    _fields_ = [
        ("dtype", C.c_int32), ("K", C.c_int32), ("p", C.c_int32), ("galerkin", C.c_int32),
        ("test", SideDesc), ("trial", SideDesc),
        ("m2l_kind", C.c_int32), ("n_pairs", C.c_int64),
        ("pair_test", I32P), ("pair_trial", I32P), ("pair_op", I32P), ("pair_sym", I32P),
        ("n_ops", C.c_int64), ("ops", VP), ("n_perms", C.c_int32), ("perms", I32P),
        ("n_near_blocks", C.c_int64), ("near_test", I32P), ("near_trial", I32P), ("near_offset", I64P),
        ("n_near_values", C.c_int64), ("near_values", VP),
        ("alpha_re", C.c_double), ("alpha_im", C.c_double),
        ("precisions", C.c_int32), ("device", C.c_int32),
    ]


class MethodSideDesc(C.Structure):
    _fields_ = [
        ("points", F64P),          # (n_dofs, 3) original order
        ("box_center", F64P),      # (n_boxes, 3)
        ("box_halfwidth", F64P),   # (n_boxes,)
    ]


class MethodDesc(C.Structure):
    _fields_ = [
        ("kernel", C.c_int32), ("wavenumber", C.c_double),
        ("p", C.c_int32), ("nodes", F64P),
        ("test", MethodSideDesc), ("trial", MethodSideDesc),
        ("fill_near", C.c_int32), ("fill_collections", C.c_int32),
    ]


class WsOpts(C.Structure):
    _fields_ = [("precision", C.c_int32), ("max_rhs", C.c_int32), ("use_graph", C.c_int32), ("timing", C.c_int32),
                ("stream", VP), ("m2l_groups", C.c_int32), ("near_cols", C.c_int32), ("persistent", C.c_int32),
                ("dense_pairs", C.c_int32)]


class TimingRecord(C.Structure):
    _fields_ = [("phase", C.c_int32), ("level", C.c_int32), ("ms", C.c_float), ("mode", C.c_int32),
                ("node", C.c_int32)]


class Summary(C.Structure):
    _fields_ = [("rows", C.c_int64), ("cols", C.c_int64), ("dtype", C.c_int32), ("m2l_kind", C.c_int32),
                ("galerkin", C.c_int32), ("n_levels", C.c_int32), ("n_nodes", C.c_int32), ("n_leaves", C.c_int32),
                ("min_leaf_dofs", C.c_int64), ("max_leaf_dofs", C.c_int64), ("mean_leaf_dofs", C.c_double),
                ("far_pairs", C.c_int64), ("unique_translations", C.c_int64), ("reuse", C.c_double),
                ("near_blocks", C.c_int64), ("near_values", C.c_int64), ("lattice_levels", C.c_int32),
                ("dense_mfma", C.c_int32), ("device_bytes", C.c_int64)]


class OpsDesc(C.Structure):
    _fields_ = [("kernel", C.c_int32), ("wavenumber", C.c_double), ("p", C.c_int32), ("nodes", F64P),
                ("root_size", C.c_double), ("m2l_kind", C.c_int32), ("n_keys", C.c_int64), ("keys", I32P),
                ("device", C.c_int32)]


class OpsOut(C.Structure):
    _fields_ = [("ops", VP), ("transfers", F64P), ("perms", I32P)]


class PlanOpts(C.Structure):
    _fields_ = [("origin", C.c_double * 3), ("size", C.c_double), ("leaf_size", C.c_int32),
                ("max_level", C.c_int32), ("symmetric", C.c_int32), ("device", C.c_int32)]


class PlanCounts(C.Structure):
    _fields_ = [("n_dofs", C.c_int64 * 2), ("n_levels", C.c_int32 * 2), ("n_boxes", C.c_int32 * 2),
                ("n_pairs", C.c_int64), ("n_keys", C.c_int64), ("n_near_pairs", C.c_int64),
                ("n_near_values", C.c_int64), ("galerkin", C.c_int32)]


class PlanTreeOut(C.Structure):
    _fields_ = [("perm", I64P), ("level_offsets", I64P), ("level", I32P), ("coord", I32P),
                ("key", C.POINTER(C.c_uint64)), ("parent", I32P), ("sector", I32P), ("begin", I64P),
                ("end", I64P), ("is_leaf", U8P)]


class PlanOut(C.Structure):
    _fields_ = [("test", PlanTreeOut), ("trial", PlanTreeOut), ("pair_test", I32P), ("pair_trial", I32P),
                ("pair_op", I32P), ("pair_sym", I32P), ("keys", I32P), ("near_test", I32P),
                ("near_trial", I32P), ("near_offset", I64P)]


class Issue(C.Structure):
    _fields_ = [("severity", C.c_int32), ("code", C.c_int32), ("count", C.c_int64), ("level", C.c_int32)]


_lib = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libh2g.so not found at {LIB_PATH}: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.h2g_core_create.argtypes = [C.POINTER(CoreDesc), C.POINTER(VP)]
        L.h2g_core_retain.argtypes = [VP]
        L.h2g_core_release.argtypes = [VP]
        L.h2g_core_size.argtypes = [VP, I64P, I64P]
        L.h2g_core_device_bytes.argtypes = [VP, I64P]
        L.h2g_ws_create.argtypes = [VP, C.POINTER(WsOpts), C.POINTER(VP)]
        L.h2g_ws_destroy.argtypes = [VP]
        L.h2g_ws_device_bytes.argtypes = [VP, I64P]
        L.h2g_ws_stream.argtypes = [VP]
        L.h2g_ws_stream.restype = VP
        L.h2g_ws_set_stream.argtypes = [VP, VP]
        L.h2g_core_summary.argtypes = [VP, C.POINTER(Summary)]
        L.h2g_core_sanitycheck.argtypes = [VP, C.POINTER(Issue), C.c_int32, I32P]
        L.h2g_ws_acquire.argtypes = [VP]
        L.h2g_ws_release.argtypes = [VP]
        L.h2g_matvec.argtypes = [VP, C.c_int, C.c_int, C.c_int, C.c_int64, VP, C.c_int64, VP, C.c_int64,
                                 VP, VP, C.c_int]
        L.h2g_matvec_noalloc.argtypes = [VP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int64, VP, C.c_int64, VP,
                                         C.c_int64, VP, VP, VP, C.c_int]
        L.h2g_ws_sync.argtypes = [VP]
        L.h2g_ws_query.argtypes = [VP]
        L.h2g_ws_timing.argtypes = [VP, C.POINTER(TimingRecord), C.c_int32, I32P]
        L.h2g_last_error.argtypes = [C.c_char_p, C.c_size_t]
        L.h2g_method_fill.argtypes = [VP, C.POINTER(MethodDesc)]
        L.h2g_plan_build.argtypes = [F64P, C.c_int64, F64P, C.c_int64, C.POINTER(PlanOpts), C.POINTER(VP)]
        L.h2g_plan_info.argtypes = [VP, C.POINTER(PlanCounts)]
        L.h2g_plan_copy.argtypes = [VP, C.POINTER(PlanOut)]
        L.h2g_plan_destroy.argtypes = [VP]
        L.h2g_method_operators.argtypes = [C.POINTER(OpsDesc), C.POINTER(OpsOut)]
        _lib = L
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(4096)
    lib().h2g_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


class H2GError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"h2g error {code}: {msg}")
        self.code = code


def check(code: int):
    if code != H2G_OK:
        raise H2GError(code, last_error())


def ptr(a: np.ndarray, ctype=None):
    """Pointer to a contiguous numpy array (the caller keeps the array alive)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]
    p = a.ctypes.data_as(C.c_void_p)
    return C.cast(p, C.POINTER(ctype)) if ctype is not None else p
