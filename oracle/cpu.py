"""ctypes wrapper and build recipe of the C/OpenMP restatement in ``oracle/h2cpu.c`` -- TEST
INFRASTRUCTURE ONLY (the CPU baseline and a second checker).  ``build()`` compiles it with gcc into
``oracle/_build/libh2cpu.so``; ``CpuH2`` runs the forward product of a real Laplace Toeplitz map
on the host cores from a host description (the arrays of ``h2g.method.build_map``)."""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libh2cpu.so")


def build():
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    # x86-64-v3 (AVX2 + FMA): runs on the build container and on the GPU boxes' hosts alike
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", os.path.join(HERE, "h2cpu.c"), "-o",
           LIB + ".tmp", "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("gcc failed building the CPU restatement:\n" + r.stderr)
    os.replace(LIB + ".tmp", LIB)
    return LIB


P = C.c_void_p


class Desc(C.Structure):
    _fields_ = [("n", C.c_int64), ("points", P), ("perm", P), ("n_levels", C.c_int32), ("n_boxes", C.c_int32),
                ("level_offsets", P), ("parent", P), ("sector", P), ("begin", P), ("end", P), ("is_leaf", P),
                ("center", P), ("halfwidth", P), ("p", C.c_int32), ("nodes", P), ("transfers", P),
                ("n_pairs", C.c_int64), ("pair_test", P), ("pair_trial", P), ("pair_op", P),
                ("n_ops", C.c_int64), ("spectra", P), ("n_near", C.c_int64), ("near_test", P), ("near_trial", P),
                ("near_offset", P), ("n_near_values", C.c_int64)]


class CpuH2:
    """Forward product y = A x on the host cores (FP64, one column)."""

    def __init__(self, arrays):
        if arrays.dtype != "f64" or arrays.m2l_kind != "toeplitz" or not arrays.galerkin:
            raise ValueError("the CPU restatement runs real Galerkin maps with Toeplitz couplings")
        if not os.path.exists(LIB):
            build()
        self.lib = C.CDLL(LIB)
        self.lib.h2cpu_setup.restype = P
        self.lib.h2cpu_setup.argtypes = [C.POINTER(Desc)]
        self.lib.h2cpu_mul.argtypes = [P, P, P]
        self.lib.h2cpu_free.argtypes = [P]
        self.threads = self.lib.h2cpu_threads()
        t = arrays.test.tree
        c = np.ascontiguousarray
        self._keep = k = dict(
            points=c(t.points, np.float64), perm=c(t.perm, np.int64), lo=c(t.level_offsets, np.int32),
            parent=c(t.parent, np.int32), sector=c(t.sector, np.int32), begin=c(t.begin, np.int64),
            end=c(t.end, np.int64), leaf=c(t.is_leaf, np.uint8), center=c(t.centers(), np.float64),
            hw=c(t.halfwidths(), np.float64), nodes=c(arrays.nodes, np.float64),
            tr=c(arrays.test.transfers, np.float64), pt=c(arrays.pairs[:, 0], np.int32),
            pc=c(arrays.pairs[:, 1], np.int32), po=c(arrays.pair_op, np.int32), ops=c(arrays.ops, np.complex128),
            nt=c(arrays.near_pairs[:, 0], np.int32), nc=c(arrays.near_pairs[:, 1], np.int32),
            no=c(arrays.near_offset, np.int64))
        d = Desc()
        ptr = lambda a: a.ctypes.data_as(P)
        d.n = t.points.shape[0]
        d.points, d.perm = ptr(k["points"]), ptr(k["perm"])
        d.n_levels, d.n_boxes = t.n_levels, t.n_boxes
        d.level_offsets, d.parent, d.sector = ptr(k["lo"]), ptr(k["parent"]), ptr(k["sector"])
        d.begin, d.end, d.is_leaf = ptr(k["begin"]), ptr(k["end"]), ptr(k["leaf"])
        d.center, d.halfwidth = ptr(k["center"]), ptr(k["hw"])
        d.p, d.nodes, d.transfers = arrays.p, ptr(k["nodes"]), ptr(k["tr"])
        d.n_pairs = len(arrays.pairs)
        d.pair_test, d.pair_trial, d.pair_op = ptr(k["pt"]), ptr(k["pc"]), ptr(k["po"])
        d.n_ops, d.spectra = len(arrays.ops), ptr(k["ops"])
        d.n_near = len(arrays.near_pairs)
        d.near_test, d.near_trial, d.near_offset = ptr(k["nt"]), ptr(k["nc"]), ptr(k["no"])
        d.n_near_values = arrays.n_near_values
        self.desc = d
        self.n = d.n
        self.handle = self.lib.h2cpu_setup(C.byref(d))
        if not self.handle:
            raise ValueError("the CPU restatement supports 2 <= p <= 6")

    def mul(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.n)
        self.lib.h2cpu_mul(self.handle, x.ctypes.data_as(P), y.ctypes.data_as(P))
        return y

    def close(self):
        if getattr(self, "handle", None):
            self.lib.h2cpu_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
