"""Runs one map through the host replay of libh2g under AddressSanitizer -- TEST INFRASTRUCTURE
ONLY.  Started by tests/test_cpu_replay.py in a child process with the ASan runtime preloaded;
prints one line per product (relative error against the oracle) and exits non-zero on a mismatch.

    python run_case.py CASE
"""

import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "h2factory.jl_amd"), os.path.join(ROOT, "tests")]

import numpy as np  # noqa: E402

import h2g._lib as H  # noqa: E402

H.LIB_PATH = os.path.join(HERE, "_build", "libh2g_replay.so")

import h2g  # noqa: E402
import h2g._lib as _H  # noqa: E402,F401
from convert import to_arrays  # noqa: E402
from oracle import build as OB  # noqa: E402
from oracle import h2 as O  # noqa: E402


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def orient(A, op):
    return {"N": A, "T": h2g.transpose(A), "C": h2g.adjoint(A)}[op]


def helmholtz_sphere():
    """The map of the Helmholtz parity test: 1,500 points on a sphere, k = 10, Chebyshev p = 3,
    dense couplings, leaves of at most 32 points."""
    pts = np.random.default_rng(8).standard_normal((1500, 3))
    pts = 0.5 + 0.45 * pts / np.linalg.norm(pts, axis=1, keepdims=True)
    core = OB.build_core(pts, p=3, nodekind="cheb", m2l_kind="dense", kernel="helmholtz", wavenumber=10.0,
                         leaf_size=32)
    x = np.random.default_rng(2).standard_normal(1500) + 1j * np.random.default_rng(3).standard_normal(1500)
    return core, x


def helmholtz_petrov_alpha():
    """Helmholtz on two point sets (Chebyshev p = 3, dense couplings: the lattice kernels do not
    run here) with a complex kernel factor: real transfers on planar rows, complex collections (the
    complex factor keeps them complex), so P2M and L2P take the complex kernels with the row
    conversions after P2M and before L2P."""
    rng = np.random.default_rng(12)
    a = rng.random((700, 3))
    b = rng.random((500, 3))
    core = OB.build_core(a, b, p=3, nodekind="cheb", m2l_kind="dense", kernel="helmholtz", wavenumber=5.0,
                         leaf_size=24, alpha=0.7 - 0.2j)
    x = rng.standard_normal((500, 2)) + 1j * rng.standard_normal((500, 2))
    return core, x


def random_petrov_symdense():
    """Random complex operators, separate test and trial trees, symmetry-reduced dense couplings
    (the box GEMM with grid permutations), two right-hand sides."""
    rng = np.random.default_rng(7)
    core = OB.random_core(rng.random((260, 3)), rng.random((190, 3)), p=2, m2l_kind="symdense", leaf_size=8,
                          seed=11, alpha=0.8 - 0.3j)
    x = rng.standard_normal((190, 2)) + 1j * rng.standard_normal((190, 2))
    return core, x


def random_petrov_diagonal():
    """Random complex operators with diagonal (elementwise) couplings, separate trees."""
    rng = np.random.default_rng(9)
    core = OB.random_core(rng.random((240, 3)), rng.random((200, 3)), p=2, m2l_kind="diagonal", leaf_size=8,
                          seed=12, alpha=1.1 + 0.4j)
    x = rng.standard_normal((200, 2)) + 1j * rng.standard_normal((200, 2))
    return core, x


def _products(name, A, core, x, ops, fns=("mul",)):
    ok = True
    for op in ops:
        M = orient(A, op)
        n_out = M.shape[0]
        for fn in fns:
            xx = x if op == "N" else _in_for(core, x, op)
            y = np.zeros((n_out,) + xx.shape[1:], dtype=np.result_type(xx, np.complex128 if core.dtype == np.complex128 else np.float64))
            getattr(h2g, fn + "_")(y, M, xx)
            ref = getattr(O, fn)(core, xx, op)
            ref = ref if np.iscomplexobj(y) else ref.real
            e = relerr(y, ref)
            print(f"{name} {fn} op={op}: rel err vs oracle {e:.2e}", flush=True)
            ok &= e < 1e-12
    return ok


def _in_for(core, x, op):
    n_in = core.test.points.shape[0]
    rng = np.random.default_rng(31)
    shp = (n_in,) + x.shape[1:]
    return rng.standard_normal(shp) + (1j * rng.standard_normal(shp) if np.iscomplexobj(x) else 0)


def main(name):
    if name == "device_plan":
        # trees and pair lists of the plan kernels (plan.hip) against the oracle
        import plan_check as PC
        from h2g.plan import device_plan

        for case, a, b, leaf, origin, size, sym in PC.point_sets():
            dp = device_plan(a, b, leaf, origin, size, symmetric=sym)
            PC.check_against_oracle(dp, a, b, leaf, origin, size, sym)
            print(f"{name} {case}: {dp.test.n_boxes} boxes, {len(dp.pairs)} M2L pairs, {len(dp.near_pairs)} near "
                  f"pairs: equal to the oracle (vs oracle 0)", flush=True)
        sys.exit(0)
    if name == "device_operators":
        # couplings, spectra, transfers and permutations of the operator kernels (operators.hip)
        import ops_check as OC

        for case in OC.CASES:
            e = OC.check(*case)
            print(f"{name} {case}: largest deviation vs oracle {e:.2e}", flush=True)
        from h2g import method as PM

        pts = np.random.default_rng(6).random((1500, 3))
        # (the lattice Toeplitz kernels run on the GPU only: a symmetry-reduced dense map here)
        core = OB.build_core(pts, p=3, nodekind="cheb", m2l_kind="symdense", leaf_size=40)
        arrays = PM.build_map(pts, p=3, nodekind="cheb", m2l_kind="symdense", leaf_size=40, plan="device",
                              ops="device", host_values=True)
        c = h2g.GpuH2Core(arrays)
        A = h2g.H2Workspace(c, max_rhs=2)
        x = np.random.default_rng(9).standard_normal((1500, 2))
        ok = _products(name, A, core, x, "NTC", ("mul",))
        A.close()
        c.close()
        sys.exit(0 if ok else 1)
    if name == "laplace_device_fill":
        # near values and leaf collections filled by the library (h2g_method_fill) against the
        # oracle's host assembly on the same points
        from h2g import method as PM

        pts = np.random.default_rng(4).random((1200, 3))
        core = OB.build_core(pts, p=3, nodekind="cheb", m2l_kind="dense", leaf_size=40)
        arrays = PM.build_map(pts, p=3, nodekind="cheb", m2l_kind="dense", leaf_size=40, host_values=False)
        c = h2g.GpuH2Core(arrays, precisions=("plain", "mixed"))
        A = h2g.H2Workspace(c, max_rhs=3)
        x = np.random.default_rng(9).standard_normal((1200, 3))
        ok = _products(name, A, core, x, "NT", ("mul", "nearmul", "farmul"))
    elif name == "diagnostics":
        # h2summary / translation_counts / sanitycheck against the oracle's tree and pair lists
        from h2g import diagnostics as D

        pts = np.random.default_rng(3).random((1500, 3))
        core = OB.build_core(pts, p=3, nodekind="cheb", m2l_kind="symdense", leaf_size=40)
        c = h2g.GpuH2Core(to_arrays(core))
        A = h2g.H2Workspace(c)
        s = D.h2summary(A)
        tr = core.test
        dofs = (tr.end - tr.begin)[tr.is_leaf.astype(bool)]
        uniq = len({(int(o), int(q)) for o, q in zip(core.pair_op, core.pair_sym)})
        ok = (s.size == (1500, 1500) and s.numlevels == tr.n_levels and s.numleaves == int(tr.is_leaf.sum())
              and s.numnodes == tr.n_boxes and s.minleafdofs == dofs.min() and s.maxleafdofs == dofs.max()
              and abs(s.meanleafdofs - dofs.mean()) < 1e-12 and s.farpairs == len(core.pairs)
              and s.uniquetranslations == uniq and s.nearblocks == len(core.near_pairs))
        tc = D.translation_counts(A)
        ok &= tc.farpairs == len(core.pairs) and abs(tc.reuse - len(core.pairs) / uniq) < 1e-12
        ok &= D.sanitycheck(A) == []
        print(f"{name}: summary {s}, counts {tc}, vs oracle {'ok' if ok else 'MISMATCH'}", flush=True)
        A.close()
        c.close()
        # no far pair: a warning; a pair translating a node with itself: an error
        c2 = h2g.GpuH2Core(to_arrays(OB.no_far_core()))
        iss = D.sanitycheck(c2)
        ok &= [i.code for i in iss] == ["no_far_pairs"] and iss[0].severity == "warning"
        bad = OB.random_core(np.random.default_rng(4).random((200, 3)), p=2, m2l_kind="dense", leaf_size=8, seed=1)
        b0 = int(bad.pairs[0, 0])
        bad.pairs = np.vstack([bad.pairs, [[b0, b0]]])
        bad.pair_op = np.concatenate([bad.pair_op, [0]])
        c3 = h2g.GpuH2Core(to_arrays(bad))
        iss = D.sanitycheck(c3)
        ok &= any(i.code == "self_far_pair" and i.severity == "error" and i.count == 1 for i in iss)
        print(f"{name}: no-far {D.sanitycheck(c2)}, self pair {iss}", flush=True)
        c2.close()
        c3.close()
        sys.exit(0 if ok else 1)
    elif name == "noalloc":
        # h2g_matvec_noalloc (buffers in the static scratch) against the workspace product and the
        # oracle: two calls in a row, the second on a smaller map with fewer columns
        import ctypes as C

        ok = True
        for n, nrhs, seed in ((1500, 3, 1), (700, 2, 2)):
            pts = np.random.default_rng(seed).random((n, 3))
            core = OB.build_core(pts, p=3, nodekind="cheb", m2l_kind="dense", leaf_size=40)
            c = h2g.GpuH2Core(to_arrays(core), precisions=("plain", "mixed"))
            A = h2g.H2Workspace(c, max_rhs=nrhs)
            for op, code in (("N", H.H2G_OP_N), ("T", H.H2G_OP_T)):
                x = np.asfortranarray(np.random.default_rng(seed + 5).standard_normal((n, nrhs)))
                y = np.asfortranarray(np.full((n, nrhs), np.nan))
                rc = H.lib().h2g_matvec_noalloc(c.handle, H.H2G_PREC_PLAIN, code, H.H2G_MUL, H.H2G_F64, nrhs,
                                                H.ptr(x), n, H.ptr(y), n, None, None, None, 0)
                assert rc == 0, H.last_error()
                e1 = relerr(y, orient(A, op) @ x)
                e2 = relerr(y, O.mul(core, x, op).real)
                print(f"{name} n={n} op={op}: vs workspace product {e1:.2e}, vs oracle {e2:.2e}", flush=True)
                ok &= e1 < 1e-12 and e2 < 1e-12
            A.close()
        c.close()
        sys.exit(0 if ok else 1)
    else:
        core, x = {"helmholtz_sphere": helmholtz_sphere, "random_petrov_symdense": random_petrov_symdense,
                   "random_petrov_diagonal": random_petrov_diagonal,
                   "helmholtz_petrov_alpha": helmholtz_petrov_alpha}[name]()
        c = h2g.GpuH2Core(to_arrays(core), precisions=("plain", "mixed"))
        A = h2g.H2Workspace(c, max_rhs=2)
        ok = _products(name, A, core, x, "NTC", ("mul", "farmul", "nearmul"))
    A.close()
    c.close()
    sys.exit(0 if ok else 1)


if __name__ == "__main__":
    main(sys.argv[1])
