"""Core data, the H2 matvec traversal and its brute-force dense counterpart -- TEST INFRASTRUCTURE ONLY.

The traversal restates H2Factory's far product (AggregateMode) and near product:

* entry rules (``src/maps/abstracth2map.jl:54-86``): 3-argument ``mul!`` is
  ``y = near*x`` then the far part added with ``alpha(A)``; 5-argument is
  ``y = a*near*x + b*y`` then ``a*alpha(A)*far*x`` added; ``farmul!`` 3-argument zeroes ``y``
  first and writes ``alpha(A)*far*x`` into initialised leaves; ``farmul!`` 5-argument writes
  ``a*alpha(A)*far*x + b*y`` into initialised leaves only and leaves other rows untouched
  (``src/farinteractions/farmul.jl:47-70``, ``src/maps/functors.jl:272-282``,
  ``src/farinteractions/traversalhelpers.jl:84-112``); ``nearmul!`` is a plain (5-arg) SpMV
  (``src/nearinteractions/nearmul.jl:8-14``);
* upward pass (``src/farinteractions/aggregation.jl:58-134``,
  ``src/farinteractions/traversalhelpers.jl:10-75``): leaf moments ``C*x[leaf]``, parent =
  sum of ``T[level(child)][sector]*child`` (first child overwrites);
* I2O (``src/farinteractions/traversalhelpers.jl:167-219``,
  ``src/farinteractions/i2otranslations/i2oprotocol.jl:87-92``): for every receiving node
  its partners are applied, first overwrite then accumulate, coupling key swapped for
  transpose/adjoint; dense kernel ``op(M)*in``
  (``i2otranslationcollection.jl:351-409``), Toeplitz ``modulate_in``/Hadamard/``modulate_out``
  with the FFT plans swapped for transpose and the spectrum conjugated for adjoint
  (``toeplitzi2ocollection.jl:110-193,248-272``), symmetry-reduced Toeplitz reads
  ``D_canon[perm]`` (``symmetryreducedcollection.jl:33-35``), symmetry-reduced dense gathers
  with the column permutation and scatters with the row one, swapped for transpose/adjoint
  (``symmetryreduceddense.jl:60-93,190-220``);
* downward pass (``src/farinteractions/disaggregation.jl:212-295``): a node's local moment is
  initialised by I2O or, when its parent is initialised, by ``op(T)^T*parent``
  (accumulate if I2O wrote it); transpose for forward/transpose maps, adjoint for the adjoint
  map (``src/maps/accessors.jl:397-416``); leaves with an initialised moment scatter
  ``alpha(A)*a*C^T*l (+ y)``;
* orientation (``src/maps/traits.jl``, ``src/maps/accessors.jl``,
  ``src/maps/translationcontext.jl:57-59``): transpose/adjoint aggregate on the test side
  and disaggregate on the trial side; adjoint conjugates the leaf collections, the I2I
  transfers and ``alpha(A)`` (``src/maps/functors.jl:206-259``,
  ``docs/src/implementing/example.md:144-145``).

``dense_matrix`` assembles the same operator pair by pair from per-box composite bases; it
shares no code with the traversal, so a wrong index, sign, transposition or orientation
rule in one of them shows up as a mismatch (``tests/test_oracle_h2.py``).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import method as M
from .tree import Tree


@dataclass
class Core:
    test: Tree
    trial: Tree
    K: int
    p: int
    dtype: type
    coll_test: np.ndarray            # (K, N_test) tree order
    coll_trial: np.ndarray           # (K, N_trial)
    transfers_test: list             # [level][sector] (K, K), indexed by the CHILD level
    transfers_trial: list
    pairs: np.ndarray                # (P, 2) forward (test box, trial box)
    pair_op: np.ndarray              # (P,)
    m2l_kind: str                    # dense | toeplitz | symdense | symtoeplitz | diagonal
    ops: np.ndarray                  # dense (n, K, K) | spectra (n, L, L, L) | diagonal (n, K)
    pair_sym: np.ndarray = None      # (P,) permutation id (symmetry-reduced kinds)
    perms: np.ndarray = None         # (n_sym, K) grid perms | (n_sym, L^3) spectrum perms
    near_pairs: np.ndarray = None    # (B, 2) (test leaf, trial leaf)
    near_values: list = field(default_factory=list)  # (n_test_leaf, n_trial_leaf) per block
    alpha: complex = 1.0

    @property
    def galerkin(self):
        return self.test is self.trial

    @property
    def L(self):
        return 2 * self.p - 1

    @property
    def shape(self):
        return (self.test.points.shape[0], self.trial.points.shape[0])


# --------------------------------------------------------------------------------------
# Coupling application for one pair, in the orientation ``op``
# --------------------------------------------------------------------------------------
def _dense_coupling(core: Core, k: int):
    """Dense matrix of pair ``k`` in forward orientation (for the dense kinds)."""
    if core.m2l_kind == "dense":
        return core.ops[core.pair_op[k]]
    if core.m2l_kind == "symdense":
        c = core.ops[core.pair_op[k]]
        pi = core.perms[core.pair_sym[k]]
        t = np.empty_like(c)
        t[np.ix_(pi, pi)] = c
        return t
    if core.m2l_kind == "diagonal":
        return np.diag(core.ops[core.pair_op[k]])
    raise ValueError(core.m2l_kind)


def _spectrum(core: Core, k: int):
    D = core.ops[core.pair_op[k]]
    if core.m2l_kind == "symtoeplitz":
        D = D.reshape(-1)[core.perms[core.pair_sym[k]]].reshape(D.shape)
    return D


def _apply_dense_pair(core, k, m, op):
    """``op(M)*m`` following the dense / symmetry-reduced dense / diagonal kernels."""
    if core.m2l_kind == "diagonal":
        # elementwise coupling (i2otranslationcollection.jl:415-495, abstractmoment.jl:420-510):
        # forward and transpose use the coefficients as stored, the adjoint conjugates them
        d = core.ops[core.pair_op[k]]
        d = np.conj(d) if op == "C" else d
        return d[:, None] * m if m.ndim == 2 else d * m
    if core.m2l_kind == "dense":
        Mk = core.ops[core.pair_op[k]]
        if op == "N":
            return Mk @ m
        if op == "T":
            return Mk.T @ m
        return Mk.conj().T @ m
    # symmetry-reduced dense: gather, canonical product, scatter
    c = core.ops[core.pair_op[k]]
    pr = pc = core.perms[core.pair_sym[k]]
    gather, scatter = (pc, pr) if op == "N" else (pr, pc)
    w = m[gather]
    y = (c if op == "N" else (c.T if op == "T" else c.conj().T)) @ w
    out = np.zeros_like(y)
    out[scatter] = y
    return out


# --------------------------------------------------------------------------------------
# Traversal
# --------------------------------------------------------------------------------------
def _sides(core: Core, op: str):
    if op == "N":
        return (core.trial, core.coll_trial, core.transfers_trial,
                core.test, core.coll_test, core.transfers_test)
    return (core.test, core.coll_test, core.transfers_test,
            core.trial, core.coll_trial, core.transfers_trial)


def far_apply(core: Core, x: np.ndarray, op: str = "N"):
    """Far part in tree-independent terms: returns (far*x in original order of the output
    side, boolean mask of output rows whose leaf moment was initialised)."""
    x = np.asarray(x)
    squeeze = x.ndim == 1
    X = x.reshape(x.shape[0], -1)
    R = X.shape[1]
    src, csrc, tsrc, dst, cdst, tdst = _sides(core, op)
    cdt = np.result_type(core.dtype, X.dtype, np.complex128 if core.m2l_kind.endswith("toeplitz") else np.float64)
    xs = X[src.perm]

    conj = (lambda a: np.conj(a)) if op == "C" else (lambda a: a)

    # upward pass
    mom = [None] * src.n_boxes
    for l in range(src.n_levels - 1, -1, -1):
        for b in range(src.level_offsets[l], src.level_offsets[l + 1]):
            if src.is_leaf[b]:
                mom[b] = conj(csrc[:, src.begin[b]:src.end[b]]) @ xs[src.begin[b]:src.end[b]]
            else:
                acc = None
                for c in range(src.child_begin[b], src.child_end[b]):
                    t = conj(tsrc[src.level[c]][src.sector[c]]) @ mom[c]
                    acc = t if acc is None else acc + t
                mom[b] = acc

    # I2O: receiving node -> list of (partner, pair index)
    if op == "N":
        recv_col, part_col = 0, 1
    else:
        recv_col, part_col = 1, 0
    partners = {}
    for k in range(core.pairs.shape[0]):
        partners.setdefault(int(core.pairs[k, recv_col]), []).append((int(core.pairs[k, part_col]), k))

    loc = [None] * dst.n_boxes
    toep = core.m2l_kind in ("toeplitz", "symtoeplitz")
    spec_in = {}
    for r in sorted(partners):
        plist = sorted(partners[r])
        if toep:
            acc = None
            for (t, k) in plist:
                if t not in spec_in:
                    spec_in[t] = np.stack(
                        [M.modulate_in(mom[t][:, j], core.p, core.L, op) for j in range(R)], axis=-1)
                D = M.hadamard_transform(_spectrum(core, k), op)
                term = D[..., None] * spec_in[t]
                acc = term if acc is None else acc + term
            loc[r] = np.stack([M.modulate_out(acc[..., j], core.p, op) for j in range(R)], axis=-1)
        else:
            acc = None
            for (t, k) in plist:
                term = _apply_dense_pair(core, k, mom[t], op)
                acc = term if acc is None else acc + term
            loc[r] = acc

    # downward pass
    init = np.zeros(dst.n_boxes, dtype=bool)
    for r in partners:
        init[r] = True
    for l in range(dst.n_levels):
        for b in range(dst.level_offsets[l], dst.level_offsets[l + 1]):
            pa = dst.parent[b]
            if pa >= 0 and init[pa]:
                T = tdst[dst.level[b]][dst.sector[b]]
                Td = T.conj().T if op == "C" else T.T
                contrib = Td @ loc[pa]
                loc[b] = contrib if loc[b] is None else loc[b] + contrib
                init[b] = True

    n_out = dst.points.shape[0]
    ys = np.zeros((n_out, R), dtype=np.result_type(cdt, core.dtype))
    mask_s = np.zeros(n_out, dtype=bool)
    for b in dst.leaves():
        if init[b]:
            cb = cdst[:, dst.begin[b]:dst.end[b]]
            cb = cb.conj() if op == "C" else cb
            ys[dst.begin[b]:dst.end[b]] = cb.T @ loc[b]
            mask_s[dst.begin[b]:dst.end[b]] = True
    y = np.zeros_like(ys)
    y[dst.perm] = ys
    mask = np.zeros(n_out, dtype=bool)
    mask[dst.perm] = mask_s
    if not toep and np.isrealobj(core.ops) and np.isrealobj(X) and np.isrealobj(core.coll_test):
        y = y.real if np.iscomplexobj(y) else y
    return (y[:, 0] if squeeze else y), mask


def near_apply(core: Core, x: np.ndarray, op: str = "N"):
    x = np.asarray(x)
    squeeze = x.ndim == 1
    X = x.reshape(x.shape[0], -1)
    n_out = core.shape[0] if op == "N" else core.shape[1]
    dt = np.result_type(core.dtype, X.dtype)
    y = np.zeros((n_out, X.shape[1]), dtype=dt)
    for (b, c), B in zip(core.near_pairs, core.near_values):
        rows = core.test.perm[core.test.begin[b]:core.test.end[b]]
        cols = core.trial.perm[core.trial.begin[c]:core.trial.end[c]]
        if op == "N":
            y[rows] += B @ X[cols]
        elif op == "T":
            y[cols] += B.T @ X[rows]
        else:
            y[cols] += B.conj().T @ X[rows]
    return y[:, 0] if squeeze else y


def core_alpha(core: Core, op: str):
    return np.conj(core.alpha) if op == "C" else core.alpha


def mul(core: Core, x, op="N", a=None, b=None, y=None):
    """``mul!(y, op(A), x[, a, b])``."""
    near = near_apply(core, x, op)
    far, _ = far_apply(core, x, op)
    if a is None:
        return near + core_alpha(core, op) * far
    return a * near + b * y + a * core_alpha(core, op) * far


def farmul(core: Core, x, op="N", a=None, b=None, y=None):
    """``farmul!(y, op(A), x[, a, b])`` including the untouched-row rule of the 5-arg form."""
    far, mask = far_apply(core, x, op)
    if a is None:
        return core_alpha(core, op) * far
    out = np.array(y, dtype=np.result_type(y, far, a, b), copy=True)
    sel = mask if out.ndim == 1 else mask[:, None]
    return np.where(sel, a * core_alpha(core, op) * far + b * out, out)


def nearmul(core: Core, x, op="N", a=None, b=None, y=None):
    near = near_apply(core, x, op)
    if a is None:
        return near
    return a * near + b * y


# --------------------------------------------------------------------------------------
# Brute-force dense assembly of the same operator
# --------------------------------------------------------------------------------------
def _out_basis(side: Tree, coll, transfers, b):
    """Rows: side dofs (tree order) of box b's subtree; maps b's local moment to them."""
    if side.is_leaf[b]:
        return coll[:, side.begin[b]:side.end[b]].T
    blocks = []
    for c in range(side.child_begin[b], side.child_end[b]):
        T = transfers[side.level[c]][side.sector[c]]
        blocks.append(_out_basis(side, coll, transfers, c) @ T.T)
    return np.vstack(blocks)


def _in_basis(side: Tree, coll, transfers, b):
    """Maps the dofs of b's subtree to b's moment."""
    if side.is_leaf[b]:
        return coll[:, side.begin[b]:side.end[b]]
    blocks = []
    for c in range(side.child_begin[b], side.child_end[b]):
        T = transfers[side.level[c]][side.sector[c]]
        blocks.append(T @ _in_basis(side, coll, transfers, c))
    return np.hstack(blocks)


def pair_matrix(core: Core, k: int):
    """Dense forward coupling of pair k, built without the FFT path for Toeplitz kinds."""
    if core.m2l_kind in ("dense", "symdense", "diagonal"):
        return _dense_coupling(core, k)
    kappa = np.fft.ifftn(_spectrum(core, k))
    return M.toeplitz_dense(kappa, core.p)


def dense_matrix(core: Core):
    """``Matrix(A)`` in original dof order: near blocks + alpha(A) * sum of far pair blocks."""
    nt, nc = core.shape
    toep = core.m2l_kind in ("toeplitz", "symtoeplitz")
    dt = np.result_type(core.dtype, np.complex128 if toep else core.dtype)
    A = np.zeros((nt, nc), dtype=dt)
    for (b, c), B in zip(core.near_pairs, core.near_values):
        A[np.ix_(core.test.begin[b] + np.arange(B.shape[0]), core.trial.begin[c] + np.arange(B.shape[1]))] += B
    F = np.zeros((nt, nc), dtype=dt)
    for k in range(core.pairs.shape[0]):
        b, c = core.pairs[k]
        W = _out_basis(core.test, core.coll_test, core.transfers_test, b)
        V = _in_basis(core.trial, core.coll_trial, core.transfers_trial, c)
        F[core.test.begin[b]:core.test.end[b], core.trial.begin[c]:core.trial.end[c]] += W @ pair_matrix(core, k) @ V
    A = A + core.alpha * F
    out = np.zeros_like(A)
    out[np.ix_(core.test.perm, core.trial.perm)] = A
    return out


def direct_sum(kind, wavenumber, targets, sources, x, rows=None, chunk=2048):
    """O(N^2) direct sum ``y_i = sum_j G(t_i - s_j) x_j`` (optionally on a subset of rows)."""
    t = targets if rows is None else targets[rows]
    X = np.asarray(x).reshape(len(sources), -1)
    out = []
    for i in range(0, len(t), chunk):
        G = M.kernel_matrix(kind, wavenumber, t[i:i + chunk], sources)
        out.append(G @ X)
    y = np.vstack(out)
    return y[:, 0] if np.asarray(x).ndim == 1 else y
