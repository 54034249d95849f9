"""Interpolation-method data for the oracle -- TEST INFRASTRUCTURE ONLY.

H2Factory defines no physical method (the downstream method package supplies trees, leaf
collections, transfers and couplings), so the method the benchmark workloads name --
tensor-product Lagrange interpolation on Chebyshev or equispaced nodes with the Laplace or
Helmholtz kernel -- is restated here from its mathematics.

Conventions shared with the product library (and with the reference's array layout):

* a moment of order ``p`` is a ``p x p x p`` array stored with the x index fastest
  (Julia column-major, ``src/moments/toeplitz/toeplitzmul.jl:308-315`` copies it into the
  leading corner of the padded array), i.e. linear index ``k = kx + p*ky + p*p*kz``.
  NumPy arrays here are indexed ``[kz, ky, kx]`` so that C-order flattening gives that
  order;
* the Toeplitz difference tensor has ``L = 2p-1`` entries per axis in wrap order
  ``0..p-1, -(p-1)..-1`` and the operator is ``T[i, j] = kappa[mod(i-j, L)]`` per axis
  (``test/H2Toeplitz/test_smoke.jl:27-36``); its spectrum is ``D = fftn(kappa)``;
* a child's sector is ``bx + 2*by + 4*bz`` (bit set = upper half on that axis);
* a lattice symmetry ``S = (perm, signs)`` acts as ``(S v)[i] = signs[i] * v[perm[i]]``
  (the rule stated at
  ``src/farinteractions/i2otranslations/symmetryreducedcollection.jl:243-258``).

Pinning (``tests/test_oracle_method.py``): Lagrange bases reproduce polynomials of degree
< p exactly; transfers equal direct evaluation of the parent basis at child nodes;
Toeplitz spectra give the same operator as the dense kernel matrix on equispaced nodes;
the symmetry group has 48 elements inducing 48 distinct permutations on a ``(2n-1)^3``
tensor, 20 of them self-inverse, and 16 accepted / 2 distinct / 2 self-inverse on a
degenerate ``(2n-1, 1, 1)`` tensor (``test/farinteractions/test_toeplitz3d.jl:107-116``);
permuted spectra and permuted dense matrices equal the directly built ones.
"""

from __future__ import annotations

import itertools

import numpy as np

FOUR_PI = 4.0 * np.pi


# --------------------------------------------------------------------------------------
# Nodes and Lagrange bases
# --------------------------------------------------------------------------------------
def nodes1d(p: int, kind: str) -> np.ndarray:
    """Interpolation nodes on [-1, 1] in increasing order.

    ``kind`` is ``"cheb"`` (first-kind Chebyshev points) or ``"equi"`` (equispaced, the
    grid on which the M2L operator is multilevel Toeplitz).  Both are symmetric about 0,
    which the symmetry-reduced dense storage relies on.
    """
    if kind == "cheb":
        k = np.arange(p)
        return -np.cos((2 * k + 1) * np.pi / (2 * p))
    if kind == "equi":
        if p == 1:
            return np.zeros(1)
        return -1.0 + 2.0 * np.arange(p) / (p - 1)
    raise ValueError(f"unknown node kind {kind!r}")


def lagrange1d(nodes: np.ndarray, t: np.ndarray) -> np.ndarray:
    """Values ``l_k(t)`` of the Lagrange basis on ``nodes``: shape ``(len(t), p)``."""
    t = np.asarray(t, dtype=np.float64)
    p = len(nodes)
    out = np.ones((t.size, p))
    for k in range(p):
        for m in range(p):
            if m != k:
                out[:, k] *= (t - nodes[m]) / (nodes[k] - nodes[m])
    return out


def tensor_basis(nodes: np.ndarray, tpts: np.ndarray) -> np.ndarray:
    """Tensor basis ``S_k(t)`` for points ``tpts`` (n, 3) in reference coordinates.

    Returns shape ``(K, n)`` with ``k = kx + p*ky + p*p*kz``.
    """
    p = len(nodes)
    lx = lagrange1d(nodes, tpts[:, 0])
    ly = lagrange1d(nodes, tpts[:, 1])
    lz = lagrange1d(nodes, tpts[:, 2])
    # [n, kz, ky, kx]
    s = lz[:, :, None, None] * ly[:, None, :, None] * lx[:, None, None, :]
    return s.reshape(tpts.shape[0], p ** 3).T.copy()


def grid_points(nodes: np.ndarray) -> np.ndarray:
    """Reference-coordinate tensor grid (K, 3), ordered x fastest."""
    p = len(nodes)
    kz, ky, kx = np.meshgrid(np.arange(p), np.arange(p), np.arange(p), indexing="ij")
    return np.stack([nodes[kx.ravel()], nodes[ky.ravel()], nodes[kz.ravel()]], axis=1)


# --------------------------------------------------------------------------------------
# Kernels
# --------------------------------------------------------------------------------------
def kernel_matrix(kind: str, wavenumber: float, xt: np.ndarray, ys: np.ndarray):
    """Dense ``G(x_i - y_j)``; Laplace ``1/(4 pi r)``, Helmholtz ``exp(i k r)/(4 pi r)``,
    both 0 at ``r == 0`` (the diagonal of a Galerkin point cloud)."""
    d = xt[:, None, :] - ys[None, :, :]
    r = np.sqrt(np.einsum("ijk,ijk->ij", d, d))
    with np.errstate(divide="ignore", invalid="ignore"):
        if kind == "laplace":
            g = np.where(r > 0, 1.0 / (FOUR_PI * np.where(r > 0, r, 1.0)), 0.0)
        elif kind == "helmholtz":
            rr = np.where(r > 0, r, 1.0)
            g = np.where(r > 0, np.exp(1j * wavenumber * rr) / (FOUR_PI * rr), 0.0)
        else:
            raise ValueError(kind)
    return g


def kernel_of_r(kind: str, wavenumber: float, r: np.ndarray):
    r = np.asarray(r, dtype=np.float64)
    rr = np.where(r > 0, r, 1.0)
    if kind == "laplace":
        return np.where(r > 0, 1.0 / (FOUR_PI * rr), 0.0)
    return np.where(r > 0, np.exp(1j * wavenumber * rr) / (FOUR_PI * rr), 0.0)


# --------------------------------------------------------------------------------------
# Operators of the interpolation method
# --------------------------------------------------------------------------------------
def leaf_collection(nodes, center, halfwidth, pts):
    """P2M/L2P collection ``C[k, j] = S_k((y_j - c)/h)``: shape (K, n)."""
    t = (np.asarray(pts) - np.asarray(center)[None, :]) / halfwidth
    return tensor_basis(nodes, t)


def transfer_1d(nodes, upper: bool) -> np.ndarray:
    """Per-axis factor ``A[k, j] = l_k((sigma + t_j)/2)`` of the parent<-child transfer."""
    sigma = 1.0 if upper else -1.0
    return lagrange1d(nodes, (sigma + nodes) / 2.0).T.copy()


def transfer(nodes, sector: int) -> np.ndarray:
    """Dense M2M matrix ``T_s[k, j]`` = parent basis ``k`` at child node ``j`` (K x K)."""
    ax = transfer_1d(nodes, bool(sector & 1))
    ay = transfer_1d(nodes, bool(sector & 2))
    az = transfer_1d(nodes, bool(sector & 4))
    return np.kron(az, np.kron(ay, ax))


def box_nodes(nodes, center, halfwidth) -> np.ndarray:
    return np.asarray(center)[None, :] + halfwidth * grid_points(nodes)


def m2l_dense(kind, wavenumber, nodes, width, offset) -> np.ndarray:
    """Dense coupling of a receiving box and a translating box ``offset`` boxes away.

    ``offset = coord(translating) - coord(receiving)`` in units of the box width.
    ``M[k, j] = G(x_k - y_j)`` with ``x`` the receiving nodes and ``y`` the translating ones.
    """
    h = width / 2.0
    xt = box_nodes(nodes, np.zeros(3), h)
    ys = box_nodes(nodes, width * np.asarray(offset, dtype=np.float64), h)
    return kernel_matrix(kind, wavenumber, xt, ys)


def wrap_differences(L: int) -> np.ndarray:
    """Signed difference held by each wrap-order index ``0..L-1`` (``0..p-1, -(p-1)..-1``)."""
    i = np.arange(L)
    return np.where(i <= L // 2, i, i - L)


def toeplitz_kappa(kind, wavenumber, p, width, offset) -> np.ndarray:
    """Difference tensor ``kappa[dz, dy, dx]`` (wrap order) of the equispaced M2L.

    With node spacing ``delta = width / (p-1)``, ``x_k - y_j = delta*((k-j) - (p-1)*offset)``,
    so ``kappa[d] = G(delta*(d - (p-1)*offset))``.
    """
    L = 2 * p - 1
    delta = width / (p - 1)
    d = wrap_differences(L).astype(np.float64)
    o = np.asarray(offset, dtype=np.float64)
    dz, dy, dx = np.meshgrid(d, d, d, indexing="ij")
    vx = delta * (dx - (p - 1) * o[0])
    vy = delta * (dy - (p - 1) * o[1])
    vz = delta * (dz - (p - 1) * o[2])
    return kernel_of_r(kind, wavenumber, np.sqrt(vx * vx + vy * vy + vz * vz))


def toeplitz_dense(kappa: np.ndarray, p: int) -> np.ndarray:
    """Dense ``K x K`` matrix of the multilevel Toeplitz operator ``T[i,j] = kappa[(i-j) mod L]``."""
    L = kappa.shape[0]
    g = np.arange(p)
    kz, ky, kx = np.meshgrid(g, g, g, indexing="ij")
    kx, ky, kz = kx.ravel(), ky.ravel(), kz.ravel()
    ix = (kx[:, None] - kx[None, :]) % L
    iy = (ky[:, None] - ky[None, :]) % L
    iz = (kz[:, None] - kz[None, :]) % L
    return kappa[iz, iy, ix]


# --------------------------------------------------------------------------------------
# Toeplitz application (pad -> FFT -> Hadamard -> iFFT -> crop)
# --------------------------------------------------------------------------------------
def pad(m: np.ndarray, p: int, L: int) -> np.ndarray:
    """Copy a length-K moment into the leading corner of a zero ``L^3`` array
    (``src/moments/toeplitz/toeplitzmul.jl:290-297,308-311``)."""
    buf = np.zeros((L, L, L), dtype=np.complex128)
    buf[:p, :p, :p] = np.asarray(m).reshape(p, p, p)
    return buf


def crop(buf: np.ndarray, p: int) -> np.ndarray:
    """Leading ``p^3`` corner back to a length-K vector (``toeplitzmul.jl:299-305,313-315``)."""
    return buf[:p, :p, :p].reshape(p ** 3).copy()


def modulate_in(m, p, L, op: str):
    """Forward and adjoint use the forward FFT plan; transpose swaps it for the inverse one
    (``src/farinteractions/i2otranslations/toeplitzi2ocollection.jl:248-272``)."""
    b = pad(m, p, L)
    return np.fft.ifftn(b) if op == "T" else np.fft.fftn(b)


def modulate_out(spec, p, op: str):
    """Inverse FFT (normalised) and crop; transpose uses the forward FFT instead."""
    return crop(np.fft.fftn(spec) if op == "T" else np.fft.ifftn(spec), p)


def hadamard_transform(D, op: str):
    """Adjoint reads the coupling conjugated (``toeplitzmul.jl:133-140``)."""
    return np.conj(D) if op == "C" else D


# --------------------------------------------------------------------------------------
# Lattice symmetries (48-element signed-permutation group of the cube)
# --------------------------------------------------------------------------------------
def symmetry_group():
    """All 48 signed coordinate permutations ``(perm, signs)`` of three axes."""
    out = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            out.append((tuple(perm), tuple(signs)))
    return out


def apply_symmetry(sym, v):
    perm, signs = sym
    v = np.asarray(v)
    return np.array([signs[i] * v[perm[i]] for i in range(3)])


def inverse_symmetry(sym):
    perm, signs = sym
    inv = [0, 0, 0]
    for i in range(3):
        inv[perm[i]] = i
    inv_signs = tuple(signs[inv[j]] for j in range(3))
    return (tuple(inv), inv_signs)


def canonical_offset(offset):
    """Canonical orbit representative ``|o|`` sorted ascending, and ``S`` with ``S*c == o``."""
    o = np.asarray(offset, dtype=np.int64)
    a = np.abs(o)
    order = np.argsort(a, kind="stable")
    c = a[order]
    perm = [0, 0, 0]
    for m in range(3):
        perm[order[m]] = m
    signs = tuple(int(-1 if o[i] < 0 else 1) for i in range(3))
    sym = (tuple(int(q) for q in perm), signs)
    assert np.array_equal(apply_symmetry(sym, c), o)
    return tuple(int(q) for q in c), sym


def toeplitz_permutation(size, sym) -> np.ndarray:
    """Linear-index permutation of a wrap-order difference tensor of ``size`` (x, y, z).

    ``perm[j]`` is the wrap index of ``S^-1 d_j``, so ``kappa_{S o}[j] = kappa_o[perm[j]]``
    for a symmetry-invariant kernel, and the same holds for the spectra
    (``symmetryreducedcollection.jl:162-194``).  Raises when ``S`` exchanges axes of
    different length.
    """
    perm, signs = sym
    for i in range(3):
        if size[i] != size[perm[i]]:
            raise ValueError("symmetry exchanges axes of different length")
    inv = inverse_symmetry(sym)
    Lx, Ly, Lz = size
    out = np.empty(Lx * Ly * Lz, dtype=np.int32)
    for iz in range(Lz):
        for iy in range(Ly):
            for ix in range(Lx):
                idx = (ix, iy, iz)
                d = [idx[a] if idx[a] <= size[a] // 2 else idx[a] - size[a] for a in range(3)]
                img = apply_symmetry(inv, d)
                src = [int(img[a]) % size[a] for a in range(3)]
                out[ix + Lx * (iy + Ly * iz)] = src[0] + Lx * (src[1] + Ly * src[2])
    return out


def lattice_grid_permutation(size, sym) -> np.ndarray:
    """Grid permutation: ``perm[i]`` indexes the image of grid point ``i`` under ``S``
    (reversal ``u -> n-1-u`` for a sign flip; ``symmetryreducedcollection.jl:206-232``)."""
    perm, signs = sym
    for i in range(3):
        if size[i] != size[perm[i]]:
            raise ValueError("symmetry exchanges axes of different length")
    nx, ny, nz = size
    out = np.empty(nx * ny * nz, dtype=np.int32)
    for iz in range(nz):
        for iy in range(ny):
            for ix in range(nx):
                src = (ix, iy, iz)
                img = [src[perm[i]] if signs[i] > 0 else size[i] - 1 - src[perm[i]] for i in range(3)]
                out[ix + nx * (iy + ny * iz)] = img[0] + nx * (img[1] + ny * img[2])
    return out
