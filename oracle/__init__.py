"""CPU oracle for the hierarchical (H2) matvec -- TEST INFRASTRUCTURE ONLY.

This package restates, in plain NumPy (FP64 / complex128), the algorithm that
H2Factory.jl runs for ``mul!(y, A, x)`` on an H2 map, so that the HIP product path can
be checked against it.  It is imported only by ``tests/``, by
``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of ``bench.py``; the product
library never imports, links or calls anything here.

Modules
-------
tree      octree over the unit cube, near/far partition and M2L interaction lists
          (H2Trees is not vendored; this is the build's own definition, see the module).
method    interpolation method data: Chebyshev / equispaced tensor nodes, leaf
          collections (P2M/L2P), level transfers (M2M/L2L), Laplace / Helmholtz
          kernels, dense M2L matrices, Toeplitz difference tensors and spectra,
          the 48-element lattice symmetry group and its permutations.
h2        the core (test/trial sides, couplings, near blocks) and the traversal:
          near SpMV, P2M, M2M, M2L (dense / Toeplitz / symmetry-reduced), L2L, L2P,
          with forward / transpose / adjoint routing and the 3- and 5-argument
          ``mul!`` / ``farmul!`` / ``nearmul!`` rules; plus a brute-force dense
          assembly of the same map (the ``Matrix(A)`` pin) and the O(N^2) direct sum.

Pinning status of each function is stated in its module header.
"""
