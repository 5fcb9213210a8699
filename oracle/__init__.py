"""CPU oracle for the L-sweeps preconditioned Helmholtz solve path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, its C++
host classes, the Python ctypes binding) imports or calls this package.  It is
used by ``tests/``, by ``__graft_entry__.smoke()`` as the checker, and by
``bench.py`` to time the CPU baseline.

It restates, in numpy/scipy, the algorithm of the reference L-Sweeps-2D code:

* ``geometry``  -- DDCheckerboard::setUp, AssemblerHelmholtz2DFD grids,
  SubdomainStructured index sets, FD/PML assembly, media, point-source RHS;
* ``sweeps``    -- DDCheckerboard::applyPrecond / applySweep and the nine
  sweep loops, applySystemMatrix;
* ``gmres``     -- GlobalIterativeSolverGMRES (reference mode, with its
  quirks) and a textbook flexible GMRES used for the mixed-precision path;
* ``output``    -- the point, element and value layout of OutputVTK::print2D.

Local sparse LU is SuperLU through ``scipy.sparse.linalg.splu`` with COLAMD
ordering, the same algorithm family as the Eigen SparseLU the reference calls
(src/LocalLinearSolverEigenSparseLU.cpp:50,75).

Pinning (how each function is checked beyond restating it) is listed in the
docstring of each module; the checks live in ``tests/test_oracle_*.py``.
"""
