"""CPU oracle for the TransiFlow hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in numpy/scipy, the reference algorithms that the HIP
backend replaces: the finite-volume Discretization and BoundaryConditions
(``transiflow/Discretization.py``, ``transiflow/BoundaryConditions.py``),
the state layout (``transiflow/utils.py``), CSR assembly and compression
(``transiflow/CrsMatrix.py``) and the SciPy backend's solve with its pressure
gauge and border (``transiflow/interface/SciPy.py``), plus a numpy FP64
convolution used to check the learned preconditioner's kernels.

It exists only to check the product.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
may import it, and only as the checker or the timed CPU baseline -- never as
a fallback of the product path, which fails loudly when its HIP library is
missing.

Pinning (what checks each function against something other than itself):

* discretization.py linear/nonlinear atoms, boundary conditions, rhs and
  jacobian -- the reference's golden matrices and rhs vectors in
  ``tests/golden/reference_data`` (exact ``==`` where the reference test is
  exact), the 2D == 3D(nz=1) equivalence, finite-difference Jacobian
  consistency and the zero-diagonal invariant on divergence-free states;
* crs.py -- the same golden matrices (sorted, zero-dropped rows) and a
  brute-force per-row assembler on tiny grids;
* layout.py -- closed-form coordinate-vector tests and the index formula;
* interface.py solve -- residual identities and scipy's dense solver on
  tiny systems; bordered matrix identity;
* conv.py -- scipy.signal.correlate on single-channel layers, a brute-force
  loop convolution on tiny inputs, finite differences of forward vs backward
  and BF16 rounding against known bit patterns.
"""
