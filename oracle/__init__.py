"""FP64 CPU oracle for the MatMRI iterative-SENSE operator chain.

TEST INFRASTRUCTURE ONLY.  This package is a plain restatement, in NumPy
float64/complex128, of the reference MATLAB algorithms on the hot path.  It is
imported only by ``tests/``, by ``__graft_entry__.smoke()`` and by the
``cpu_baseline`` leg of ``bench.py``, and only as the checker.  The product
path (the HIP library behind ``include/matmri_hip.h``) never calls into it.

Modules and the reference files they follow (paths relative to the MatMRI
repository root):

* ``fftnc``      -- ``general/fftnc.m``, ``general/ifftnc.m``
* ``nufft``      -- ``operators/@nufftOp/nufftOp.m:138-408``
* ``rcvr``       -- ``operators/@rcvrOp/rcvrOp.m:104-159``
* ``basis``      -- ``computeHarmonics/basisFuncHarm.m``, ``basisFuncConc.m``
* ``highorder``  -- ``operators/@sampHighOrder/sampHighOrder.m:361-622``
* ``mrsamp``     -- ``operators/mrSampFunc.m:82-159``
* ``solvers``    -- ``iterativeSolvers/cgne.m``, ``bfista.m``,
  ``powermethod.m``, ``general/softthresh.m``
* ``warmstart``  -- the learned warm-start network (no reference counterpart;
  defines the function the device network computes)

How each function is pinned (known-answer tests, closed forms, invariants) is
listed in the docstring of each module and exercised by ``tests/test_oracle_*``.
"""
