"""CPU FP64 oracle for the variational-sweep hot path (TEST INFRASTRUCTURE ONLY).

This package is a from-scratch NumPy/SciPy restatement of gcross/Carcassonne's
2D sweep algorithm.  It exists to CHECK the HIP implementation in
``carcassonne_amd``: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker
(or as the timed CPU baseline).  The product never imports it.

Modules and the reference code each one follows:

* ``tags``      -- operator-channel tags and the sparse-dict contraction loop
                   (carcassonne/sparse.py:9-345)
* ``dense``     -- dense index maps: absorptions, stage 1/2/3, dense matrices
                   (carcassonne/tensors/_2d/dense.py:10-225)
* ``sparse2d``  -- tag rules lifting the dense maps (carcassonne/tensors/_2d/sparse.py:12-161)
* ``linalg``    -- normalizeAxis / unitize / enlargener / product compressor
                   (carcassonne/data/__init__.py:43-50,148-150,263-352;
                   carcassonne/utils.py:879-882; carcassonne/compression.py:11-46)
* ``solver``    -- relaxOver (carcassonne/utils.py:805-878), the Multiplier
                   cost rule (utils.py:180-207), and a residual-controlled
                   Lanczos / generalized-eigh reference for the GPU solvers
* ``system``    -- the 2D System, its observables, sweep driver and policies
                   (carcassonne/system/_2d.py, system/base.py, policies.py)
* ``policies``  -- the sweep policies (carcassonne/policies.py:46-242)

Pinning (which check fixes each function) is recorded in each module's
docstring and exercised by tests/test_oracle_dense.py, test_oracle_linalg.py and
test_oracle_system.py; functions without an independent check say "parity
unpinned".
"""
