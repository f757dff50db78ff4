"""FP64 CPU oracle for the density-assignment -> P(k)/B(k) path.

TEST INFRASTRUCTURE ONLY. Nothing in the product path (``pylians_amd``) imports,
calls or links anything in this package.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it,
and only as the checker / CPU baseline, never as the thing that is measured.

The oracle restates the reference algorithm (Pylians, Python-2 ``library/`` layout)
in float64 arithmetic:

* ``oracle.mas``  -- MAS_library.MA and the NGP/CIC/TSC/PCS(+W, 2-D, _d) stencils,
  MAS_library.pyx:57-545, and the OpenMP MAS_c.c variants, MAS_c.c:8-266.
* ``oracle.pk``   -- Pk_library helpers and the ``Pk`` estimator,
  Pk_library.pyx:59-181 and 266-425.
* ``oracle.bk``   -- the ``Bk`` estimator, bispectrum_library.pyx:32-196.
* ``oracle.adjoint`` -- FP64 analytic gradient of sum_k w_k P0(k) w.r.t. positions
  (builder-defined; no reference counterpart).
* ``oracle.cpu`` -- ctypes front-end of the C restatements and of the
  reference's MAS_c.c built from source into ``oracle/_ref`` (``oracle/Makefile``).

How each function is pinned (golden vectors, known-answer tests, brute force,
the compiled reference) is listed in each module's header and exercised by
``tests/test_oracle_*.py``.
"""
