"""CPU oracle for the near-Toeplitz tridiagonal hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped package imports, links or
calls this package: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker (or, for
``cpu_baseline``, as the timed CPU comparator).  The product path runs the HIP
kernels and fails loudly when they are missing.

Modules
-------
``cr``         near-Toeplitz cyclic reduction restated in NumPy FP64
               (reference ``neato.py:199-274`` coefficients, ``neato.py:34-123``
               phase order), generalised to any n >= 2; SciPy banded solve
               (reference ``parallelcompact/misc.py:5-19``).
``compact``    compact 4th-order RHS stencil (``compact/rhs.py:21-36``),
               permute (``compact/permute.py:15-19,45-46``), full derivative.
``adi``        Peaceman-Rachford ADI: explicit RHS (``adi/rhs.py:21-24,46-49``),
               transpose (``adi/transpose.py``), the step loop
               (``experiments/run-adi.py:40-71``) and its closed-form eigenmode
               amplification factor.
``partition``  the partition (superposition) method: secondary solves
               (``parallelcompact/misc.py:21-46``), reduced system
               (``parallelcompact/reduced.py:163-178``), correction sum
               (``parallelcompact/sum.py:30``), ghosted RHS
               (``parallelcompact/rhs.py:33-52``).
``fields``     the seeded synthetic inputs of the measured workloads.

Parity pins (each exercised in ``tests/test_oracle.py``)
-------------------------------------------------------
* CR solve: SciPy ``solve_banded`` (LAPACK gtsv) of the same matrix for many n
  (power-of-2 and not) and coefficient sets; dense ``numpy.linalg.solve`` brute
  force for n <= 16.
* level coefficients: hand values of the compact matrix, agreement with an
  independent literal restatement of ``neato.py:241-272``, and a numerical
  elimination step applied to the dense matrix.
* compact RHS / derivative: exactness on cubic polynomials, the reference's own
  known answer f = x*y*z (``experiments/run-parallel-compact.py:89-90``),
  4th/3rd-order convergence ratios (``experiments/run-compact-errors.py``).
* ADI: per-line SciPy solves of each half step, and the closed-form
  amplification factor of a discrete sine eigenmode.
* partition method: equality with the monolithic SciPy solve of the whole line.
* permute / transpose: equality with ``numpy.transpose``.

No oracle function is "parity unpinned".
"""
