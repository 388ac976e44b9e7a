"""CPU oracle for the Heston ADI path -- TEST INFRASTRUCTURE ONLY.

This package is a NumPy/SciPy restatement of the reference's CPU engine
(FiniteDifference/FiniteDifferenceEngine.pyx, FiniteDifference/BandedOperator.pyx
and the Heston coefficient/boundary dictionaries of FiniteDifference/heston.py).
It exists only to check the HIP product path.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker; the product package ``fd_adi_amd`` never
imports, links or executes anything in here.

Pinning (what fixes each function independently of this restatement):

* ``banded.centercoeffs`` -- exact on quadratics on nonuniform meshes
  (tests/test_oracle_operators.py) and the reference test's closed-form
  9-point construction (FiniteDifference/tests/test_BandedOperatorGPU.py:349-420).
* ``banded.BandedOperator.fold_*`` / ``fold_vector`` -- the 5x5 known-answer
  matrix of FiniteDifference/tests/test_BandedOperatorCPU.py:341-353 and the
  elimination-matrix identity of utils.foldMatFor (FiniteDifference/utils.py:505-519).
* ``apply`` / ``solve`` -- dense numpy.linalg brute force on small blocks.
* ``applyboundary`` -- the explicit boundary rows written out in
  FiniteDifference/tests/test_FiniteDifferenceEngine.py:124-208 style
  hand-built matrices and the Dirichlet invariants.
* Heston coefficient fields -- closed forms of heston.py:303-326.
* Full schemes -- a semi-analytic Heston price (``heston_cf``, an independent
  Gauss-Laguerre/quad integration of the characteristic function) at rtol 1e-3,
  the Black-Scholes limit, and the observed time-convergence order
  (1 for implicit, 2 for HV/CS/MCS with theta=1/2).
* ``solve_craigsneyd2`` (modified CS) -- pinned only by the analytic price and
  convergence order, like the other schemes; the reference's GPU engine has
  no counterpart.
"""
