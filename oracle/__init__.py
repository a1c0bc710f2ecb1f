# -*- coding: utf-8 -*-
"""CPU oracle for the batched GGA forward solve and its implicit adjoint.

TEST INFRASTRUCTURE ONLY.  This package is a float64 numpy/scipy
restatement of the reference's CPU path (dgga/solver.py ``GGASolver.solve``
in dense mode, and dgga/autodiff.py's CPU and reduced adjoints).  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker.  The product package
``wdsgpu_amd`` never imports it: its solver runs on the HIP library and
fails loudly when that library is missing.

Pinning (how the oracle itself is checked; see tests/test_oracle.py):
  * coefficient formulas: closed forms (single pipe, series pipes, textbook
    SI Hazen-Williams), derivative identities, brute-force finite differences;
  * dense GGA: converged solutions satisfy every link energy equation,
    emitter equation and junction mass balance; batch == single bitwise;
    CSR assembly == dense assembly bitwise; iteration counts stable;
  * the reduced adjoint equals the full-Jacobian adjoint solved with scipy
    ``splu`` (an independent library routine) and central finite
    differences;
  * Woodbury row replacement equals a dense solve of the replaced matrix.
"""
