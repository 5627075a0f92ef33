"""CPU oracle: a restatement of the reference's torch eager semantics.

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py``; never by the ``stenpy_amd`` package.  ``ops`` restates each
operator (FP64 oracle or FP32 twin by input dtype), ``graph_eval`` runs whole graphs.

Pinning (tests/test_oracle.py): every function is checked against closed forms (discrete
Fourier symbols, exact derivatives of low-order polynomials, exact Simpson integrals), library
routines (numpy.linalg, math.fsum) or brute-force loops on tiny grids.  The reference itself
cannot be imported here (h5py and zict are missing) and ships no tests or fixtures.
"""
