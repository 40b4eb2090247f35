"""CPU FP64 oracle for the closed-shell coupled-cluster path of ElemCo.jl.

TEST INFRASTRUCTURE ONLY.  This package is the checker, never the product:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  The product path (``elemco.jl_amd``) never routes
through it and fails loudly when its HIP library is missing.

It restates, in numpy, the algorithm of the reference files (paths relative to
the ElemCo.jl v0.14.0 source tree):

* ``src/integrals/dump.jl``       FCIDUMP reader/writer            -> fcidump.py
* ``src/integrals/dumptools.jl``  freeze_orbs_in_dump              -> fcidump.py
* ``src/cc/cc.jl``                dressing, residual, energies,
                                  MP2, iteration loop               -> ccsd.py
* ``src/cc/cctools.jl``           Fock/HF energy, updates, norms    -> ccsd.py
* ``src/solvers/diis.jl``         Pulay DIIS                        -> ccsd.py
* ``src/cc/cc_triples.jl``        (T) and Lambda-(T)                -> triples.py
* ``src/cc/cc_lagrange.jl``       Lambda residual, 1-RDM, Lambda loop -> lagrange.py
* ``src/integrals/dftools.jl``,   DF preparation (AO -> MO factors,
  ``src/tools/tensortools.jl``    sqrtinvchol of the metric)        -> dfprep.py

and generates the seeded synthetic density-fitted systems (synthetic.py).

Parity pins (tests in ``tests/test_oracle.py``):
  - reference-held energies of ``test/h2o.jl`` and ``test/h2o_st1.jl`` (1e-6 Eh):
    HF, MP2, CCSD, CCSD(T), DCSD, Lambda-CCSD(T), Hermitian and ST files;
  - two-electron systems, where CCSD and DCSD equal full CI (closed form) and
    E(T) vanishes (every triplet has i = j = k);
  - brute-force [T] from explicit T3 (the reference's ``cc_tests.jl:54-75``);
  - finite differences of the CC Lagrangian for the Lambda residual and of the
    energy for the 1-RDM;
  - splitmix64 published vectors for the synthetic generator;
  - the DF preparation against its closed form C^T C^T (mu nu|P) pinv((P|Q)) (P|ls) C C
    and M M^T = pinv((P|Q)), full-rank and rank-deficient metrics.
"""
