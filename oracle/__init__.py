"""CPU FP64 oracle for the ZMPO-DMRG two-site sigma / renormalisation path.

THIS PACKAGE IS TEST INFRASTRUCTURE. Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the
checker. The product package ``zmpo_dmrg_amd`` never imports it.

It is a numpy restatement of the reference (zhendongli2008/zmpo_dmrg, Python 2)
algorithm for this path; every function cites the reference file:line it
follows (paths relative to the reference root).

Pinning (what checks each module against something other than itself):
  qtensor.blks_allowed     -- compiled reference ``ctypes/qsym.c`` (oracle/_ref)
                              and brute force over the closed-form rule.
  qtensor.tensordot*       -- dense ``numpy.tensordot`` of toDense() operands.
  mpo (dense W factors)    -- the reference's hand-written matrices
                              (qtensor_opers.py:535-562) and the brute-force
                              Fock-space Hamiltonian (oracle.fci) at K<=4.
  kernels (HVec/HDiag/renorm/genHops)
                           -- dense H_eff built from the full MPO contracted
                              with the MPS (Hermiticity, diag == HDiag,
                              <psi|H|psi> == Davidson eigenvalue).
  davidson                 -- scipy.linalg.eigh on the explicit matrix.
  fci.sector_fci_energy    -- fci.fci_energy (Fock-space build) at K = 4, and the
                              BLOCK energy the reference records for its FCIDUMP
                              case (itools/fromFCIDUMP/eBLOCK:1) to 6e-9.
  fci.fci_rdms             -- E = sum h gamma + 1/2 sum (pr|qs) Gamma, tr gamma = N,
                              and central differences of the FCI energy
                              (Hellmann-Feynman) along h and symmetric eri.
  sweep                    -- exact FCI (oracle.fci) at K<=6 with D >= full
                              dimension, and the H6/STO-3G FCI energy recorded
                              in the reference (samps/example_conversion/
                              0.h6_nqt/README:2).
  decimation tie rule      -- parity unpinned against the reference itself (its
                              rule depends on Python-2 dict order and unstable
                              argsort, mpo_dmrg_qparser.py:220-233,259); the
                              oracle and the device path share the documented
                              deterministic rule below and are pinned against
                              each other and against FCI.
"""
