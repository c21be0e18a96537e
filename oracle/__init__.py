"""CPU oracle for the NEVPT2 RDM-contraction hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker or the timed CPU baseline -- never as part of the product path, which
runs on the GPU through ``libnevpt2_amd.so`` and fails loudly without it.

Contents (each function cites the reference file:line it restates):

* ``eexx``   -- accessor closed forms (eexx_functions.F90, qdnevpt.F:6448-6519),
                scalar and as full tensors.
* ``stage1`` -- ctypes entry points over ``c/oracle_stage1.c``, the C/OpenMP
                restatement of the stage-1 Koopmans builds (koop_matrices.F90,
                koopro4QD.F90:3900-5049).
* ``stage2`` -- Fock/canonicalisation, Cholesky transform, the eight perturber
                classes in SC and PC form, Heff (qdnevpt.F).
* ``qd``     -- the QD state-pair couplings and the Heff post-processing.
* ``brute``  -- second-quantisation brute force on tiny active spaces; pins the
                RDM conventions and accessor closed forms independently.
* ``fock``   -- Fock-space brute force (sparse Jordan-Wigner operators): every
                Koopmans matrix as a commutator expectation value, and SC/PC
                NEVPT2 per class and the QD couplings from the first-order
                functions themselves; pins stage1, stage2 and qd
                (tests/test_oracle_brute.py).
* ``ref/``   -- recipe that compiles the reference's own accessor modules
                (eexx_functions.F90, indices_norm.F90, ord_utils.F90,
                swap_util.F90) into ``oracle/_ref`` to pin ``eexx``.

Parity status per function is recorded in each module header.
"""
