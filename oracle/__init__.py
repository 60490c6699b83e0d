"""CPU oracle for the DMRG sweep hot path -- TEST INFRASTRUCTURE ONLY.

This package is a plain numpy/scipy restatement of the reference's single-GPU
DMRG sweep (two-site ``DMRG2GPU`` and single-site ``DMRGGPU``) and of the
``cpu/pdmrg`` numpy primitives.  It exists to check the HIP engine and to time
the reference's CPU path.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as a checker or as
the timed CPU baseline.  The product (``libdmrg_amd.so`` and the Python
binding over it) never imports, links or calls anything in here.

Modules
-------
models      MPO builders (Heisenberg, random-field Heisenberg, TFIM, Josephson)
            in the reference GPU layout, plus independent dense Hamiltonians
            built from Kronecker products for exact diagonalisation.
prng        counter-based splitmix64 generator shared bit-exactly with the
            engine's ``dmrg_init_mps_random``.
dmrg        two-site / single-site sweep restated from the reference GPU
            engines (Lanczos with full reorthogonalisation, truncation rules,
            sweep order, returned energy).
pdmrg_serial  the reference CPU path (tensordot H_eff + ARPACK eigsh +
            numpy SVD), used as the timed CPU baseline.

The reader/writer of the reference's ``.bin`` MPS/MPO format is product code
(``dmrg_amd/io.py``), not part of the oracle.

Pinning status of each function is listed in the module docstrings; every
function here is checked in ``tests/test_oracle_*.py`` against the
reference's known answers, einsum identities, exact diagonalisation or
invariants.
"""
