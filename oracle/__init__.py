"""CPU oracle: FP64 restatement of the reference's DF Fock build, SCF loop and
SOS-AO-MP2 algorithms.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (chem_amd/, include/) may
import or call this package; only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py use it, as the checker.

Every function cites the reference file:line it follows.  The reference itself
cannot be built here (DBCSR, libcint, ScaLAPACK and a Fortran compiler are
absent), so these restatements are pinned by closed forms, invariants and
alternative formulations in tests/ (see the header of each module for what
pins it, and where a part is "parity unpinned").
"""
