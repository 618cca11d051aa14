"""CPU FP64 oracle for torch-fem's assemble-and-solve path.

TEST INFRASTRUCTURE ONLY. This package restates the reference's algorithm
(meyer-nils/torch-fem v0.9.0, file:line citations in each function) in plain
torch-on-CPU / numpy / SciPy FP64 so that the HIP path can be checked against
it. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import it, and only as the checker: the product path (torch-fem_amd/) never
imports, calls or falls back to anything here.

Pinning (what fixes each function independently of its own formulas) is
listed in tests/test_oracle.py; see the module docstrings for the reference
lines each function follows.
"""

from .fem_oracle import *  # noqa: F401,F403
