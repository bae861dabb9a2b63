"""FP64 CPU oracle for ugacc's CCSD and (T) (test infrastructure only).

Only the test suite, the smoke check and the CPU-baseline leg of the
benchmark import this package, and only as the checker or as the timed CPU
baseline.  The product library never calls it.
"""
from .oracle import Oracle, build, libpath  # noqa: F401
