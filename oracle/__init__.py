"""CPU oracle for the path-tracing hot path (TEST INFRASTRUCTURE ONLY).

A plain-C restatement of the reference shader code (oracle/oracle.c), bound
with ctypes.  Only tests/, the smoke check and the CPU-baseline leg of the
benchmark use it, as the checker; the product library never calls it.
"""
from .pyoracle import Oracle, OracleStats, build_if_missing, load  # noqa: F401
