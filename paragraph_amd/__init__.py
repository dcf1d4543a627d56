"""MI355X-native ParaGraph GPU index-build stage.

The compute path is libparagraph_hip.so (hand-written HIP kernels for
gfx950 behind the C ABI in include/paragraph.h).  This package holds the
ctypes binding, the Builder mirror of the reference interface, the
synthetic-workload generator, index I/O and the recall evaluator.
"""
from ._native import (Context, PgError, PgParams, PgStats, make_params,  # noqa: F401
                      check_params, PG_FP32, PG_BF16, PG_FP8E4M3, TOMB)
from .builder import Builder  # noqa: F401

__all__ = ["Builder", "Context", "PgError", "PgParams", "PgStats",
           "make_params", "check_params", "PG_FP32", "PG_BF16", "PG_FP8E4M3",
           "TOMB"]
