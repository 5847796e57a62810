"""MI355X-native batched LUP-determinant engine (the LU-fqhe hot path).

The package directory name contains a hyphen, so import it with
``importlib.import_module("lion-codes_amd")``.
"""
from .lucore import (  # noqa: F401
    ARITHS,
    POLICIES,
    DetResult,
    LufqheError,
    alloc_result,
    det_batched,
    det_inplace,
    global_reduction,
    load,
    pair_sum,
)
