"""MI355X-native batched LUP-determinant engine (the LU-fqhe hot path).

The package directory name contains a hyphen, so import it with
``importlib.import_module("lion-codes_amd")``.
"""
from .lucore import (  # noqa: F401
    ARITHS,
    POLICIES,
    DetResult,
    LufqheError,
    MixedPlan,
    Server,
    alloc_result,
    cpu_crout_det,
    cpu_pair_sum,
    det_batched,
    apply_shape_table,
    det_inplace,
    file_chunk,
    fqhe_build,
    fqhe_det,
    get_shape,
    global_reduction,
    load,
    mc_lu_run,
    mc_split,
    pair_sum,
    read_matrices,
    set_shape,
    shape_count,
    shape_info,
)
