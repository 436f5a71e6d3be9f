"""MI355X-native DVC per-point correlation path (CCPi DVC Search / Interpolate /
ObjectiveFunctions / strain fit) behind a C ABI (include/dvc_amd.h).

The compute path is libdvc_amd.so (hand-written gfx950 HIP).  This package holds its
sources (csrc/), the C++ host programs (host/), a ctypes binding (_lib) and the
synthetic-data generator used by tests and benchmarks (synth).
"""
from ._lib import (CONVG_FAIL, ERR_RANGE, NOT_SEARCH, POINT_GOOD, RANGE_FAIL, RESULT_DTYPE, START_CHAIN, START_GIVEN,
                   START_JUDGED, Context, DeviceArray, DvcError, exported_symbols, knn_order_neighbors, load, make_params,
                   sort_order_neighbors)

__all__ = ["Context", "DeviceArray", "DvcError", "make_params", "sort_order_neighbors", "knn_order_neighbors", "load", "exported_symbols",
           "RESULT_DTYPE", "ERR_RANGE", "POINT_GOOD", "RANGE_FAIL", "CONVG_FAIL", "NOT_SEARCH", "START_CHAIN", "START_GIVEN",
           "START_JUDGED"]
