"""The C ABI library loads, exports every entry point include/dvc_amd.h declares,
refuses to run without a gfx950 device, and its host-side ordering matches the oracle
(DataCloud::sort_order_neighbors, Core/DataCloud.cpp:69-144) bit for bit."""
import ctypes
import json
from pathlib import Path

import numpy as np
import pytest

import digitalvolumecorrelation_amd as D
from digitalvolumecorrelation_amd import synth


def test_library_exports_header_symbols():
    lib = D.load()
    syms = D.exported_symbols()
    assert len(syms) >= 16 and "dvc_last_error" in syms and "dvc_search_points" in syms
    for s in syms:
        assert hasattr(lib, s), s


def test_struct_layouts_match_header():
    assert ctypes.sizeof(D._lib.RunParams) == 8 * 4 + 8 + 24 + 24
    assert ctypes.sizeof(D._lib.PointResult) == 8 + 8 + 96


def test_sort_matches_oracle_and_known_answer(oracle):
    g = json.loads((Path(__file__).parent / "golden" / "datacloud_10nn.json").read_text())
    order, neigh = D.sort_order_neighbors(np.array(g["points"]))
    assert np.array_equal(neigh, np.array(g["neighbours"]))
    for pts in (synth.grid_points(7, 8.0, (50, 50, 50)),
                np.random.default_rng(1).uniform(0, 100, size=(500, 3)).round(1)):
        o1, n1 = D.sort_order_neighbors(pts, start=pts[17])
        o2, n2 = oracle.sort_order_neighbors(pts, start=pts[17])
        assert np.array_equal(o1, o2) and np.array_equal(n1, n2)


def test_no_device_is_an_error():
    # without a gfx950 device dvc_create fails with an error code, never a crash
    try:
        c = D.Context(0)
    except D.DvcError as e:
        assert e.code == -5
        return
    c.close()
    pytest.skip("a GPU is present")
