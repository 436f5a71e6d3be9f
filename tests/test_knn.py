"""Grid k-nearest-neighbour builder (dvc_knn_order_neighbors) against the reference's
full sort per point (DataCloud::sort_order_neighbors, Core/DataCloud.cpp:69-144, as
dvc_sort_order_neighbors) and against a brute-force stable sort: exact on clouds with
distinct distances, ties in index order on lattices, duplicates, planes and lines."""
import time

import numpy as np
import pytest

import digitalvolumecorrelation_amd as D
from digitalvolumecorrelation_amd import synth


def _brute(points, k):
    """k nearest by (distance, index), distances as the reference computes them."""
    d = points[:, None, :] - points[None, :, :]
    dist = np.sqrt(d[..., 0] ** 2 + d[..., 1] ** 2 + d[..., 2] ** 2)
    idx = np.arange(len(points))
    out = np.empty((len(points), k), np.int32)
    for i in range(len(points)):
        out[i] = np.lexsort((idx, dist[i]))[:k]
    return out


def test_random_cloud_equals_full_sort():
    rng = np.random.default_rng(7)
    pts = rng.uniform(0, 200, size=(2500, 3))
    o1, n1 = D.sort_order_neighbors(pts, pts[5])
    o2, n2 = D.knn_order_neighbors(pts, pts[5])
    assert np.array_equal(o1, o2)
    assert np.array_equal(n1, n2)


@pytest.mark.parametrize("shape", ["grid", "plane", "line", "duplicates"])
def test_lattices_ties_in_index_order(shape):
    if shape == "grid":
        pts = synth.grid_points(11, 7.0, (50.0, 50.0, 50.0))
    elif shape == "plane":
        g = np.stack(np.meshgrid(np.arange(30.0), np.arange(25.0), indexing="ij"), -1).reshape(-1, 2)
        pts = np.c_[g * 3.0, np.full(len(g), 9.0)]
    elif shape == "line":
        pts = np.c_[np.arange(400.0) * 0.5, np.zeros(400), np.zeros(400)]
    else:
        rng = np.random.default_rng(2)
        base = rng.uniform(0, 40, size=(300, 3))
        pts = np.r_[base, base[::3], base[::7]]
    for k in (75, 20):
        _, nb = D.knn_order_neighbors(pts, k=k)
        assert np.array_equal(nb, _brute(pts, min(k, len(pts))))
    # the reference's unstable sort may order equal distances differently, never the distances
    _, ref = D.sort_order_neighbors(pts)
    _, nb = D.knn_order_neighbors(pts)
    d = lambda i, j: np.sqrt(((pts[i] - pts[j]) ** 2).sum(-1))
    rows = np.arange(len(pts))[:, None]
    assert np.array_equal(d(rows, ref), d(rows, nb))


def test_small_clouds_and_errors():
    pts = np.array([[1.0, 2.0, 3.0], [1.0, 2.0, 3.5], [4.0, 2.0, 3.0]])
    o, nb = D.knn_order_neighbors(pts)
    assert nb.shape == (3, 3) and np.array_equal(nb, _brute(pts, 3))
    with pytest.raises(D.DvcError):
        D.knn_order_neighbors(np.array([[0.0, np.nan, 1.0]]))


def test_large_cloud_speed():
    rng = np.random.default_rng(11)
    pts = rng.uniform(0, 512, size=(200_000, 3))
    t = time.perf_counter()
    order, nb = D.knn_order_neighbors(pts, pts[0], nthreads=8)
    sec = time.perf_counter() - t
    print(f"200k points, 75 neighbours: {sec:.2f} s")
    sample = rng.choice(len(pts), 20, replace=False)
    for i in sample:
        dist = np.sqrt(((pts - pts[i]) ** 2).sum(1))
        assert np.array_equal(nb[i], np.lexsort((np.arange(len(pts)), dist))[:75])
    assert sec < 60
