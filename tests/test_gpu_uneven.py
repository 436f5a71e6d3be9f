"""Per-point subvolume geometry (dvc_set_point_geometry) on the GPU against the oracle's
loop with one RunControl per point (orc_dvc_run_geo).

Bar: statuses and iteration counts equal, displacements within 1e-3 voxel, objectives
within 1e-5 relative (points that end at the 20-evaluation limit: 1e-4, as in
test_gpu_start_modes).  Every point uniform: results identical to the one-geometry
search.
"""
import numpy as np
import pytest

import digitalvolumecorrelation_amd as D
from digitalvolumecorrelation_amd import synth

pytestmark = pytest.mark.gpu


def _close(g, o, label):
    assert np.array_equal(g["status"], o["status"]), label
    assert np.array_equal(g["iters"], o["iters"]), label
    good = o["status"] == 0
    conv = good & (o["iters"] < 20)
    limit = good & (o["iters"] >= 20)
    du = np.abs(g["par"][:, :3] - o["par"][:, :3]).max()
    rel = np.abs(g["obj"] - o["obj"]) / np.where(good, np.abs(o["obj"]), 1.0)
    rc = rel[conv].max() if conv.any() else 0.0
    rl = rel[limit].max() if limit.any() else 0.0
    print(f"  {label}: max |du| {du:.2e}  obj rel {rc:.2e} / {rl:.2e} at the limit  good {good.sum()}/{len(good)}")
    assert du <= 1e-3 and rc <= 1e-5 and rl <= 1e-4, label


@pytest.fixture(scope="module")
def case():
    return synth.translation_u8(size=80, n=4, spacing=8)


@pytest.mark.parametrize("obj", ["znssd", "ssd"])
def test_mixed_geometry_6dof(ctx, oracle, case, obj):
    rng = np.random.default_rng(17)
    n = len(case.points)
    geom = rng.integers(0, 2, n)
    size = rng.integers(11, 22, n)
    npp = rng.integers(300, 1500, n)
    params = dict(case.params, ndof=6, obj=obj)
    order, neigh = D.sort_order_neighbors(case.points)
    run = oracle.make_run(shape=case.ref.shape, bytes_per=1, **params)
    ores, _, _ = oracle.dvc_run_geo(run, case.ref, case.tar, case.points, order, neigh, geom, size, npp, nthreads=8)
    ctx.upload(0, case.ref)
    ctx.upload(1, case.tar)
    ctx.set_point_geometry(geom, size, npp)
    try:
        ctx.prepare(D.make_params(**params), case.points, order, neigh)
        ctx.execute()
        g = ctx.fetch()
    finally:
        ctx.set_point_geometry()
    _close(g, ores, f"mixed geometry, {obj}")


def test_uniform_geometry_matches_one_geometry_search(ctx, case):
    n = len(case.points)
    p = case.params
    order, neigh = D.sort_order_neighbors(case.points)
    ctx.upload(0, case.ref)
    ctx.upload(1, case.tar)
    rp = D.make_params(**p)
    # the per-point path runs one workgroup per point: compare it with the same launches
    # of the uniform path (slice launches sum the samples in another order)
    ctx.set_slice_launches(-1)
    try:
        plain = ctx.search(rp, case.points, order, neigh)
        ctx.set_point_geometry([p["geom"]] * n, [p["subvol_size"]] * n, [p["subvol_npts"]] * n)
        ctx.prepare(rp, case.points, order, neigh)
        ctx.execute()
        g = ctx.fetch()
    finally:
        ctx.set_point_geometry()
        ctx.set_slice_launches(0)
    for f in ("status", "iters", "obj", "par"):
        assert np.array_equal(g[f], plain[f]), f


def test_sphere_geometries_judged(ctx, oracle, case):
    # per-point geometry with the one-launch modes: given starts, every point a sphere of
    # its own size
    rng = np.random.default_rng(18)
    n = len(case.points)
    geom = np.ones(n, np.int32)
    size = rng.integers(13, 22, n)
    npp = rng.integers(400, 1000, n)
    params = dict(case.params, ndof=6)
    order, neigh = D.sort_order_neighbors(case.points)
    starts = np.zeros((n, 12))
    starts[:, :3] = case.truth + rng.uniform(-0.5, 0.5, (n, 3))
    ctx.upload(0, case.ref)
    ctx.upload(1, case.tar)
    ctx.set_point_geometry(geom, size, npp)
    ctx.set_start_mode(D.START_GIVEN)
    try:
        ctx.prepare(D.make_params(**params), case.points, order, neigh)
        ctx.set_starts(starts)
        ctx.execute()
        g = ctx.fetch()
    finally:
        ctx.set_point_geometry()
        ctx.set_start_mode(D.START_CHAIN)
    assert (g["status"] == 0).all()
    assert np.abs(g["par"][:, :3] - case.truth).max() < 0.1
