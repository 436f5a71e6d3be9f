"""Slice launches of the serial chain on the GPU (dvc_set_slice_launches): a chain
launch's points run one launch per Gauss-Newton evaluation with each point's samples
split over several workgroups, then the one-workgroup search resumes the points still
searching.  Checked against the oracle's serial dvc loop (Core/dvc.cpp:229-320 around
Search::process_point, Core/Search.cpp:95-246) and against the one-workgroup search of
the same library.

Bar (per assertion): statuses and iteration counts equal to the oracle's, |du| <= 1e-3
voxel, objectives <= 1e-5 relative; against the one-workgroup search (the same FP32
sample sums in another order) |dpar| <= 1e-6.
"""
import numpy as np
import pytest

import digitalvolumecorrelation_amd as D
from digitalvolumecorrelation_amd import synth

pytestmark = pytest.mark.gpu


def _run(ctx, case, params, order, neigh, slices):
    ctx.set_slice_launches(*slices)
    try:
        ctx.prepare(D.make_params(**params), case.points, order, neigh)
        ctx.execute()
        res = ctx.fetch()
        info = ctx.slice_info()
    finally:
        ctx.set_slice_launches(0, 0, 0)
    return res, info


def _check(g, o, w):
    assert np.array_equal(g["status"], o["status"])
    assert np.array_equal(g["iters"], o["iters"])
    good = o["status"] == 0
    assert np.abs(g["par"][:, :3] - o["par"][:, :3]).max() <= 1e-3
    if good.any():
        assert (np.abs(g["obj"][good] - o["obj"][good]) / np.abs(o["obj"][good])).max() <= 1e-5
        assert np.abs(g["par"][good] - w["par"][good]).max() <= 1e-6
    assert np.array_equal(g["status"], w["status"])
    return good


@pytest.fixture(scope="module")
def affine():
    # 12-DOF sphere clouds of the affine workload's shape on a smaller volume
    return synth.affine_u16(size=128, n=5, spacing=12, npts=3000, diameter=31)


@pytest.mark.parametrize("slices", [(0, 4, 512), (0, 1, 512), (0, 2, 256), (0, 4, 1024)])
def test_sliced_affine_u16_12dof(ctx, oracle, affine, slices):
    case = affine
    p = dict(case.params, disp_max=8)
    order, neigh = D.sort_order_neighbors(case.points, case.start)
    ctx.upload(0, case.ref)
    ctx.upload(1, case.tar)
    g, info = _run(ctx, case, p, order, neigh, slices)
    w, winfo = _run(ctx, case, p, order, neigh, (-1, 0, 0))
    run = oracle.make_run(shape=case.ref.shape, bytes_per=2, **p)
    o, _, _ = oracle.dvc_run(run, case.ref, case.tar, case.points, order, neigh, nthreads=8)
    print(f"slices {slices}: {info[0]} sliced launches, {info[1]} workgroups per point, {info[2]} slice launches each")
    assert info[0] > 0 and winfo[0] == 0
    good = _check(g, o, w)
    assert good.all()


@pytest.mark.parametrize("ndof,obj", [(6, "ssd"), (3, "znssd"), (6, "nssd")])
def test_sliced_dofs_objectives_range_fail(ctx, oracle, ndof, obj):
    # the outer ring of points has boxes leaving the volume (Range_Fail at set-up)
    case = synth.translation_u8(size=72, n=5, spacing=11, npts=2000)
    p = dict(case.params, ndof=ndof, obj=obj, subvol_size=21, disp_max=4)
    order, neigh = D.sort_order_neighbors(case.points, case.start)
    ctx.upload(0, case.ref)
    ctx.upload(1, case.tar)
    g, info = _run(ctx, case, p, order, neigh, (0, 3, 512))
    w, _ = _run(ctx, case, p, order, neigh, (-1, 0, 0))
    run = oracle.make_run(shape=case.ref.shape, bytes_per=1, **p)
    o, _, _ = oracle.dvc_run(run, case.ref, case.tar, case.points, order, neigh, nthreads=8)
    assert info[0] > 0
    _check(g, o, w)
    assert (o["status"] == -1).any() and (o["status"] == 0).any()


@pytest.mark.parametrize("ndof,slices", [(6, (0, 3, 256)), (12, (0, 4, 512))])
def test_sliced_per_point_geometry(ctx, oracle, ndof, slices):
    """Per-point geometry (dvc_set_point_geometry) in slice launches: each point is split
    into as many workgroups as its own sample count needs (1-6 here), a chain launch into
    as many as its largest point's.  Against the oracle's loop with one RunControl per point
    (orc_dvc_run_geo) and against the same search without slice launches."""
    case = synth.translation_u8(size=80, n=4, spacing=8)
    rng = np.random.default_rng(23)
    n = len(case.points)
    geom = rng.integers(0, 2, n)
    size = rng.integers(13, 22, n)
    npp = rng.integers(200, 1500, n)
    p = dict(case.params, ndof=ndof, obj="znssd")
    order, neigh = D.sort_order_neighbors(case.points)
    ctx.upload(0, case.ref)
    ctx.upload(1, case.tar)
    ctx.set_point_geometry(geom, size, npp)
    try:
        g, info = _run(ctx, case, p, order, neigh, slices)
        w, winfo = _run(ctx, case, p, order, neigh, (-1, 0, 0))
    finally:
        ctx.set_point_geometry()
    run = oracle.make_run(shape=case.ref.shape, bytes_per=1, **p)
    o, _, _ = oracle.dvc_run_geo(run, case.ref, case.tar, case.points, order, neigh, geom, size, npp, nthreads=8)
    print(f"per-point geometry, {ndof} DOF, slices {slices}: {info[0]} sliced launches, up to {info[1]} "
          f"workgroups per point")
    assert info[0] > 0 and winfo[0] == 0
    assert np.array_equal(g["status"], o["status"])
    assert np.array_equal(g["iters"], o["iters"])
    good = o["status"] == 0
    conv = good & (o["iters"] < 20)
    assert np.abs(g["par"][good, :3] - o["par"][good, :3]).max() <= 1e-3
    assert (np.abs(g["obj"][conv] - o["obj"][conv]) / np.abs(o["obj"][conv])).max() <= 1e-5
    assert np.array_equal(g["status"], w["status"])
    assert np.abs(g["par"][good] - w["par"][good]).max() <= 1e-6


def test_wide_launches_take_fewer_slices(ctx):
    """The 15^3 affine workload (256^3 u16, sphere 41 / 5000 samples, 12-DOF): its widest
    chain launches hold 25-30 points, 4 per XCD, so at 10 slices per point an XCD would
    need 40 workgroups for its 32 CUs; those launches run 8 slices per point instead
    (the kernels deal each point's 64-sample chunks over however many slices it has).
    Sliced results against the one-workgroup search on all 3375 points: statuses equal,
    |du| <= 1e-3 voxel everywhere, and |dpar| <= 1e-6 wherever the iteration counts agree
    (they may differ where a stopping test sits within the FP32 summation-order
    difference; at most 2 points)."""
    case = synth.affine_u16()
    order, neigh = D.sort_order_neighbors(case.points, case.start)
    ctx.upload(0, case.ref)
    ctx.upload(1, case.tar)
    g, info = _run(ctx, case, case.params, order, neigh, (0, 4, 512))
    w, winfo = _run(ctx, case, case.params, order, neigh, (-1, 0, 0))
    print(f"wide launches: {info[0]} sliced launches, up to {info[1]} workgroups per point")
    assert info[0] == 256 and info[1] == 10 and winfo[0] == 0
    assert np.array_equal(g["status"], w["status"]) and (g["status"] == 0).all()
    assert np.abs(g["par"][:, :3] - w["par"][:, :3]).max() <= 1e-3
    same = g["iters"] == w["iters"]
    assert (~same).sum() <= 2
    assert np.abs(g["par"][same] - w["par"][same]).max() <= 1e-6
