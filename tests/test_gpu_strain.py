"""GPU strain-window fit (dvc_strain_fit) against the oracle (Core/strain.cpp:402-591)
and against exact fields.  Both sides are FP64: strains agree within 1e-9 (the
required bar is 1e-5)."""
import numpy as np
import pytest

import digitalvolumecorrelation_amd as D
from digitalvolumecorrelation_amd import synth

pytestmark = pytest.mark.gpu


def _cmp(ctx, oracle, pts, disp, st, om, ndp, **kw):
    order, neigh = D.sort_order_neighbors(pts)
    o = oracle.strain(pts, disp, st, om, neigh, ndp=ndp, **kw)
    g = ctx.strain(pts, disp, st, om, neigh, ndp=ndp, **kw)
    assert np.array_equal(o["npts"], g["npts"])
    for key, tol in (("DG", 1e-9), ("E", 1e-9), ("L", 1e-9), ("disp", 1e-7), ("swr", 1e-9)):
        assert np.abs(o[key] - g[key]).max() <= tol, key
    return g


def test_linear_and_quadratic_fields(ctx, oracle):
    pts = synth.grid_points(8, 9.0, (70.0, 60.0, 80.0))
    G = np.array([[0.010, 0.002, -0.003], [0.001, -0.004, 0.002], [0.0, 0.003, 0.005]])
    disp = 0.7 + pts @ G.T
    n = len(pts)
    g = _cmp(ctx, oracle, pts, disp, np.zeros(n, np.int32), np.full(n, 0.01), 27)
    assert np.abs(g["DG"] - G.reshape(1, 9)).max() < 1e-9
    q = 1e-4 * pts[:, [0]] * pts[:, [1]] * np.array([[1.0, -0.5, 0.25]])
    _cmp(ctx, oracle, pts, disp + q, np.zeros(n, np.int32), np.full(n, 0.01), 30)


def test_filtering_refill_and_planar(ctx, oracle):
    rng = np.random.default_rng(4)
    pts = synth.grid_points(7, 10.0, (60.0, 70.0, 80.0))
    n = len(pts)
    disp = rng.normal(0, 0.3, (n, 3)) + pts @ np.diag([0.01, -0.02, 0.005]).T
    st = np.where(rng.random(n) < 0.1, -1, 0).astype(np.int32)
    om = rng.random(n) * 0.05
    for ndp in (10, 25, 75):
        for refill in (False, True):
            _cmp(ctx, oracle, pts, disp, st, om, ndp, objt=0.03, refill=refill)
    plane = pts[np.isclose(pts[:, 1], pts[0, 1])]
    dp = plane @ np.array([[0.01, 0, 0.003], [0, 0, 0], [0.002, 0, -0.004]]).T
    _cmp(ctx, oracle, plane, dp, np.zeros(len(plane), np.int32), np.zeros(len(plane)), 20)


def test_strain_after_search_on_device(ctx):
    """dvc_search_set_strain: the fit runs on the device after every execute, from the
    search's own results and neighbour table; it equals dvc_strain_fit on the fetched
    results (the same kernel on the same inputs), not-searched points included."""
    case = synth.translation_u8(size=80, n=4, spacing=8)
    order, neigh = D.sort_order_neighbors(case.points, case.start)
    ctx.upload(0, case.ref)
    ctx.upload(1, case.tar)
    ctx.set_search_strain(ndp=27)
    try:
        ctx.prepare(D.make_params(**case.params), case.points, order, neigh, nproc=40)
        ctx.execute()
        res = ctx.fetch()
        dev = ctx.fetch_search_strain()
    finally:
        ctx.set_search_strain(ndp=0)
    assert (res["status"] == D.NOT_SEARCH).sum() == len(case.points) - 40
    host = ctx.strain(case.points, res["par"][:, :3].copy(), res["status"], res["obj"], neigh, ndp=27)
    for f in ("disp", "npts", "swr", "E", "L", "DG"):
        assert np.array_equal(dev[f], host[f]), f
