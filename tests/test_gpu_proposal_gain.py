"""What the trained proposal network gains, on a held-out pair.

The pair (seed 900; training used seeds 100-147, the held-out loss 1100-1101) is made like the
training pairs: 64^3 tiles, each under its own random affine map (translation up to
8 voxels per axis, rotation up to 3 degrees, strain up to 2 %), 125 points around each
tile centre.  The plain search started at each point's imposed translation gives the
converged warp the method can reach.  Every point is then searched at once from
(a) the network's proposal and (b) the method without it: one rigid_trans start for
all points (the plain loop's start before any neighbour has converged); and (c) the
reference's serial chain, whose neighbour starts cross tile boundaries.  Bars: the
network's starts are much closer than the rigid start, at least 90 % of the points
reach the method's warp from them (more than from either alternative), in under half
the Gauss-Newton evaluations of the rigid start.
"""
import numpy as np
import pytest

import digitalvolumecorrelation_amd as D
from digitalvolumecorrelation_amd import proposal as P

pytestmark = pytest.mark.gpu


def test_network_starts_beat_the_method_without_it(ctx):
    if not P.WEIGHTS.exists():
        pytest.skip("no trained weights in the package")
    tp = P.tiled_pair(900, size=192, tile=64)
    target, _ = P.label_pair(ctx, tp, 64)              # leaves the pair in the context
    params = dict(P.LABEL_PARAMS, rigid_trans=tuple(float(v) for v in np.rint(tp.disp.mean(axis=0))))
    rp = D.make_params(**params)
    order, neigh = D.knn_order_neighbors(tp.points, start=tp.points[len(tp.points) // 2])

    def given(starts):
        ctx.set_start_mode(D.START_GIVEN)
        try:
            ctx.prepare(rp, tp.points, order, neigh)
            ctx.set_starts(starts)
            ctx.execute()
            return ctx.fetch()
        finally:
            ctx.set_start_mode(D.START_CHAIN)

    exact = np.zeros((len(tp.points), 12))
    exact[:, :3] = tp.disp
    ref = given(exact)
    ok = (ref["status"] == 0) & (ref["iters"] < 20)
    net = P.ProposalNet(ctx, max_batch=512, params=P.ProposalNet.load_params())
    prop = net.predict_points(tp.points)
    flat = np.zeros_like(prop)
    flat[:, :3] = params["rigid_trans"]
    rn, rf = given(prop), given(flat)
    rc = ctx.search(rp, tp.points, order, neigh)

    def reach(r):
        return (ok & (r["status"] == 0) & (np.abs(r["par"][:, :3] - ref["par"][:, :3]).max(axis=1) <= 1e-2)).sum() / ok.sum()

    en = np.linalg.norm(prop[ok, :3] - ref["par"][ok, :3], axis=1)
    ef = np.linalg.norm(flat[ok, :3] - ref["par"][ok, :3], axis=1)
    fn, ff, fc = reach(rn), reach(rf), reach(rc)
    print(f"\n  points {len(tp.points)}, converged from the imposed translation {ok.sum()}")
    print(f"  start error (voxels): network median {np.median(en):.3f} (90th pct {np.percentile(en, 90):.3f}); "
          f"one rigid start median {np.median(ef):.3f}")
    print(f"  reach the method's warp within 1e-2 voxel: network starts {fn:.3f}, rigid start {ff:.3f}, "
          f"serial chain {fc:.3f}")
    print(f"  mean evaluations: from imposed translation {ref['iters'][ok].mean():.2f}, network starts "
          f"{rn['iters'][ok].mean():.2f}, rigid start {rf['iters'][ok].mean():.2f}")
    assert np.median(en) < 0.4 * np.median(ef)
    assert fn >= 0.9 and fn > ff and fn > fc
    assert rn["iters"][ok].mean() < 0.5 * rf["iters"][ok].mean()
