"""Argument checks of the search entry points through the C ABI: every invalid input is
an error return (DVC_ERR_ARG, -1) with a message and no launch, and the context stays
usable afterwards.  The reference CLI rejects the same inputs while reading them
(InputRead.cpp:477-630); here they are checked at dvc_search_prepare."""
import numpy as np
import pytest

import digitalvolumecorrelation_amd as D
from digitalvolumecorrelation_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = D.Context(0)
    yield c
    c.close()


def test_rejected_inputs_then_a_good_run(ctx):
    case = synth.translation_u8(size=72, n=3, spacing=8)
    order, neigh = D.sort_order_neighbors(case.points)
    ctx.upload(0, case.ref)
    ctx.upload(1, case.tar)
    rp = D.make_params(**case.params)
    n = len(case.points)
    dup = order.copy()
    dup[1] = dup[0]
    nb_bad = neigh.copy()
    nb_bad[2, 0] = n
    nb_neg = neigh.copy()
    nb_neg[0, 1] = -1
    nan_pts = case.points.copy()
    nan_pts[4, 2] = np.nan
    wide = np.tile(neigh[:, :1], (1, n + 1))
    bad = {
        "no points": (np.zeros((0, 3)), np.zeros(0, np.int32), np.zeros((0, 1), np.int32)),
        "order not a permutation": (case.points, dup, neigh),
        "neighbour index past the end": (case.points, order, nb_bad),
        "negative neighbour index": (case.points, order, nb_neg),
        "non-finite point": (nan_pts, order, neigh),
        "more neighbours than points": (case.points, order, wide),
    }
    for name, (p, o, nb) in bad.items():
        with pytest.raises(D.DvcError) as e:
            ctx.prepare(rp, p, o, nb)
        assert e.value.code == -1, name
        print(f"  {name}: {e.value}")
    with pytest.raises(D.DvcError):
        ctx.execute()                                    # nothing prepared
    ctx.prepare(rp, case.points, order, neigh)
    ctx.execute()
    r = ctx.fetch()
    assert (r["status"] == 0).all()
    assert np.abs(r["par"][:, :3] - case.truth).max() < 0.05


def test_single_point(ctx, oracle):
    """The smallest input the reference runs: one point, its own neighbour list of one
    (itself, never earlier in the order), so the start is the rigid translation."""
    case = synth.translation_u8(size=72, n=3, spacing=8)
    pt = case.points[13:14]
    order = np.zeros(1, np.int32)
    neigh = np.zeros((1, 1), np.int32)
    ctx.upload(0, case.ref)
    ctx.upload(1, case.tar)
    g = ctx.search(D.make_params(**case.params), pt, order, neigh)
    run = oracle.make_run(shape=case.ref.shape, bytes_per=1, jacobian=0, **case.params)
    o, _, _ = oracle.dvc_run(run, case.ref, case.tar, pt, order, neigh, nthreads=4)
    assert g["status"][0] == o["status"][0] == 0
    assert g["iters"][0] == o["iters"][0]
    assert np.abs(g["par"][0, :3] - o["par"][0, :3]).max() <= 1e-3
    assert np.abs(g["par"][0, :3] - case.truth[13]).max() < 0.05

