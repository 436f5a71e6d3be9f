"""Neighbour tables above the exact-sort limit.  The dvc program sorts every point's
distances as DataCloud::sort_order_neighbors does (Core/DataCloud.cpp:90-140, std::sort
of DualSort records, whose order among equal distances is the library's) up to 20,000
points and switches to the grid k-NN above, where equal distances keep point-index
order.  On a 30^3 lattice (27,000 points, about the largest the full sort handles in
seconds) both tables are built and both are searched on the GPU:

* the processing order is the same (one sort of distances from the start point);
* the neighbour lists differ in the order of equal distances, and as sets wherever the
  75th neighbour falls inside a shell of equal distances;
* the start guesses (means of the good earlier neighbours' translations) then differ,
  but every point converges to the same minimum: statuses equal and displacements
  within the 1e-3 voxel parity bar (measured: 8e-7 at most).
"""
import numpy as np
import pytest

import digitalvolumecorrelation_amd as D
from digitalvolumecorrelation_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(300)
def test_lattice_exact_vs_grid_tables(ctx):
    case = synth.translation_u8(size=288, n=30, spacing=8)    # every subvolume inside the volume
    start = np.full(3, 144.0)
    n = len(case.points)
    assert n == 27_000
    oe, ne = D.sort_order_neighbors(case.points, start=start, nthreads=16)
    og, ng = D.knn_order_neighbors(case.points, start=start, k=75, nthreads=16)
    assert np.array_equal(oe, og)
    lists = int((ne != ng).any(axis=1).sum())
    sets = int(sum(set(a) != set(b) for a, b in zip(ne, ng)))
    # the same distances in both tables: every list is a tie-order permutation of the
    # other's distance sequence
    d = lambda nb: np.linalg.norm(case.points[nb] - case.points[:, None, :], axis=2)
    assert np.array_equal(d(ne), d(ng))
    ctx.upload(0, case.ref)
    ctx.upload(1, case.tar)
    rp = D.make_params(**case.params)
    ge = ctx.search(rp, case.points, oe, ne)
    gg = ctx.search(rp, case.points, og, ng)
    du = np.abs(ge["par"][:, :3] - gg["par"][:, :3]).max(axis=1)
    err = np.abs(gg["par"][:, :3] - case.truth).max(axis=1)
    print(f"  27,000-point lattice: neighbour lists differing {lists}, as sets {sets}; results: statuses equal "
          f"{np.array_equal(ge['status'], gg['status'])}, |du| max {du.max():.2e}, median {np.median(du):.2e}, "
          f"points beyond 1e-3: {int((du > 1e-3).sum())}; error against the imposed shift {err.max():.3f}")
    assert np.array_equal(ge["status"], gg["status"]) and (ge["status"] == 0).all()
    assert du.max() <= 1e-3
    assert err.max() <= 0.05
