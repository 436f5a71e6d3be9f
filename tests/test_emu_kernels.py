"""The search and strain kernels of dvc_kernels.hip, built for the CPU (tests/emu), against
the oracle: the kernel logic (dependency levels, start guesses, range checks, analytic
Gauss-Newton, MFMA moment layout, LDS solver) is checked without a GPU.  Same bar as the
GPU tests: statuses and iterations equal, |du| <= 1e-3 voxel, objective <= 1e-5 relative."""
import sys
from pathlib import Path

import numpy as np
import pytest

import digitalvolumecorrelation_amd as D
from digitalvolumecorrelation_amd import synth

sys.path.insert(0, str(Path(__file__).resolve().parent / "emu"))
import emu  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def built():
    emu.build()


def _cmp(oracle, case, params, threads=64, nproc=0, gather32=True, precompute=False):
    order, neigh = D.sort_order_neighbors(case.points, start=case.start)
    run = oracle.make_run(shape=case.ref.shape, bytes_per=case.ref.dtype.itemsize, **params)
    o, _, _ = oracle.dvc_run(run, case.ref, case.tar, case.points, order, neigh, nproc=nproc, nthreads=4)
    g = emu.search(D.make_params(**params), case.ref, case.tar, case.points, order, neigh, nproc=nproc, threads=threads,
                   gather32=gather32, precompute=precompute)
    assert np.array_equal(g["status"], o["status"])
    assert np.array_equal(g["iters"], o["iters"])
    good = o["status"] == 0
    assert np.abs(g["par"][:, :3] - o["par"][:, :3]).max() <= 1e-3
    if good.any():
        assert (np.abs(g["obj"][good] - o["obj"][good]) / o["obj"][good]).max() <= 1e-5
    return g, o


@pytest.mark.parametrize("gather32", [True, False])
def test_translation_cube_u8(oracle, gather32):
    case = synth.translation_u8(size=56, n=2, spacing=8, npts=1000)
    _cmp(oracle, case, dict(case.params, subvol_size=21, disp_max=4), gather32=gather32)


def test_odd_row_width_u8(oracle):
    """A row width that is not a whole number of dwords takes the 64-bit row addressing."""
    case = synth.translation_u8(size=58, n=2, spacing=8, npts=1000)
    ref, tar = case.ref[:, :, :57].copy(), case.tar[:, :, :57].copy()
    case = synth.Case(case.name, ref, tar, case.points, case.truth, case.params)
    _cmp(oracle, case, dict(case.params, subvol_size=21, disp_max=4))


@pytest.mark.parametrize("precompute", [False, True])
def test_affine_sphere_u16_12dof(oracle, precompute):
    case = synth.affine_u16(size=72, n=2, spacing=10, npts=600, diameter=21)
    g, o = _cmp(oracle, case, dict(case.params, disp_max=5), threads=128, precompute=precompute)
    assert np.all(o["status"] == 0)


@pytest.mark.parametrize("obj", ["ssd", "zssd", "nssd", "sad"])
def test_6dof_objectives_and_partial(oracle, obj):
    case = synth.affine_u16(size=72, n=2, spacing=10, npts=600, diameter=21)
    _cmp(oracle, case, dict(case.params, disp_max=5, ndof=6, obj=obj), nproc=5)


@pytest.mark.parametrize("n", [3, 6, 12])
def test_gn_solve_matches_oracle_bitwise(oracle, n):
    """The wave-parallel column-pivoted QR of the search kernel against the oracle's
    serial one (Eigen 3.4 ColPivHouseholderQR, Search.cpp:234): identical bits on
    well-conditioned, symmetric positive definite, rank-deficient (duplicate and zero
    columns), pivot-tie and all-zero systems."""
    rng = np.random.default_rng(40 + n)
    cases = []
    A = rng.normal(size=(n, n))
    cases.append(A)
    J = rng.normal(size=(50, n)) * np.logspace(0, 4, n)
    cases.append(J.T @ J)
    B = rng.normal(size=(n, n))
    B[:, n // 2] = B[:, 0]
    cases.append(B)
    Z = rng.normal(size=(n, n))
    Z[:, -1] = 0.0
    Z[-1, :] = 0.0
    cases.append(Z)
    cases.append(np.eye(n) * 2.0)                   # every column norm ties
    cases.append(np.zeros((n, n)))
    for A in cases:
        b = rng.normal(size=n)
        want = oracle.colpiv_solve(A, b)
        got = emu.gn_solve(A, b)
        assert np.array_equal(np.isnan(want), np.isnan(got))
        ok = ~np.isnan(want)
        assert np.array_equal(got[ok], want[ok]), (A, got, want)


@pytest.mark.parametrize("uniform", [False, True])
@pytest.mark.parametrize("n", [3, 6, 12])
def test_gn_solve_cholesky(oracle, n, uniform):
    """The search kernel's own solve: Cholesky of the symmetric normal matrix, with the
    column-pivoted QR taking over when a pivot is not above 1e-9 of the largest diagonal.
    Well-conditioned and badly scaled systems agree with an LU solve to the conditioning
    bound; singular, near-singular and all-zero systems give the QR's bits.  uniform: the
    slice launches' Gauss-Jordan with one row per lane (same pivot rule)."""
    rng = np.random.default_rng(70 + n)
    for scale in (np.ones(n), np.logspace(0, 3, n)):
        J = rng.normal(size=(40, n)) * scale
        A = J.T @ J
        b = rng.normal(size=n)
        got = emu.gn_solve(A, b, chol=True, uniform=uniform)
        want = np.linalg.solve(A, b)
        assert np.linalg.norm(got - want) <= 1e-13 * np.linalg.cond(A) * np.linalg.norm(want)
    singular = []
    J = rng.normal(size=(40, n))
    J[:, n // 2] = J[:, 0]
    singular.append(J.T @ J)                        # duplicate column
    J = rng.normal(size=(40, n))
    J[:, -1] *= 1e-6
    singular.append(J.T @ J)                        # pivot ratio ~1e-12
    Z = J.T @ J
    Z[:, -1] = 0.0
    Z[-1, :] = 0.0
    singular.append(Z)
    singular.append(np.zeros((n, n)))
    for A in singular:
        b = rng.normal(size=n)
        want = oracle.colpiv_solve(A, b)
        got = emu.gn_solve(A, b, chol=True, uniform=uniform)
        assert np.array_equal(np.isnan(want), np.isnan(got))
        ok = ~np.isnan(want)
        assert np.array_equal(got[ok], want[ok]), (A, got, want)


@pytest.mark.parametrize("cap", [1, 7, 40])
def test_capacity_schedule(cap):
    """Launches of at most `cap` points: every point once, and every neighbour earlier in
    the processing order in an earlier launch (so each start guess reads finished
    results); never more launches than the rounds of the dependency levels, and with
    cap 1 exactly the processing order."""
    rng = np.random.default_rng(cap)
    pts = rng.uniform(0, 100, (3000, 3))
    order, neigh = D.knn_order_neighbors(pts, start=np.full(3, 50.0))
    pos = np.empty(len(order), np.int64)
    pos[order] = np.arange(len(order))
    launches = emu.schedule(order, neigh, cap)
    levels = emu.schedule(order, neigh, 0)
    at = np.full(len(order), -1)
    for t, L in enumerate(launches):
        assert 0 < len(L) <= cap
        assert (at[L] == -1).all()
        at[L] = t
    assert (at >= 0).all()
    earlier = pos[neigh] < pos[:, None]
    assert (at[neigh][earlier] < np.broadcast_to(at[:, None], neigh.shape)[earlier]).all()
    assert len(launches) <= sum(-(-len(L) // cap) for L in levels)
    assert len(launches) >= max(len(levels), -(-len(order) // cap))
    if cap == 1:
        assert np.array_equal(np.concatenate(launches), order)


@pytest.mark.parametrize("n", [1, 7, 1000])
def test_reference_block_order(n):
    """The reference kernel's workgroup order is a permutation of the processing
    positions; the blocks one XCD runs (b, b + 8, ...) hold one contiguous run of the
    points sorted by Morton code of their 16-voxel cell, so neighbouring points share
    that XCD's L2."""
    rng = np.random.default_rng(n)
    pts = rng.uniform(0, 500, (n, 3))
    order = rng.permutation(n).astype(np.int32)
    out = emu.spatial_order(pts, order)
    assert np.array_equal(np.sort(out), np.arange(n))

    def morton(q):
        c = np.floor(q / 16.0).astype(np.int64)
        k = 0
        for bit in range(21):
            for d in range(3):
                k |= ((c[d] >> bit) & 1) << (3 * bit + d)
        return k

    keys = np.array([morton(pts[order[i]]) for i in range(n)])
    runs = [out[x::8] for x in range(8)]
    flat = np.concatenate(runs)
    assert (np.diff(keys[flat]) >= 0).all()       # XCD 0's run, then XCD 1's, ... in key order


def test_strain_kernel(oracle):
    pts = synth.grid_points(5, 9.0, (40.0, 50.0, 60.0))
    rng = np.random.default_rng(3)
    disp = pts @ np.array([[0.01, 0.002, 0.0], [0.0, -0.004, 0.003], [0.001, 0.0, 0.006]]).T + rng.normal(0, 0.05, pts.shape)
    st = np.where(rng.random(len(pts)) < 0.1, -1, 0).astype(np.int32)
    om = rng.random(len(pts)) * 0.02
    order, neigh = D.sort_order_neighbors(pts)
    for ndp, refill in ((10, False), (27, True), (75, False)):
        o = oracle.strain(pts, disp, st, om, neigh, ndp=ndp, objt=0.015, refill=refill)
        g = emu.strain(pts, disp, st, om, neigh, ndp=ndp, objt=0.015, refill=refill)
        assert np.array_equal(o["npts"], g["npts"])
        for k in ("DG", "E", "L", "disp", "swr"):
            assert np.abs(o[k] - g[k]).max() <= 1e-9, k


def test_12dof_range_failures(oracle):
    """Range_Fail at set-up (boxes leaving the volume) and inside the Gauss-Newton loop
    (the moving cloud leaving the active box with disp_max 1), where the bounding box
    comes from the speculative sample pass."""
    case = synth.affine_u16(size=72, n=3, spacing=20, npts=600, diameter=21)
    g, o = _cmp(oracle, case, dict(case.params, disp_max=5), threads=128, precompute=True)
    assert (o["status"] == D.RANGE_FAIL).any() and (o["status"] == D.POINT_GOOD).any()
    g, o = _cmp(oracle, case, dict(case.params, disp_max=1, rigid_trans=(0.0, 0.0, 0.0)), threads=128)
    assert np.all(o["status"] == D.RANGE_FAIL) and np.all(o["iters"] >= 3)


def _sliced(oracle, case, params, slices, threads=64, nproc=0):
    """Slice launches (one launch per evaluation, samples split over workgroups, the last
    workgroup of a point taking the step) against the oracle and against the
    one-workgroup search of the same build."""
    order, neigh = D.sort_order_neighbors(case.points, start=case.start)
    run = oracle.make_run(shape=case.ref.shape, bytes_per=case.ref.dtype.itemsize, **params)
    o, _, _ = oracle.dvc_run(run, case.ref, case.tar, case.points, order, neigh, nproc=nproc, nthreads=4)
    pr = D.make_params(**params)
    g, ns = emu.search(pr, case.ref, case.tar, case.points, order, neigh, nproc=nproc, threads=threads,
                       precompute=True, slices=slices)
    w = emu.search(pr, case.ref, case.tar, case.points, order, neigh, nproc=nproc, threads=threads, precompute=True)
    assert ns > 0
    for other in (o, w):
        assert np.array_equal(g["status"], other["status"])
        assert np.array_equal(g["iters"], other["iters"])
    good = o["status"] == 0
    assert np.abs(g["par"][:, :3] - o["par"][:, :3]).max() <= 1e-3
    if good.any():
        assert (np.abs(g["obj"][good] - o["obj"][good]) / o["obj"][good]).max() <= 1e-5
        # only the summation order of the per-point sums differs from the one-workgroup search
        assert np.abs(g["par"][good] - w["par"][good]).max() <= 1e-5
    return g, o


@pytest.mark.parametrize("evals", [1, 2, 4])
def test_sliced_affine_sphere_u16_12dof(oracle, evals):
    """600 samples in three slices of 256; with fewer slice launches than evaluations the
    one-workgroup search resumes the points from their carried state."""
    case = synth.affine_u16(size=72, n=2, spacing=10, npts=600, diameter=21)
    g, o = _sliced(oracle, case, dict(case.params, disp_max=5), (0, evals, 256), threads=128)
    assert np.all(o["status"] == 0)


@pytest.mark.parametrize("ndof,obj", [(6, "ssd"), (6, "zssd"), (3, "znssd"), (6, "nssd"), (6, "sad")])
def test_sliced_dofs_objectives(oracle, ndof, obj):
    case = synth.affine_u16(size=72, n=2, spacing=10, npts=700, diameter=21)
    _sliced(oracle, case, dict(case.params, disp_max=5, ndof=ndof, obj=obj), (0, 3, 256), nproc=5)


def test_sliced_translation_cube_u8_range_fail(oracle):
    """Cube clouds (grid samples); the outer points' boxes leave the volume (Range_Fail at
    set-up), in slice launches with two slices of 512 samples."""
    case = synth.translation_u8(size=56, n=3, spacing=12, npts=1000)
    g, o = _sliced(oracle, case, dict(case.params, subvol_size=21, disp_max=4), (0, 4, 512))
    assert (o["status"] == -1).any() and (o["status"] == 0).any()
