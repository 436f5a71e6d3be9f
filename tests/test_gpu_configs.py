"""The two reference-scale workloads on the GPU against the oracle.

* 128^3 uint8 translation case (10^3 grid at spacing 8, cube 31 / 2000 -> 13^3 samples,
  ZNSSD, tricubic, 3 DOF, disp_max 5) in full: every point against the oracle's serial
  loop, and the 27-point strain window of the result.
* 256^3 uint16 12-DOF affine case (15^3 grid at spacing 12, sphere 41 / 5000): the
  first 160 points of the processing order against the oracle, and, over all 3375
  points, properties that do not need the oracle: every point converges, displacements
  match the imposed affine map, and the strain fit recovers its gradient.
Bar: statuses and iterations equal, |du| <= 1e-3 voxel, objectives <= 1e-5 relative,
strain <= 1e-9 (FP64 on both sides)."""
import numpy as np
import pytest

import digitalvolumecorrelation_amd as D
from digitalvolumecorrelation_amd import synth

pytestmark = pytest.mark.gpu


def _check(g, o):
    assert np.array_equal(g["status"], o["status"])
    assert np.array_equal(g["iters"], o["iters"])
    good = o["status"] == 0
    assert np.abs(g["par"][:, :3] - o["par"][:, :3]).max() <= 1e-3
    assert (np.abs(g["obj"][good] - o["obj"][good]) / np.abs(o["obj"][good])).max() <= 1e-5
    return good


def test_translation_u8_128_full(ctx, oracle):
    case = synth.translation_u8()
    p = case.params
    order, neigh = D.sort_order_neighbors(case.points, case.start)
    ctx.upload(0, case.ref)
    ctx.upload(1, case.tar)
    g = ctx.search(D.make_params(**p), case.points, order, neigh)
    run = oracle.make_run(shape=case.ref.shape, bytes_per=1, **p)
    o, sec, n = oracle.dvc_run(run, case.ref, case.tar, case.points, order, neigh, nthreads=8)
    print(f"oracle {n} points in {sec:.1f} s; gpu kernel {ctx.kernel_ms():.2f} ms")
    good = _check(g, o)
    assert good.all()
    assert np.abs(g["par"][:, :3] - case.truth).max() <= 0.02
    gs = ctx.strain(case.points, g["par"][:, :3], g["status"], g["obj"], neigh, ndp=27)
    os_ = oracle.strain(case.points, g["par"][:, :3], g["status"], g["obj"], neigh, ndp=27)
    for k in ("E", "L", "DG", "disp", "swr"):
        assert np.abs(gs[k] - os_[k]).max() <= 1e-9 * max(1.0, np.abs(os_[k]).max()), k
    assert np.array_equal(gs["npts"], os_["npts"])
    assert np.abs(gs["DG"]).max() <= 2e-3          # a rigid translation has no strain


def test_affine_u16_256_prefix_and_properties(ctx, oracle):
    case = synth.affine_u16()
    p = case.params
    order, neigh = D.sort_order_neighbors(case.points, case.start)
    ctx.upload(0, case.ref)
    ctx.upload(1, case.tar)
    g = ctx.search(D.make_params(**p), case.points, order, neigh)
    run = oracle.make_run(shape=case.ref.shape, bytes_per=2, **p)
    M = 160
    o, sec, n = oracle.dvc_run(run, case.ref, case.tar, case.points, order, neigh, nproc=M, nthreads=8)
    idx = order[:M]
    print(f"oracle {n} points in {sec:.1f} s")
    _check(g[idx], o[idx])
    assert np.all(g["status"] == 0)
    assert np.abs(g["par"][:, :3] - case.truth).max() <= 0.02
    gs = ctx.strain(case.points, g["par"][:, :3], g["status"], g["obj"], neigh, ndp=27)
    # imposed map: 1 degree about x, then 0.5 % along z; displacement gradient F - I
    c, s = np.cos(np.radians(1.0)), np.sin(np.radians(1.0))
    F = np.array([[1, 0, 0], [0, c, -s], [0, s, c]]) @ np.diag([1.0, 1.0, 1.005])
    inner = np.all(np.abs(case.points - 128.0) <= 60.0, axis=1)
    DG = gs["DG"].reshape(-1, 3, 3)[inner]
    assert np.abs(DG - (F - np.eye(3))).max() <= 2e-3
