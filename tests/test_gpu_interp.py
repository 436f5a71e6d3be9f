"""GPU interpolation and objectives against the oracle (Core/Interpolate.cpp,
Core/ObjectiveFunctions.cpp).  The GPU evaluates the interpolant in FP32: values
agree with the FP64 oracle within 2e-6 of the volume's full range (tolerance stated
per test); objectives are FP64 on both sides."""
import numpy as np
import pytest

import digitalvolumecorrelation_amd as D
from digitalvolumecorrelation_amd import synth

pytestmark = pytest.mark.gpu


def _pair(oracle, vol, **kw):
    return (oracle.make_run(shape=vol.shape, bytes_per=vol.dtype.itemsize, **kw), D.make_params(**kw))


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_tricubic_matches_oracle(ctx, oracle, dtype):
    vol = synth.smoothed_noise((64, 64, 64), 1.5, dtype, seed=11)
    ctx.upload(0, vol)
    run, par = _pair(oracle, vol, subvol_size=31, disp_max=5)
    rng = np.random.default_rng(1)
    for center in ([32.3, 31.7, 32.9], [30.0, 33.0, 31.5]):
        c = np.array(center)
        # whole active box, including the face cells (act box = est box - 1 voxel)
        xyz = c + rng.uniform(-20.4, 20.4, size=(6000, 3))
        bb = np.r_[xyz.min(0), xyz.max(0)]
        ov, og = oracle.interp(run, vol, c, xyz, bb, 2, grad=True)
        gv, gg = ctx.interpolate(0, par, c, xyz, bb, 2, grad=True)
        assert ov is not None and gv is not None
        top = float(np.iinfo(dtype).max)
        assert np.abs(gv - ov).max() <= 2e-6 * top
        assert np.abs(gg - og).max() <= 2e-6 * top


def test_face_rule_closed_form(ctx, oracle):
    n = 48
    z, y, x = np.meshgrid(*[np.arange(n)] * 3, indexing="ij")
    vol = (x.astype(np.int64) ** 2).astype(np.uint16)
    ctx.upload(0, vol)
    par = D.make_params(subvol_size=10, disp_max=2)
    t = np.linspace(0.05, 0.95, 19)
    xs = 27.0 + t
    pts = np.stack([xs, np.full_like(xs, 20.3), np.full_like(xs, 19.7)], 1)
    got = ctx.interpolate(0, par, [20.0, 20.0, 20.0], pts, np.r_[pts.min(0), pts.max(0)])
    assert np.abs(got - (xs ** 2 + 2 * t ** 2 * (1 - t))).max() <= 2e-6 * 65535


def test_trilinear_nearest_and_range(ctx, oracle):
    vol = synth.smoothed_noise((40, 40, 40), 1.0, np.uint8, seed=12)
    ctx.upload(1, vol)
    run, par = _pair(oracle, vol, subvol_size=10, disp_max=2)
    c = np.array([20.2, 19.9, 20.4])
    xyz = c + np.random.default_rng(2).uniform(-6.5, 6.5, size=(800, 3))
    bb = np.r_[xyz.min(0), xyz.max(0)]
    for m in (0, 1):
        ov = oracle.interp(run, vol, c, xyz, bb, m)
        gv = ctx.interpolate(1, par, c, xyz, bb, m)
        assert np.abs(gv - ov).max() <= (0.0 if m == 0 else 2e-6 * 255)
    c0 = np.array([20.0, 20.0, 20.0])
    inside = np.array([[12.0, 15.0, 15.0], [27.5, 20.0, 26.0]])
    for pts, ok in ((inside, True), (inside - [0.001, 0, 0], False), (inside + [0, 0, 2.0], False),
                    (np.array([[28.0, 15.0, 15.0]]), False)):
        bb = np.r_[pts.min(0), pts.max(0)]
        assert (ctx.interpolate(1, par, c0, pts, bb) is not None) == ok
        assert (oracle.interp(run, vol, c0, pts, bb) is not None) == ok


@pytest.mark.parametrize("kind", ["sad", "ssd", "zssd", "nssd", "znssd"])
def test_objectives(ctx, oracle, kind):
    rng = np.random.default_rng(5)
    r = rng.random(5003) * 250
    t = 1.1 * r + rng.normal(0, 3, len(r))
    ov, ores = oracle.objective(kind, r, t, residual=True)
    gv, gres = ctx.objective(kind, r, t, residual=True)
    assert gv == pytest.approx(ov, rel=1e-12)
    assert np.abs(gres - ores).max() <= 1e-12 * np.abs(ores).max()
