"""Device-made synthetic volumes (dvc_synth_noise, dvc_synth_warp_field).

Checks: the smoothed noise spans the full uint16 range after its min-max rescale and
is smooth (neighbour correlation near 1, as for the host generator); an identity
field (no bumps, no strain) reproduces the reference exactly (Catmull-Rom at integer
nodes); a bump far wider than the volume is a uniform translation and matches the
host's separable Catmull-Rom translation (synth.translate) to within one rounding
level; and the full-scale case's imposed field is what bump_field says.
"""
import numpy as np
import pytest

from digitalvolumecorrelation_amd import synth

pytestmark = pytest.mark.gpu


def test_noise_range_and_smoothness(ctx):
    shape = (40, 48, 56)
    ctx.synth_noise(0, shape, np.uint16, 2.0, 7)
    v = ctx.download(0, shape, np.uint16).astype(np.float64)
    assert v.min() == 0 and v.max() == 65535
    h = synth.smoothed_noise(shape, 2.0, np.uint16, seed=7).astype(np.float64)

    def corr(a):
        x = a[:, :, :-1].ravel() - a.mean()
        y = a[:, :, 1:].ravel() - a.mean()
        return (x * y).mean() / (a.var())

    assert corr(v) > 0.9 and abs(corr(v) - corr(h)) < 0.05
    ctx.synth_noise(1, shape, np.uint16, 2.0, 8)
    w = ctx.download(1, shape, np.uint16)
    assert not np.array_equal(w, v.astype(np.uint16))        # another seed, another volume


def test_identity_and_uniform_translation(ctx):
    shape = (32, 36, 40)
    ctx.synth_noise(0, shape, np.uint16, 1.5, 9)
    ref = ctx.download(0, shape, np.uint16)
    c = np.array([20.0, 18.0, 16.0])
    ctx.synth_warp_field(np.zeros((0, 7)), 0.0, c)
    assert np.array_equal(ctx.download(1, shape, np.uint16), ref)
    # a sinusoid of zero amplitude is the identity too
    ctx.synth_warp_field(np.zeros((0, 7)), 0.0, c, waves=np.array([[0.1, 0.2, 0.3, 1.0, 0.0, 0.0, 0.0]]))
    assert np.array_equal(ctx.download(1, shape, np.uint16), ref)
    t = np.array([1.25, -0.5, 2.75])
    ctx.synth_warp_field(np.array([[0, 0, 0, 1e9, *t]]), 0.0, c)
    got = ctx.download(1, shape, np.uint16).astype(np.int64)
    want = synth.translate(ref, t).astype(np.int64)
    assert np.abs(got - want).max() <= 1
    assert (got == want).mean() > 0.999


def test_full_scale_field_matches_host_formula(ctx):
    case = synth.full_scale_u16(ctx, size=96, n=3, spacing=10.0)
    u = synth.bump_field(case.points, case.bumps, case.strain, np.full(3, 47.5))
    assert np.allclose(u, case.truth)
    assert abs((1 + case.strain) ** 3 - 1.01) < 1e-12
    assert case.params["subvol_npts"] == 8000
