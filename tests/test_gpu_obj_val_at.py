"""dvc_obj_val_at -- Search::obj_val_at(x[, residual]) in the state search_pt_setup leaves
(Core/Search.cpp:248-275, 448-488) -- on the GPU against the oracle's orc_obj_val_at,
over both cloud shapes, the five objectives, the three interpolators and 3 / 6 / 12
parameters.

Bar: objectives within 1e-5 relative; residuals within 1e-5 of their largest magnitude
plus the FP32 interpolation error (2e-6 of the uint16 range, scaled as the objective
scales the sample values); DVC_ERR_RANGE exactly where the oracle throws."""
import itertools

import numpy as np
import pytest

import digitalvolumecorrelation_amd as D
from digitalvolumecorrelation_amd import synth

pytestmark = pytest.mark.gpu

T = np.array([1.4, -0.6, 2.2])


@pytest.fixture(scope="module")
def pair(ctx):
    ref = synth.smoothed_noise((72, 72, 72), 1.5, np.uint16, seed=21)
    tar = synth.translate(ref, T)
    ctx.upload(0, ref)
    ctx.upload(1, tar)
    return ref, tar


def _x(ndof, off):
    x = np.zeros(ndof)
    x[:3] = T + off
    if ndof > 3:
        x[3:6] = [0.01, -0.02, 0.015]
    if ndof > 6:
        x[6:12] = [0.004, -0.003, 0.002, 0.001, -0.0015, 0.0005]
    return x


CASES = [c for c in itertools.product(["cube", "sphere"], ["sad", "ssd", "zssd", "nssd", "znssd"],
                                      ["tricubic"], [3, 6, 12])] + \
        [c for c in itertools.product(["cube", "sphere"], ["ssd", "znssd"], ["trilinear", "nearest"], [6])]


@pytest.mark.parametrize("geom,obj,interp,ndof", CASES)
def test_matches_oracle(ctx, oracle, pair, geom, obj, interp, ndof):
    ref, tar = pair
    kw = dict(ndof=ndof, obj=obj, interp=interp, geom=geom, subvol_size=19, subvol_npts=1500, disp_max=4)
    rp = D.make_params(**kw)
    run = oracle.make_run(shape=ref.shape, bytes_per=2, **kw)
    pt = np.array([35.3, 36.8, 34.1])
    start = np.rint(T)
    x = _x(ndof, np.array([0.25, -0.15, 0.2]))
    g, gres = ctx.obj_val_at(rp, pt, x, start=start, residual=True)
    o, ores = oracle.obj_val_at(run, ref, tar, pt, start, x, residual=True)
    assert ctx.obj_val_at(rp, pt, x, start=start) == g
    scale = {"sad": 1.0, "ssd": 1.0, "zssd": 1.0}.get(obj)
    if scale is None:                          # normalised samples: the values' norm divides them
        scale = 1.0 / np.sqrt((ref.astype(np.float64) ** 2).mean() * len(ores))
    rerr = np.abs(gres - ores).max()
    print(f"  {geom} {obj} {interp} {ndof}: obj {g:.6e} vs {o:.6e} (rel {abs(g - o) / abs(o):.2e}), "
          f"max residual error {rerr:.2e} of {np.abs(ores).max():.2e}")
    assert abs(g - o) <= 1e-5 * abs(o)
    assert rerr <= 1e-5 * np.abs(ores).max() + 2e-6 * 65535 * scale


def test_start_defaults_to_rigid_trans_and_range_rule(ctx, oracle, pair):
    ref, tar = pair
    kw = dict(ndof=3, obj="ssd", interp="tricubic", geom="sphere", subvol_size=19, subvol_npts=1500, disp_max=4,
              rigid_trans=tuple(np.rint(T)))
    rp = D.make_params(**kw)
    run = oracle.make_run(shape=ref.shape, bytes_per=2, **kw)
    pt = np.array([35.3, 36.8, 34.1])
    x = T + 0.1
    assert ctx.obj_val_at(rp, pt, x) == ctx.obj_val_at(rp, pt, x, start=np.rint(T))
    # a warp beyond the box around the start translation: Range_Fail on both sides
    for far in (T + [5.5, 0.0, 0.0], T - [0.0, 0.0, 5.5], np.array([np.nan, 0.0, 0.0])):
        assert oracle.obj_val_at(run, ref, tar, pt, np.rint(T), far) is None
        with pytest.raises(D.DvcError) as e:
            ctx.obj_val_at(rp, pt, far)
        assert e.value.code == D.ERR_RANGE
    # a point whose reference cloud leaves the volume
    with pytest.raises(D.DvcError) as e:
        ctx.obj_val_at(rp, np.array([4.0, 36.0, 36.0]), x)
    assert e.value.code == D.ERR_RANGE
    assert oracle.obj_val_at(run, ref, tar, np.array([4.0, 36.0, 36.0]), np.rint(T), x) is None
