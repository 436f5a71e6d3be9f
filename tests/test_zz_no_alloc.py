"""dvc_obj_val_at_noalloc (no counterpart in the reference: dvc_obj_val_at for callers
that cannot allocate) against dvc_obj_val_at, through the C ABI.

The two entry points run the same kernels on the same inputs, so objectives and
residuals must be equal bit for bit; two calls in a row, the second on a smaller cloud
(the scratch keeps the first call's contents), then a range failure and the first
geometry again."""
import ctypes as C

import numpy as np
import pytest

import digitalvolumecorrelation_amd as D
from digitalvolumecorrelation_amd import synth

pytestmark = pytest.mark.gpu

T = np.array([1.4, -0.6, 2.2])


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _call(ctx, name, params, pt, start, x):
    fn = getattr(ctx.lib, name)
    n = C.c_int32(0)
    o = C.c_double()
    rc = fn(ctx.h, C.byref(params), _dp(pt), _dp(start), _dp(x), len(x), C.byref(o), None, 0, C.byref(n))
    if rc != 0:
        return rc, None, None
    res = np.zeros(n.value)
    rc = fn(ctx.h, C.byref(params), _dp(pt), _dp(start), _dp(x), len(x), C.byref(o), _dp(res), len(res), C.byref(n))
    return rc, o.value, res


@pytest.fixture(scope="module")
def noalloc_ctx():
    c = D.Context(0)
    ref = synth.smoothed_noise((72, 72, 72), 1.5, np.uint16, seed=21)
    c.upload(0, ref)
    c.upload(1, synth.translate(ref, T))
    yield c
    c.close()


def test_noalloc_matches_allocating_entry(noalloc_ctx):
    ctx = noalloc_ctx
    pt = np.array([35.3, 36.8, 34.1])
    start = np.rint(T)
    x12 = np.r_[T + [0.2, -0.1, 0.15], 0.01, -0.02, 0.015, 0.004, -0.003, 0.002, 0.001, -0.0015, 0.0005]
    calls = [
        (dict(ndof=12, obj="znssd", interp="tricubic", geom="sphere", subvol_size=25, subvol_npts=6000, disp_max=4), x12),
        (dict(ndof=6, obj="ssd", interp="tricubic", geom="cube", subvol_size=15, subvol_npts=1000, disp_max=4), x12[:6]),
    ]
    for kw, x in calls + calls[:1]:
        p = D.make_params(**kw)
        rc_a, oa, ra = _call(ctx, "dvc_obj_val_at", p, pt, start, np.ascontiguousarray(x))
        rc_b, ob, rb = _call(ctx, "dvc_obj_val_at_noalloc", p, pt, start, np.ascontiguousarray(x))
        print(f"  {kw['geom']} {kw['subvol_npts']}: {len(ra)} samples, obj {oa:.9e} / {ob:.9e}")
        assert rc_a == 0 and rc_b == 0
        assert ob == oa and np.array_equal(rb, ra)
    # a range failure leaves the entry point usable
    p = D.make_params(**calls[1][0])
    far = np.ascontiguousarray(np.r_[T + [6.0, 0.0, 0.0], np.zeros(3)])
    assert _call(ctx, "dvc_obj_val_at_noalloc", p, pt, start, far)[0] == D.ERR_RANGE
    rc, o, r = _call(ctx, "dvc_obj_val_at_noalloc", p, pt, start, np.ascontiguousarray(calls[1][1]))
    rc2, o2, r2 = _call(ctx, "dvc_obj_val_at", p, pt, start, np.ascontiguousarray(calls[1][1]))
    assert rc == 0 and rc2 == 0 and o == o2 and np.array_equal(r, r2)
