"""Proposal network kernels (dvc_net_*) on the GPU against the FP64 oracle
(oracle/proposal_net.py, with BF16 rounding at the points where the kernels store
BF16).

Bars, written per assertion: crops within one BF16 unit in the last place of the
oracle's value; head outputs and the loss within 1e-2 relative; every parameter
gradient within 3e-2 relative (L2 norm of the difference over the norm of the
oracle's gradient: FP32 accumulation in another order flips some BF16 roundings of
the activations); training drives the loss of a fixed batch down; proposals written
by dvc_search_propose equal the network's forward pass on the same points.
"""
import numpy as np
import pytest

import digitalvolumecorrelation_amd as D
from digitalvolumecorrelation_amd import proposal as P
from digitalvolumecorrelation_amd import synth
import proposal_net as PN

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pair():
    return synth.translation_u8(size=64, n=2, spacing=8)


def _bits(d):
    u = d.download().astype(np.uint32) << 16
    return u.view(np.float32).astype(np.float64)


def _crops(ctx, net, pts, slots=None, flips=None):
    dp = D.DeviceArray.from_host(ctx, np.ascontiguousarray(pts, dtype=np.float64))
    out = D.DeviceArray(ctx, (len(pts), P.CROP_ELEMS), np.uint16)
    sl = D.DeviceArray.from_host(ctx, np.asarray(slots, np.int32)) if slots is not None else None
    fl = D.DeviceArray.from_host(ctx, np.asarray(flips, np.int32)) if flips is not None else None
    net.crop(dp.ptr, len(pts), out.ptr, slot_dev=sl.ptr if sl else None, flip_dev=fl.ptr if fl else None)
    ctx.synchronize()
    return out


def test_crops_match_oracle(ctx, pair):
    ctx.upload(0, pair.ref)
    ctx.upload(1, pair.tar)
    net = P.ProposalNet(ctx, max_batch=8)
    pts = np.array([[31.3, 30.8, 33.9], [2.5, 60.2, 10.0], [63.9, 0.1, 40.4]])   # two clamped at faces
    g = _bits(_crops(ctx, net, pts)).reshape(3, 32, 32, 32, 2)
    o = PN.bf16(PN.crops(pair.ref, pair.tar, pts))
    assert np.all(np.abs(g - o) <= 2.0 ** -7 * np.abs(o) + 1e-6)
    # slot table: crop i lands in slot [2, 0, 1][i]
    h = _bits(_crops(ctx, net, pts, slots=[2, 0, 1])).reshape(3, 32, 32, 32, 2)
    assert np.array_equal(h[[2, 0, 1]], g)
    # mirrors: bit 0 x (last array axis), bit 1 y, bit 2 z
    f = _bits(_crops(ctx, net, pts, flips=[1, 6, 7])).reshape(3, 32, 32, 32, 2)
    assert np.array_equal(f[0], g[0][:, :, ::-1])
    assert np.array_equal(f[1], g[1][::-1, ::-1, :])
    assert np.array_equal(f[2], g[2][::-1, ::-1, ::-1])


def test_forward_and_gradients_match_oracle(ctx, pair):
    ctx.upload(0, pair.ref)
    ctx.upload(1, pair.tar)
    params = P.init_params(7)
    rng = np.random.default_rng(8)
    params[:] += rng.normal(0, 0.01, params.size).astype(np.float32) * (np.abs(params) == 0)   # nonzero biases
    net = P.ProposalNet(ctx, max_batch=8, training=True, params=params)
    pts = rng.uniform(20, 44, (4, 3))
    crops = _crops(ctx, net, pts)
    labels = (rng.normal(size=(4, 12)) * P.SCALE).astype(np.float32)
    lab = D.DeviceArray.from_host(ctx, labels)
    y = D.DeviceArray(ctx, (4, 12), np.float32)
    net.forward(crops.ptr, 4, None, y.ptr)
    loss, grad = net.gradients(crops.ptr, lab.ptr, 4)
    x = _bits(crops).reshape(4, 32, 32, 32, 2)
    p = PN.unpack(params.astype(np.float64))
    yo, _ = PN.forward(p, x, bf=True)
    lo, go = PN.loss_and_grads(p, x, labels.astype(np.float64), bf=True)
    yg = y.download().astype(np.float64)
    assert np.linalg.norm(yg - yo) <= 1e-2 * np.linalg.norm(yo)
    assert abs(loss - lo) <= 1e-2 * lo
    gg = PN.unpack(grad.astype(np.float64))
    for name in go:
        err = np.linalg.norm(gg[name] - go[name]) / np.linalg.norm(go[name])
        print(f"  {name}: relative gradient error {err:.2e}")
        assert err <= 3e-2, name


def test_training_fits_a_fixed_batch(ctx, pair):
    ctx.upload(0, pair.ref)
    ctx.upload(1, pair.tar)
    net = P.ProposalNet(ctx, max_batch=8, training=True, params=P.init_params(9))
    rng = np.random.default_rng(10)
    crops = _crops(ctx, net, rng.uniform(20, 44, (8, 3)))
    lab = D.DeviceArray.from_host(ctx, (rng.normal(size=(8, 12)) * P.SCALE).astype(np.float32))
    first = net.train_step(crops.ptr, lab.ptr, 8, 1e-3, want_loss=True)
    for _ in range(80):
        last = net.train_step(crops.ptr, lab.ptr, 8, 1e-3, want_loss=True)
    print(f"  loss {first:.4f} -> {last:.4f}")
    assert last < 0.1 * first


def test_propose_equals_forward(ctx):
    case = synth.translation_u8(size=80, n=3, spacing=8)
    ctx.upload(0, case.ref)
    ctx.upload(1, case.tar)
    net = P.ProposalNet(ctx, max_batch=8, params=P.init_params(11))
    order, neigh = D.sort_order_neighbors(case.points)
    starts = net.predict_points(case.points)
    ctx.set_start_mode(D.START_GIVEN)
    try:
        ctx.prepare(D.make_params(**case.params), case.points, order, neigh)
        ctx.set_starts(starts)
        ctx.execute()
        a = ctx.fetch()
        net.propose()            # batches of 8 in processing order, into the start table
        ctx.execute()
        b = ctx.fetch()
    finally:
        ctx.set_start_mode(D.START_CHAIN)
    assert np.array_equal(a["status"], b["status"])
    assert np.array_equal(a["par"], b["par"]) and np.array_equal(a["obj"], b["obj"])


def test_inference_forward_equals_training_forward_and_oracle(ctx, pair):
    # inference nets gather patches inside the GEMM (conv_fwd_kernel); training nets
    # store the im2col matrix first: same values in the same k order, so the outputs
    # are identical; both match the oracle
    ctx.upload(0, pair.ref)
    ctx.upload(1, pair.tar)
    params = P.init_params(12)
    rng = np.random.default_rng(13)
    params[:] += rng.normal(0, 0.02, params.size).astype(np.float32) * (params == 0)
    inf = P.ProposalNet(ctx, max_batch=8, params=params)
    trn = P.ProposalNet(ctx, max_batch=8, training=True, params=params)
    crops = _crops(ctx, inf, rng.uniform(16, 48, (6, 3)))
    ya = D.DeviceArray(ctx, (6, 12), np.float32)
    yb = D.DeviceArray(ctx, (6, 12), np.float32)
    inf.forward(crops.ptr, 6, None, ya.ptr)
    trn.forward(crops.ptr, 6, None, yb.ptr)
    a, b = ya.download(), yb.download()
    assert np.array_equal(a, b)
    yo, _ = PN.forward(PN.unpack(params.astype(np.float64)), _bits(crops).reshape(6, 32, 32, 32, 2), bf=True)
    assert np.linalg.norm(a - yo) <= 1e-2 * np.linalg.norm(yo)


def test_tile_warp_matches_catmull_rom(ctx):
    rng = np.random.default_rng(14)
    ref = synth.smoothed_noise((40, 48, 56), 1.5, np.uint8, seed=3)
    tile = 24
    nt = [(s + tile - 1) // tile for s in (56, 48, 40)]           # x, y, z
    maps = []
    for _ in range(nt[0] * nt[1] * nt[2]):
        F = np.eye(3) + rng.uniform(-0.03, 0.03, (3, 3))
        maps.append(np.concatenate([F.ravel(), rng.uniform(-4, 4, 3)]))
    maps = np.array(maps)
    ctx.upload(0, ref)
    ctx.upload(1, ref)
    ctx.warp_tiles(tile, maps)
    got = ctx.download(1, ref.shape, np.uint8).astype(np.float64)
    zz, yy, xx = np.meshgrid(np.arange(40), np.arange(48), np.arange(56), indexing="ij")
    t = ((zz // tile) * nt[1] + yy // tile) * nt[0] + xx // tile
    X = np.stack([xx, yy, zz], -1).reshape(-1, 3).astype(np.float64)
    m = maps[t.reshape(-1)]
    src = np.einsum("nij,nj->ni", m[:, :9].reshape(-1, 3, 3), X) + m[:, 9:]
    want = np.clip(np.rint(synth.catmull_rom(ref, src)), 0, 255).reshape(ref.shape)
    # FP64 sums in another order: a value at a rounding tie may land one level away
    assert np.abs(got - want).max() <= 1.0
    assert (got == want).mean() > 0.999


def test_tiled_labels_recover_imposed_warps(ctx):
    tp = P.tiled_pair(5, size=128, tile=64)
    pts, lab = P.label_pair(ctx, tp, 64)
    assert len(pts) >= 0.5 * len(tp.points)
    idx = np.array([np.flatnonzero((tp.points == q).all(axis=1))[0] for q in pts])
    err = np.abs(lab[:, :3] - tp.disp[idx]).max(axis=1)
    print(f"  labelled {len(pts)} of {len(tp.points)}; max |label - imposed| {err.max():.3f} voxel")
    assert np.median(err) < 0.05
