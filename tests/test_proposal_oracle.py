"""Pins of the proposal network's CPU oracle (oracle/proposal_net.py).

The network has no reference counterpart, so the oracle is pinned by what the
mathematics fixes: its im2col convolution against a brute-force loop over the
definition (stride 2, padding 1, tap-major weights), col2im as the exact adjoint of
im2col, every analytic gradient against central finite differences of the loss in
FP64, the crop rule against direct slicing, and the BF16 rounding against known
bit patterns.
"""
import numpy as np
import pytest

import proposal_net as PN


def test_conv_matches_brute_force():
    rng = np.random.default_rng(0)
    B, D, C, Co = 2, 4, 3, 5
    x = rng.normal(size=(B, D, D, D, C))
    W = rng.normal(size=(Co, 27 * C))
    col = PN.im2col(x)
    y = col @ W.T
    Do = D // 2
    ref = np.zeros((B, Do, Do, Do, Co))
    for b in range(B):
        for oz in range(Do):
            for oy in range(Do):
                for ox in range(Do):
                    for co in range(Co):
                        s = 0.0
                        for kz in range(3):
                            for ky in range(3):
                                for kx in range(3):
                                    z, yy, xx = 2 * oz + kz - 1, 2 * oy + ky - 1, 2 * ox + kx - 1
                                    if 0 <= z < D and 0 <= yy < D and 0 <= xx < D:
                                        tap = (kz * 3 + ky) * 3 + kx
                                        s += (W[co, tap * C:(tap + 1) * C] * x[b, z, yy, xx]).sum()
                        ref[b, oz, oy, ox, co] = s
    assert np.allclose(y, ref, rtol=0, atol=1e-12)


def test_col2im_is_adjoint_of_im2col():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 8, 8, 8, 4))
    c = rng.normal(size=(2, 4, 4, 4, 27 * 4))
    lhs = (PN.im2col(x) * c).sum()
    rhs = (x * PN.col2im(c, 8, 4)).sum()
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(2)
    p = PN.unpack(PN.init_params(3).astype(np.float64))
    for l in range(4):
        p[f"b{l}"] = rng.normal(0, 0.05, p[f"b{l}"].shape)      # exercise the bias paths
    p["fcb"] = rng.normal(0, 0.1, 12)
    x = rng.normal(size=(2, 32, 32, 32, 2))
    label = rng.normal(size=(2, 12)) * PN.SCALE
    loss, g = PN.loss_and_grads(p, x, label)
    h = 1e-6
    for name in ["w0", "b0", "w1", "b1", "w2", "b2", "w3", "b3", "fcw", "fcb"]:
        flat = p[name].reshape(-1)
        for idx in rng.choice(flat.size, 3, replace=False):
            old = flat[idx]
            flat[idx] = old + h
            lp = PN.loss_and_grads(p, x, label)[0]
            flat[idx] = old - h
            lm = PN.loss_and_grads(p, x, label)[0]
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            an = g[name].reshape(-1)[idx]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)) + 1e-8, (name, idx, fd, an)


def test_loss_is_scaled_squared_error():
    p = PN.unpack(PN.init_params(4).astype(np.float64))
    x = np.random.default_rng(5).normal(size=(4, 32, 32, 32, 2))
    y, _ = PN.forward(p, x)
    lab = (y + 0.5) * PN.SCALE            # every output 0.5 below its label / scale
    loss, _ = PN.loss_and_grads(p, x, lab)
    assert abs(loss - 12 * 0.25) <= 1e-12


def test_crops_slice_clamp_and_normalise():
    rng = np.random.default_rng(6)
    ref = rng.integers(0, 255, (40, 50, 60)).astype(np.uint8)
    tar = rng.integers(0, 255, (40, 50, 60)).astype(np.uint8)
    inner = np.array([[30.7, 25.2, 20.9]])
    c = PN.crops(ref, tar, inner)[0]
    box = ref[20 - 15:20 + 17, 25 - 15:25 + 17, 30 - 15:30 + 17].astype(np.float64)
    assert np.allclose(c[..., 0], (box - box.mean()) / box.std())
    assert abs(c[..., 1].mean()) < 1e-12 and abs(c[..., 1].std() - 1) < 1e-12
    corner = PN.crops(ref, tar, np.array([[0.5, 0.5, 0.5]]))[0]
    # x, y, z below 0 repeat the face voxel: the first 16 entries on each axis are equal
    assert np.all(corner[:16, :16, :16, 0] == corner[0, 0, 0, 0])


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-3], dtype=np.float64)
    r = PN.bf16(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0                     # tie, rounds to even
    assert r[2] == 1.0 + 2 ** -6           # tie, rounds up to even
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-3) <= 3.0e-3 * 2 ** -8


def test_mirrored_labels_mirror_the_warp(oracle):
    # flip_labels (augmentation of mirrored crop pairs) against the oracle's warp
    # (FloatingCloud::affine_to): warping the cube cloud with the mirrored parameters
    # gives the mirror image, about the point, of the cloud warped with the originals
    from digitalvolumecorrelation_amd.proposal import flip_labels
    run = oracle.make_run(ndof=12, geom="cube", subvol_size=11, subvol_npts=27)
    pt = np.array([40.0, 41.0, 42.0])
    stable, _, _ = oracle.cloud(run, pt)
    rng = np.random.default_rng(15)
    for fl in range(8):
        par = np.concatenate([rng.uniform(-3, 3, 3), rng.uniform(-0.05, 0.05, 3), rng.uniform(-0.03, 0.03, 6)])
        d = np.array([-1.0 if fl & b else 1.0 for b in (1, 2, 4)])
        moved, _ = oracle.affine(run, pt, par, 12)
        moved_f, _ = oracle.affine(run, pt, flip_labels(par[None], np.array([fl]))[0], 12)
        # stable point j and its mirror image about pt share an index map on the grid
        mirror = pt + (stable - pt) * d
        idx = [int(np.argmin(np.abs(stable - q).sum(axis=1))) for q in mirror]
        assert np.allclose(stable[idx], mirror, atol=1e-9)
        assert np.allclose(moved_f[idx], pt + (moved - pt) * d, atol=1e-9), fl
