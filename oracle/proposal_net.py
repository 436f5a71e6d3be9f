"""FP64 numpy restatement of the learned start-warp proposal network.

TEST INFRASTRUCTURE ONLY: the expected values for the network kernels in
digitalvolumecorrelation_amd/csrc/dvc_net.hip (crops, forward pass, loss and every
parameter gradient).  The reference has no counterpart of this network; what fixes
the expected values is the mathematics of the layers themselves, pinned in
tests/test_proposal_oracle.py by a brute-force convolution loop and by central finite
differences of the loss.

Layout, as the kernels use it: activations channels-last [B, D, D, D, C]; convolution
weights [Cout, 27*Cin] with k = tap*Cin + c and tap = (kz*3 + ky)*3 + kx; stride 2,
padding 1; the head averages the last 2^3 voxels and applies a 128 -> 12 linear map;
the loss is the batch mean of sum_j (y_j - label_j / scale_j)^2.

With ``bf16=True`` values are rounded to bfloat16 (round to nearest even) where the
kernels store them in BF16: the input crops, the convolution weights, each ReLU output
and the masked output gradients that feed the weight and input gradient products.
"""
from __future__ import annotations

import numpy as np

CIN = (2, 16, 32, 64)
COUT = (16, 32, 64, 128)
DIN = (32, 16, 8, 4)
FEAT, NOUT = 128, 12
SCALE = np.array([8.0] * 3 + [0.05] * 3 + [0.02] * 6)


def layout():
    """Offsets of every tensor in the flat FP32 parameter vector."""
    off = {}
    o = 0
    for l in range(4):
        k = 27 * CIN[l]
        off[f"w{l}"] = (o, (COUT[l], k)); o += COUT[l] * k
        off[f"b{l}"] = (o, (COUT[l],)); o += COUT[l]
    off["fcw"] = (o, (NOUT, FEAT)); o += NOUT * FEAT
    off["fcb"] = (o, (NOUT,)); o += NOUT
    return off, o


def unpack(flat):
    off, n = layout()
    assert flat.size == n
    return {k: np.asarray(flat[o:o + int(np.prod(s))], dtype=np.float64).reshape(s) for k, (o, s) in off.items()}


def pack(p):
    off, n = layout()
    flat = np.zeros(n, np.float64)
    for k, (o, s) in off.items():
        flat[o:o + int(np.prod(s))] = np.asarray(p[k]).reshape(-1)
    return flat


def bf16(x):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def crops(ref, tar, points):
    """[n, 32, 32, 32, 2]: the 32^3 boxes starting 15 voxels below floor(point) on each
    axis (x, y, z), clamped to the volume, each channel normalised to zero mean and unit
    variance (population variance)."""
    T, H, W = ref.shape
    out = np.zeros((len(points), 32, 32, 32, 2))
    r = np.arange(32)
    for i, p in enumerate(np.asarray(points, dtype=np.float64)):
        o = np.floor(p).astype(np.int64) - 15
        xs = np.clip(o[0] + r, 0, W - 1)
        ys = np.clip(o[1] + r, 0, H - 1)
        zs = np.clip(o[2] + r, 0, T - 1)
        for c, v in enumerate((ref, tar)):
            box = v[np.ix_(zs, ys, xs)].astype(np.float64)
            m = box.mean()
            var = (box * box).mean() - m * m
            out[i, ..., c] = (box - m) / np.sqrt(var) if var > 1e-12 else 0.0
    return out


def im2col(x):
    """x [B, D, D, D, C] -> col [B, D/2, D/2, D/2, 27*C] (tap-major)."""
    B, D, _, _, C = x.shape
    Do = D // 2
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    parts = []
    for kz in range(3):
        for ky in range(3):
            for kx in range(3):
                parts.append(xp[:, kz:kz + 2 * Do:2, ky:ky + 2 * Do:2, kx:kx + 2 * Do:2, :])
    return np.concatenate(parts, axis=-1)


def col2im(dcol, D, C):
    """Adjoint of im2col: dcol [B, Do, Do, Do, 27*C] -> dx [B, D, D, D, C]."""
    B, Do = dcol.shape[0], dcol.shape[1]
    dxp = np.zeros((B, D + 2, D + 2, D + 2, C))
    t = 0
    for kz in range(3):
        for ky in range(3):
            for kx in range(3):
                dxp[:, kz:kz + 2 * Do:2, ky:ky + 2 * Do:2, kx:kx + 2 * Do:2, :] += dcol[..., t * C:(t + 1) * C]
                t += 1
    return dxp[:, 1:-1, 1:-1, 1:-1, :]


def forward(p, x, bf=False):
    """Returns (y [B, 12] unscaled head output, cache)."""
    rnd = bf16 if bf else (lambda v: v)
    a = rnd(x)
    acts, cols = [a], []
    for l in range(4):
        col = im2col(a)
        z = col @ rnd(p[f"w{l}"]).T + p[f"b{l}"]
        a = rnd(np.maximum(z, 0.0))
        cols.append(col)
        acts.append(a)
    feat = acts[-1].reshape(len(x), 8, FEAT).mean(axis=1)
    y = feat @ p["fcw"].T + p["fcb"]
    return y, {"acts": acts, "cols": cols, "feat": feat}


def loss_and_grads(p, x, label, bf=False):
    """Loss and the gradient of every parameter (dict like p)."""
    rnd = bf16 if bf else (lambda v: v)
    y, c = forward(p, x, bf)
    B = len(x)
    e = y - label / SCALE
    loss = float((e * e).sum() / B)
    dy = 2.0 * e / B
    g = {"fcw": dy.T @ c["feat"], "fcb": dy.sum(axis=0)}
    dfeat = dy @ p["fcw"]
    da = np.repeat(dfeat[:, None, :] / 8.0, 8, axis=1).reshape(c["acts"][-1].shape)
    for l in range(3, -1, -1):
        act = c["acts"][l + 1]
        dz = np.where(act > 0.0, da, 0.0)
        g[f"b{l}"] = dz.reshape(-1, COUT[l]).sum(axis=0)
        dzr = rnd(dz)
        col = c["cols"][l]
        g[f"w{l}"] = dzr.reshape(-1, COUT[l]).T @ col.reshape(-1, col.shape[-1])
        if l > 0:
            dcol = dzr @ rnd(p[f"w{l}"])
            da = col2im(dcol, DIN[l], CIN[l])
    return loss, g


def init_params(seed=0):
    """He-normal convolution weights, small head weights, zero biases (FP32)."""
    rng = np.random.default_rng(seed)
    p = {}
    for l in range(4):
        k = 27 * CIN[l]
        p[f"w{l}"] = rng.normal(0.0, np.sqrt(2.0 / k), (COUT[l], k))
        p[f"b{l}"] = np.zeros(COUT[l])
    p["fcw"] = rng.normal(0.0, 0.1 / np.sqrt(FEAT), (NOUT, FEAT))
    p["fcb"] = np.zeros(NOUT)
    return pack(p).astype(np.float32)
